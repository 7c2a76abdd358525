/*
 * vcl_oracle.h -- CPU restatement of ViennaCL 1.7.1's host_based (OpenMP)
 * backend for the sparse Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against, and the CPU baseline that bench.py reports beside the GPU
 * figure.  Nothing in the product (viennacl-dev_amd/) links, loads or calls it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * reference root).  Floating-point expressions keep the reference's evaluation
 * order and the library is built with -ffp-contract=off, so single-threaded
 * results are bit-identical to the reference built the same way
 * (checked against oracle/_ref by tests/test_oracle_golden.py).
 *
 * Pinning:
 *   - SpMV (CSR/SELL), row_info, dots, norms, pipelined kernels, all solvers:
 *     pinned bit-for-bit against the reference build (oracle/_ref) and the
 *     golden fixtures under tests/golden/, plus closed forms (Laplacian
 *     eigenpairs, row sums) and dense numpy solves in tests/test_oracle_math.py.
 *   - orc_mixed_cg2_csr mode 1 (reliable updates), orc_batched_cg2_csr (pipelined CG per
 *     block, FP64 or FP32), orc_adjoint_grad_csr: no reference counterpart; pinned by the FP64
 *     true residual, dense numpy solves and central finite differences (tests/test_oracle_math.py).
 *     orc_mixed_cg2_csr mode 0 with acc64 = 0 is orc_mixed_cg_csr (pinned against the reference).
 *
 * Index type is uint32_t as in compressed_matrix<T> (compressed_matrix.hpp:57-58).
 */
#ifndef VCL_ORACLE_H
#define VCL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP thread count for the reductions (1 = reference single-thread order). */
void orc_set_num_threads(int n);
int  orc_get_num_threads(void);
/* 0 (default): the reference's reduction order.  k > 0: every inner product is summed as k
 * contiguous blocks (sequential inside, block sums added in order), the order of the reference's
 * OpenMP reduction with k threads, without its 5000-element threshold. */
void orc_set_reduction_blocks(int k);

/* ---------------------------------------------------------------- SpMV --- */
/* host_based/sparse_matrix_operations.hpp:110-186 */
void orc_csr_spmv_f64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                      const double* x, size_t x_start, size_t x_stride, double alpha,
                      double* y, size_t y_start, size_t y_stride, double beta);
void orc_csr_spmv_f32(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                      const float* x, size_t x_start, size_t x_stride, float alpha,
                      float* y, size_t y_start, size_t y_stride, float beta);

/* host_based/sparse_matrix_operations.hpp:1796-1868, iterating over ceil(n/C) slices
 * (the reference's n/C+1 reads one slice past the arrays when C divides n). */
void orc_sell_spmv_f64(uint32_t n_rows, uint32_t C, const uint32_t* columns_per_block,
                       const uint32_t* column_indices, const uint32_t* block_start, const double* elements,
                       const double* x, size_t x_start, size_t x_stride, double alpha,
                       double* y, size_t y_start, size_t y_stride, double beta);
void orc_sell_spmv_f32(uint32_t n_rows, uint32_t C, const uint32_t* columns_per_block,
                       const uint32_t* column_indices, const uint32_t* block_start, const float* elements,
                       const float* x, size_t x_start, size_t x_stride, float alpha,
                       float* y, size_t y_start, size_t y_stride, float beta);

/* sliced_ell_matrix.hpp:140-214: CSR -> SELL-C (sigma = 1).  Sizes first
 * (num_slices = ceil(n/C), stored entries), then fill. */
void orc_sell_sizes(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, uint32_t* num_slices, uint64_t* stored);
void orc_sell_from_csr_f64(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                           uint32_t* columns_per_block, uint32_t* column_indices, uint32_t* block_start, double* elements);

/* compressed_matrix.hpp:1152-1188: greedy 1024-nnz row blocks; returns block count,
 * row_blocks must hold n_rows+1 entries. */
uint32_t orc_row_blocks(uint32_t n_rows, const uint32_t* row_ptr, uint32_t* row_blocks);

/* host_based/sparse_matrix_operations.hpp:52-98; info: 0 inf-norm, 1 one-norm, 2 two-norm, 3 diagonal */
void orc_csr_row_info_f64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                          double* out, int info);

/* --------------------------------------------------------- vector ops --- */
/* host_based/vector_operations.hpp:465-575 (dot accumulates in the value type) */
double orc_dot_f64(size_t n, const double* x, const double* y);
float  orc_dot_f32(size_t n, const float* x, const float* y);
/* host_based/vector_operations.hpp:727-833 (sqrt of the value-type sum of squares) */
double orc_norm2_f64(size_t n, const double* x);
float  orc_norm2_f32(size_t n, const float* x);

/* ----------------------------------------------- pipelined operations --- */
/* host_based/iterative_operations.hpp:58-103 (CSR) and :223-292 (SELL): Ap = A p, with
 * buf[chunk] = <Ap,Ap>, buf[2 chunk] = <p,Ap>, buf[offset] = <Ap,r0*> when r0star != NULL. */
void orc_pipelined_prod_csr_f64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                                const double* p, double* Ap, const double* r0star,
                                double* buf, size_t chunk, size_t offset);
void orc_pipelined_prod_sell_f64(uint32_t n_rows, uint32_t C, const uint32_t* columns_per_block,
                                 const uint32_t* column_indices, const uint32_t* block_start, const double* elements,
                                 const double* p, double* Ap, const double* r0star,
                                 double* buf, size_t chunk, size_t offset);
/* host_based/iterative_operations.hpp:378-418 */
void orc_pipelined_cg_vector_update_f64(size_t n, double* x, double alpha, double* p, double* r,
                                        const double* Ap, double beta, double* buf);
/* host_based/iterative_operations.hpp:518-563 */
void orc_pipelined_bicgstab_update_s_f64(size_t n, double* s, const double* r, const double* Ap,
                                         double* buf, size_t chunk, size_t offset);
/* host_based/iterative_operations.hpp:572-621 */
void orc_pipelined_bicgstab_vector_update_f64(size_t n, double* x, double alpha, double* p, double omega,
                                              const double* s, double* r, const double* As, double beta,
                                              const double* Ap, const double* r0star, double* buf);

/* ------------------------------------------------------------ solvers --- */
typedef struct orc_csr64 {
  uint32_t n_rows;
  const uint32_t* row_ptr;
  const uint32_t* col;
  const double* val;
} orc_csr64;

typedef struct orc_sell64 {
  uint32_t n_rows;
  uint32_t C;
  const uint32_t* columns_per_block;
  const uint32_t* column_indices;
  const uint32_t* block_start;
  const double* elements;
} orc_sell64;

/* coordinate_matrix (format 2: coords = 2*nnz (row, col) pairs in row order), ell_matrix
 * (format 3) and hyb_matrix (format 4: ELL part plus csr_rows / csr_cols / csr_elements); the ELL
 * parts are column-major with leading dimension ld and width columns. */
#define ORC_FORMAT_COO 2
#define ORC_FORMAT_ELL 3
#define ORC_FORMAT_HYB 4
typedef struct orc_fmt64 {
  int format;
  uint32_t n_rows;
  uint32_t nnz;
  uint32_t ld;
  uint32_t width;
  const uint32_t* coords;
  const double* elements;
  const uint32_t* csr_rows;
  const uint32_t* csr_cols;
  const double* csr_elements;
} orc_fmt64;

typedef struct orc_result {
  uint32_t iters;
  double error;
} orc_result;

/* cg.hpp:128-187 pipelined CG (no preconditioner).  history[i] = rr of iteration i (nullable). */
orc_result orc_cg_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                uint32_t max_iters, double* history);
orc_result orc_cg_pipelined_sell(const orc_sell64* A, const double* b, double* x, double tol, double abs_tol,
                                 uint32_t max_iters, double* history);
/* cg.hpp:257-322 classic preconditioned CG; diag != NULL -> Jacobi (jacobi_precond.hpp:103-130). */
orc_result orc_cg_classic_csr(const orc_csr64* A, const double* b, double* x, const double* diag,
                              double tol, double abs_tol, uint32_t max_iters, double* history);
/* bicgstab.hpp:97-215 pipelined BiCGStab (no preconditioner). */
orc_result orc_bicgstab_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                      uint32_t max_iters, double* history);
orc_result orc_bicgstab_pipelined_sell(const orc_sell64* A, const double* b, double* x, double tol, double abs_tol,
                                       uint32_t max_iters, double* history);
/* bicgstab.hpp:398-489 classic left-preconditioned BiCGStab with restarts (diag != NULL -> Jacobi);
 * diag == NULL runs the unpreconditioned classic variant bicgstab.hpp:295-383. */
orc_result orc_bicgstab_classic_csr(const orc_csr64* A, const double* b, double* x, const double* diag,
                                    double tol, double abs_tol, uint32_t max_iters, uint32_t max_iters_before_restart,
                                    double* history);
/* mixed_precision_cg.hpp:95-187: FP32 classic CG with FP64 defect-correction restarts. */
orc_result orc_mixed_cg_csr(const orc_csr64* A, const double* b, double* x, double tol, uint32_t max_iters,
                            float inner_tol, uint32_t* restarts);
/* Mixed-precision CG of the HIP path.
 *   mode 0: the schedule of mixed_precision_cg.hpp:95-187 (FP32 classic CG, restart of the Krylov
 *           space with an FP64 defect correction when <r,r> has fallen by inner_tol); acc64 != 0
 *           accumulates the FP32 inner products in double (the HIP kernels do), acc64 == 0 is
 *           orc_mixed_cg_csr.  diag must be NULL.
 *   mode 1: reliable updates.  FP32 pipelined CG (cg.hpp:128-187 recurrences, Jacobi by
 *           diag != NULL as <r, D^-1 r>) with FP64-accumulated inner products; when the FP32
 *           recursive ||r||^2 has fallen by inner_tol since the last update, or says the tolerance is
 *           met, x += x32 and r = b - A x are formed in FP64, ||b - A x|| is tested, r32 is replaced
 *           by the true residual and p is corrected by D^-1 (r_true - r32); the search direction is
 *           kept.  No reference counterpart.
 * history (optional, max_iters entries): FP32 recursive <r,r> per iteration.
 * updates receives the number of FP64 corrections. */
orc_result orc_mixed_cg2_csr(const orc_csr64* A, const double* b, double* x, const double* diag, double tol,
                             uint32_t max_iters, float inner_tol, int mode, int acc64, uint32_t* updates,
                             double* history);

/* prod_impl for the format (host_based/sparse_matrix_operations.hpp:1221-1246 COO, :1502-1537
 * ELL, :1872-1929 HYB), contiguous vectors. */
void orc_fmt_spmv_f64(const orc_fmt64* A, const double* x, double alpha, double* y, double beta);
/* pipelined_cg_prod / pipelined_bicgstab_prod for the format (host_based/iterative_operations.hpp:
 * 113-156, 166-210, 302-365). */
void orc_fmt_pipelined_prod_f64(const orc_fmt64* A, const double* p, double* Ap, const double* r0star,
                                double* buf, size_t chunk, size_t offset);
/* cg.hpp:128-187 / bicgstab.hpp:97-215 with the format's SpMV and fused product
 * (cg.hpp:204-254, bicgstab.hpp:231-285 send these formats to the pipelined solvers). */
orc_result orc_cg_pipelined_fmt(const orc_fmt64* A, const double* b, double* x, double tol, double abs_tol,
                                uint32_t max_iters, double* history);
orc_result orc_bicgstab_pipelined_fmt(const orc_fmt64* A, const double* b, double* x, double tol, double abs_tol,
                                      uint32_t max_iters, double* history);

/* gmres.hpp:182-370: pipelined GMRES(krylov_dim) without preconditioner, classical Gram-Schmidt.
 * gs_skip_tail = 1 reproduces the reference's single-thread Gram-Schmidt range, which leaves out
 * the last vector entry (host_based/iterative_operations.hpp:820-822); 0 uses all entries.
 * history (optional): the error estimate |rho * rho_0| / ||b|| after every counted step. */
orc_result orc_gmres_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                   uint32_t max_iters, uint32_t krylov_dim, int gs_skip_tail, double* history);

/* The pipelined GMRES operations of host_based/iterative_operations.hpp, single-threaded.  The
 * basis holds vector i at basis + i * ld. */
/* :734-779: norm = sqrt(sum buf[chunk..2chunk)); R[offset_in_R] = norm; v_k /= norm;
 * r_dot_vk[chunk_offset] = <residual, v_k> */
void orc_gmres_normalize_vk(size_t n, double* v_k, const double* residual, double* R, size_t offset_in_R,
                            const double* buf, double* r_dot_vk, size_t chunk, size_t chunk_offset);
/* :787-842: vi_in_vk[j*chunk] = <v_j, v_k>, j < k, over [0, n - 1) if skip_tail else [0, n) */
void orc_gmres_stage1(const double* basis, size_t n, size_t ld, size_t k, double* vi_in_vk, size_t chunk,
                      int skip_tail);
/* :851-893: h_j = sum of chunk j; v_k -= sum_j h_j v_j; R[j + k*krylov_dim] = h_j; buf[chunk] = ||v_k||^2 */
void orc_gmres_stage2(double* basis, size_t n, size_t ld, size_t k, const double* vi_in_vk, double* R,
                      size_t krylov_dim, double* buf, size_t chunk);
/* :898-926: x += coef[0] residual + sum_{j=1}^{k-1} coef[j] v_{j-1} */
void orc_gmres_update_result(double* x, const double* residual, const double* basis, size_t n, size_t ld,
                             const double* coef, size_t k);

/* Batched CG: independent pipelined CG (cg.hpp:128-187) per system i on rows
 * [sys_ptr[i], sys_ptr[i+1]) of a block-diagonal CSR with global column indices.
 * iters/errors hold one entry per system.  No reference counterpart. */
void orc_batched_cg_csr(uint32_t num_systems, const uint32_t* sys_ptr, const orc_csr64* A, const double* b,
                        double* x, double tol, uint32_t max_iters, uint32_t* iters, double* errors);
/* single != 0: the same on FP32 copies of the values and b, inner products accumulated in double. */
void orc_batched_cg2_csr(uint32_t num_systems, const uint32_t* sys_ptr, const orc_csr64* A, const double* b,
                         double* x, double tol, uint32_t max_iters, int single, uint32_t* iters, double* errors);

/* Adjoint of x = A^{-1} b on the CSR pattern: dA[k] = -lambda[row(k)] * x[col(k)]. */
void orc_adjoint_grad_csr(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col,
                          const double* lambda, const double* x, double* dA);

#ifdef __cplusplus
}
#endif

#endif
