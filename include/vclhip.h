/*
 * vclhip.h -- C ABI of the MI355X (gfx950) backend for ViennaCL's sparse Krylov path.
 *
 * This is the drop-in boundary: the HIP memory domain and the hot-path operations that
 * ViennaCL's backend-agnostic dispatchers call for each memory domain.  Each entry point
 * names the reference interface it stands in for (paths relative to the reference root,
 * viennacl/viennacl-dev 1.7.1).  The C++ headers in viennacl-dev_amd/viennacl/ call only
 * this ABI; so does the Python binding.
 *
 * Conventions
 *   - Every function returns int: 0 (VCLHIP_SUCCESS), a negative VCLHIP_ERROR_* code, or
 *     a positive hipError_t value.  Nothing throws across the ABI.  The text of the last
 *     failure on the calling thread is returned by vclhip_last_error().
 *     (reference: cuda_exception backend/cuda.hpp:38,59-68; VIENNACL_CUDA_LAST_ERROR_CHECK
 *     cuda/common.hpp:30,165-175; ViennaCLStatus libviennacl/include/viennacl.hpp:98-102)
 *   - Pointers named *_dev are device pointers (allocated by vclhip_malloc or any HIP
 *     allocator on the current device); *_host are host pointers.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.  Kernels are
 *     enqueued asynchronously unless a function says it synchronizes.
 *   - Indices are uint32_t like compressed_matrix<T> (compressed_matrix.hpp:57-58); vectors
 *     carry (start, stride) like vector_base (detail/vector_def.hpp:104-349).
 *   - x and y of an SpMV must not alias (the C++ layer copies through a temporary, as the
 *     reference's op_executor does, compressed_matrix.hpp:1231-1246).
 */
#ifndef VCLHIP_H
#define VCLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCLHIP_SUCCESS 0
#define VCLHIP_ERROR_INVALID_VALUE (-1)   /* bad size, null pointer, unsupported parameter */
#define VCLHIP_ERROR_ZERO_DIAGONAL (-2)   /* zero_on_diagonal_exception, forwards.h:621-625 */
#define VCLHIP_ERROR_NOT_SUPPORTED (-3)
#define VCLHIP_ERROR_NO_DEVICE (-4)
#define VCLHIP_ERROR_TOO_LARGE (-5)       /* size beyond a documented maximum */

/* ------------------------------------------------------------------ runtime ---- */
const char* vclhip_version(void);
const char* vclhip_last_error(void);
const char* vclhip_status_string(int status);
int vclhip_device_count(int* count);
int vclhip_set_device(int device);
int vclhip_get_device(int* device);
/* name (NUL-terminated, truncated to name_len), HBM bytes, CUs, L2 bytes, last-level cache bytes */
int vclhip_device_info(int device, char* name, size_t name_len, size_t* total_mem, int* num_cu,
                       int* l2_bytes, int* llc_bytes);

/* ---------------------------------------------------------- memory domain ---- */
/* memory_create (backend/memory.hpp:87-119 -> backend/cuda.hpp:103-117) */
int vclhip_malloc(void** ptr_dev, size_t bytes);
/* handle deleter (backend/cuda.hpp:85-93) */
int vclhip_free(void* ptr_dev);
/* memory_write (backend/memory.hpp:220-249 -> backend/cuda.hpp:154-177); returns when done */
int vclhip_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes, void* stream);
/* memory_read (backend/memory.hpp:261-292 -> backend/cuda.hpp:183-200); returns when done */
int vclhip_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes, void* stream);
/* memory_copy (backend/memory.hpp:140-171 -> backend/cuda.hpp:126-146); asynchronous */
int vclhip_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes, void* stream);
/* zero fill with a kernel (safe inside captured graphs); asynchronous */
int vclhip_fill_zero(void* dst_dev, size_t bytes, void* stream);
/* backend::finish() (backend/memory.hpp:54-62) for one stream */
int vclhip_stream_create(void** stream);
int vclhip_stream_destroy(void* stream);
int vclhip_stream_synchronize(void* stream);
int vclhip_device_synchronize(void);
/* events for timing on a given stream (no reference counterpart; tools/timer.hpp:90-115 is wall clock) */
int vclhip_event_create(void** event);
int vclhip_event_destroy(void* event);
int vclhip_event_record(void* event, void* stream);
int vclhip_event_elapsed_ms(void* start, void* stop, float* ms);
/* Measurement hook (no counterpart in the reference): the next SpMV (vclhip_csr_spmv_*,
 * vclhip_sell_spmv_*) or vclhip_axpby_f64 launched by the calling thread records start_event and
 * stop_event at the kernel's own start and end (hipExtLaunchKernelGGL), so that one kernel can be
 * timed without the latency of separate event packets.  The request applies to one launch. */
int vclhip_time_next_launch(void* start_event, void* stop_event);
/* Read-only streaming reference (no counterpart in the reference; measurement only): num_sums
 * workgroups (a multiple of 8, at most 2^20) each read a contiguous chunk of x[0..n) with 16-byte
 * loads and store its sum in sums[0..num_sums).  x must be 16-byte aligned.  Its kernel time over
 * 8 n bytes is the rate at which the device reads HBM; vclhip_time_next_launch applies to it. */
int vclhip_stream_read_f64(uint64_t n, const double* x_dev, double* sums_dev, uint32_t num_sums, void* stream);
/* Gather reference (no counterpart in the reference; measurement only): num_sums workgroups (a
 * multiple of 8, at most 2^20) each sum x[idx[i]] over a contiguous chunk of idx[0..m) and store
 * the sum in sums[0..num_sums); n is the length of x (an index at or beyond it reads x[0]).
 * With a CSR matrix's column indices as idx, its kernel time is that of the SpMV's x gathers
 * alone; vclhip_time_next_launch applies to it. */
int vclhip_gather_read_f64(uint32_t m, const uint32_t* idx_dev, const double* x_dev, uint32_t n, double* sums_dev,
                           uint32_t num_sums, void* stream);
/* Latency references (no counterpart in the reference; measurement only).  vclhip_empty_launch
 * launches a kernel of `workgroups` workgroups that does nothing: back to back, its launches time
 * the floor of a kernel boundary.  vclhip_grid_sync_probe launches `workgroups` workgroups (at most
 * what the device holds at once, and at most 4096; VCLHIP_ERROR_INVALID_VALUE otherwise) that
 * synchronise `steps` times device-wide (each publishes its stores to all XCDs and an arrival flag;
 * workgroup 0 collects the flags and releases the others): the price that a single-launch solver
 * would pay per iteration in place of a kernel boundary.  flags_dev holds workgroups + 2 words
 * (zeroed by the call); every wait is bounded, and flags_dev[workgroups + 1] is nonzero afterwards
 * if a workgroup gave up waiting. */
int vclhip_empty_launch(uint32_t workgroups, void* stream);
int vclhip_grid_sync_probe(uint32_t workgroups, uint32_t steps, uint32_t* flags_dev, void* stream);

/* ------------------------------------------------------------ sparse formats ---- */
/* CSR matrix on the device (compressed_matrix<T>: handle1 = row_ptr, handle2 = col, handle = values,
 * handle3 = row blocks; compressed_matrix.hpp:1088-1104).  `blocks` is the kernel's row-block plan
 * built by vclhip_csr_plan_blocks; `values` is double* or float* according to the function suffix. */
typedef struct vclhip_csr {
  uint32_t n_rows;
  uint32_t n_cols;
  uint32_t nnz;
  uint32_t num_blocks;
  const uint32_t* row_ptr;
  const uint32_t* col_idx;
  const void* values;
  const uint32_t* blocks;
} vclhip_csr;

/* SELL-C-sigma (sigma = 1) matrix on the device (sliced_ell_matrix<T>: handle1 = columns_per_block,
 * handle2 = column_indices, handle3 = block_start, handle = elements; sliced_ell_matrix.hpp:109-119).
 * Slice s, entry k, row-in-slice i is at elements[block_start[s] + k*C + i]. */
typedef struct vclhip_sell {
  uint32_t n_rows;
  uint32_t n_cols;
  uint32_t C;
  uint32_t num_slices;
  const uint32_t* columns_per_block;
  const uint32_t* column_indices;
  const uint32_t* block_start;
  const void* elements;
} vclhip_sell;

/* COO matrix on the device (coordinate_matrix<T>: handle12 = (row, col) pairs, handle = values;
 * coordinate_matrix.hpp:46-99, 185-260).  The entries are ordered by row, as copy() writes them;
 * row_ptr (n_rows + 1) and blocks are the row pointers of that order and the CSR kernel's row-block
 * plan, both built on the host by vclhip_coo_plan (the reference keeps its own auxiliary
 * group_boundaries for its device kernels, coordinate_matrix.hpp:68-91). */
typedef struct vclhip_coo {
  uint32_t n_rows;
  uint32_t n_cols;
  uint32_t nnz;
  uint32_t num_blocks;
  const uint32_t* coords;    /* 2 * nnz: row, col */
  const void* elements;
  const uint32_t* row_ptr;
  const uint32_t* blocks;
} vclhip_coo;

/* ELL matrix on the device (ell_matrix<T>: handle2 = coords, handle = elements; ell_matrix.hpp:
 * 123-171).  Column-major: entry k of row i is at [k * internal_size1 + i], k < internal_maxnnz;
 * padding entries hold the value 0. */
typedef struct vclhip_ell {
  uint32_t n_rows;
  uint32_t n_cols;
  uint32_t internal_size1;
  uint32_t internal_maxnnz;
  const uint32_t* coords;
  const void* elements;
} vclhip_ell;

/* HYB matrix on the device (hyb_matrix<T>: an ELL part of internal_ellnnz columns, handle2 =
 * coords, handle = elements, and a CSR part for the rest of each row, handle3 = row pointers,
 * handle4 = columns, handle5 = elements; hyb_matrix.hpp:136-240).  `blocks` is optional: the
 * row-block plan of the CSR part (vclhip_csr_plan_blocks on csr_rows), with which long CSR parts
 * are summed by a wavefront; with blocks = NULL one lane sums each row, ELL part then CSR part,
 * in the reference's order. */
typedef struct vclhip_hyb {
  uint32_t n_rows;
  uint32_t n_cols;
  uint32_t internal_size1;
  uint32_t internal_ellnnz;
  const uint32_t* ell_coords;
  const void* ell_elements;
  const uint32_t* csr_rows;
  const uint32_t* csr_cols;
  const void* csr_elements;
  uint32_t num_blocks;
  const uint32_t* blocks;
} vclhip_hyb;

/* A matrix of any of the formats, for the format-generic entry points below. */
#define VCLHIP_FORMAT_CSR 0
#define VCLHIP_FORMAT_SELL 1
#define VCLHIP_FORMAT_COO 2
#define VCLHIP_FORMAT_ELL 3
#define VCLHIP_FORMAT_HYB 4
typedef struct vclhip_matrix {
  int format;        /* VCLHIP_FORMAT_* */
  const void* desc;  /* the vclhip_csr / _sell / _coo / _ell / _hyb descriptor */
} vclhip_matrix;

/* Row-block plan for the CSR kernels (host function).  Replaces the reference's 1024-nnz row
 * blocks (compressed_matrix::generate_row_block_information, compressed_matrix.hpp:1152-1188) with
 * blocks sized for 256-lane workgroups: each entry packs the first row (bits 0-27) and a mode
 * (bits 28-31: 0 = every row of at most 32 entries, 1 = some longer rows, 7 = one row of more than
 * 2048 entries read directly).
 * `blocks_host` must hold n_rows + 1 entries; *num_blocks receives the count (entry num_blocks is
 * the end sentinel n_rows).  n_rows must be below 2^28. */
int vclhip_csr_plan_blocks(const uint32_t* row_ptr_host, uint32_t n_rows, uint32_t* blocks_host,
                           uint32_t* num_blocks);
/* The reference's own row-block partition (compressed_matrix.hpp:1152-1188), for API parity. */
int vclhip_csr_reference_row_blocks(const uint32_t* row_ptr_host, uint32_t n_rows, uint32_t* row_blocks_host,
                                    uint32_t* num_blocks);
/* CSR -> SELL-C conversion on the host (sliced_ell_matrix.hpp:140-214).  First call with
 * NULL outputs to get num_slices and the stored entry count. */
int vclhip_sell_from_csr_f64(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, const uint32_t* col,
                             const double* val, uint32_t* num_slices, uint64_t* stored,
                             uint32_t* columns_per_block, uint32_t* column_indices, uint32_t* block_start,
                             double* elements);
/* COO row pointers and CSR-kernel plan (host function): row_ptr_host (n_rows + 1) and blocks_host
 * (n_rows + 1) from the (row, col) pairs; fails unless the rows are nondecreasing and below
 * n_rows. */
int vclhip_coo_plan(const uint32_t* coords_host, uint32_t nnz, uint32_t n_rows, uint32_t* row_ptr_host,
                    uint32_t* blocks_host, uint32_t* num_blocks);
/* CSR -> ELL on the host as ell_matrix's copy() lays it out (ell_matrix.hpp:123-171):
 * internal_size1 = n_rows rounded up to `align`; first call with NULL outputs for the width. */
int vclhip_ell_from_csr_f64(uint32_t n_rows, uint32_t align, const uint32_t* row_ptr, const uint32_t* col,
                            const double* val, uint32_t* internal_maxnnz, uint32_t* coords, double* elements);
/* CSR -> HYB on the host as hyb_matrix's copy() splits it (hyb_matrix.hpp:136-240): the ELL width is
 * the smallest row length that at least csr_threshold * n_rows rows do not exceed, rounded up to
 * `align`; longer rows continue in the CSR part (one zero entry when it is empty).  First call with
 * NULL outputs for the sizes; when *internal_ellnnz is 0 the ELL arrays may stay NULL. */
int vclhip_hyb_from_csr_f64(uint32_t n_rows, double csr_threshold, uint32_t align, const uint32_t* row_ptr,
                            const uint32_t* col, const double* val, uint32_t* internal_ellnnz, uint32_t* csr_nnz,
                            uint32_t* ell_coords, double* ell_elements, uint32_t* csr_rows, uint32_t* csr_cols,
                            double* csr_elements);

/* ---------------------------------------------------------------- SpMV ---- */
/* y = alpha*A*x + beta*y.  linalg::prod_impl(compressed_matrix, x, alpha, y, beta)
 * (linalg/sparse_matrix_operations.hpp:90-121; host_based/sparse_matrix_operations.hpp:145-186;
 * cuda/sparse_matrix_operations.hpp:262-396).  beta is treated as zero when !(beta<0||beta>0). */
int vclhip_csr_spmv_f64(const vclhip_csr* A, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                        double alpha, double* y_dev, uint32_t y_start, uint32_t y_stride, double beta,
                        void* stream);
int vclhip_csr_spmv_f32(const vclhip_csr* A, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                        float alpha, float* y_dev, uint32_t y_start, uint32_t y_stride, float beta,
                        void* stream);
/* linalg::prod_impl(sliced_ell_matrix, ...) (host_based/sparse_matrix_operations.hpp:1796-1868;
 * cuda/sparse_matrix_operations.hpp:2247-2289).  Zero entries are skipped as in the reference. */
int vclhip_sell_spmv_f64(const vclhip_sell* A, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                         double alpha, double* y_dev, uint32_t y_start, uint32_t y_stride, double beta,
                         void* stream);
int vclhip_sell_spmv_f32(const vclhip_sell* A, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                         float alpha, float* y_dev, uint32_t y_start, uint32_t y_stride, float beta,
                         void* stream);
/* linalg::prod_impl(coordinate_matrix, ...) (host_based/sparse_matrix_operations.hpp:1221-1246;
 * cuda/sparse_matrix_operations.hpp:1067-1227): rows summed in entry order. */
int vclhip_coo_spmv_f64(const vclhip_coo* A, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                        double alpha, double* y_dev, uint32_t y_start, uint32_t y_stride, double beta,
                        void* stream);
int vclhip_coo_spmv_f32(const vclhip_coo* A, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                        float alpha, float* y_dev, uint32_t y_start, uint32_t y_stride, float beta,
                        void* stream);
/* linalg::prod_impl(ell_matrix, ...) (host_based/sparse_matrix_operations.hpp:1502-1537): zero
 * entries skipped as in the reference. */
int vclhip_ell_spmv_f64(const vclhip_ell* A, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                        double alpha, double* y_dev, uint32_t y_start, uint32_t y_stride, double beta,
                        void* stream);
int vclhip_ell_spmv_f32(const vclhip_ell* A, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                        float alpha, float* y_dev, uint32_t y_start, uint32_t y_stride, float beta,
                        void* stream);
/* linalg::prod_impl(hyb_matrix, ...) (host_based/sparse_matrix_operations.hpp:1872-1929): the ELL
 * part (zero entries skipped), then the row's CSR part. */
int vclhip_hyb_spmv_f64(const vclhip_hyb* A, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                        double alpha, double* y_dev, uint32_t y_start, uint32_t y_stride, double beta,
                        void* stream);
int vclhip_hyb_spmv_f32(const vclhip_hyb* A, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                        float alpha, float* y_dev, uint32_t y_start, uint32_t y_stride, float beta,
                        void* stream);
/* Diagonal of a matrix of any format (row_info with SPARSE_ROW_DIAGONAL, the Jacobi
 * preconditioner's setup, jacobi_precond.hpp:115-118; host_based :1156-1205 for COO): 0 where a
 * row has no nonzero diagonal entry. */
int vclhip_matrix_diag_f64(const vclhip_matrix* A, double* d_dev, void* stream);
int vclhip_matrix_diag_f32(const vclhip_matrix* A, float* d_dev, void* stream);
/* linalg::detail::row_info (linalg/sparse_matrix_operations.hpp:48-74; host_based :52-98):
 * info 0 = inf-norm, 1 = 1-norm, 2 = 2-norm, 3 = diagonal (0 when absent). */
int vclhip_csr_row_info_f64(const vclhip_csr* A, double* out_dev, int info, void* stream);
int vclhip_csr_row_info_f32(const vclhip_csr* A, float* out_dev, int info, void* stream);
/* Gradient of x = A^{-1} b with respect to A's values on A's pattern, given the adjoint solution
 * lambda = A^{-T} dL/dx: dA[k] = -lambda[row(k)] * x[col(k)] (dL/db is lambda itself).  One kernel
 * over the row-block plan.  No reference counterpart (the differentiable solve is new here). */
int vclhip_csr_adjoint_grad_f64(const vclhip_csr* A, const double* lambda, const double* x, double* dA,
                                void* stream);
int vclhip_csr_adjoint_grad_f32(const vclhip_csr* A, const float* lambda, const float* x, float* dA, void* stream);

/* ----------------------------------------------------------- vector ops ---- */
/* inner_prod_impl / inner_prod_cpu (linalg/vector_operations.hpp:500-590;
 * host_based/vector_operations.hpp:556-575).  Result written to *result_host; synchronizes. */
int vclhip_dot_f64(uint32_t n, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                   const double* y_dev, uint32_t y_start, uint32_t y_stride, double* result_host, void* stream);
int vclhip_dot_f32(uint32_t n, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                   const float* y_dev, uint32_t y_start, uint32_t y_stride, float* result_host, void* stream);
/* FP32 inputs, FP64 accumulation (no reference counterpart: the mixed-precision CG's inner dots) */
int vclhip_dot_f32_acc64(uint32_t n, const float* x_dev, const float* y_dev, double* result_host, void* stream);
/* norm_2_impl / norm_2_cpu (linalg/vector_operations.hpp:812-840; host_based :820-833) */
int vclhip_norm2_f64(uint32_t n, const double* x_dev, uint32_t x_start, uint32_t x_stride,
                     double* result_host, void* stream);
int vclhip_norm2_f32(uint32_t n, const float* x_dev, uint32_t x_start, uint32_t x_stride,
                     float* result_host, void* stream);
/* z = a*x + b*y  (avbv, linalg/vector_operations.hpp:106-141; host_based :127-194).
 * Contiguous vectors; z may alias x or y. */
int vclhip_axpby_f64(uint32_t n, double a, const double* x_dev, double b, const double* y_dev, double* z_dev,
                     void* stream);
int vclhip_axpby_f32(uint32_t n, float a, const float* x_dev, float b, const float* y_dev, float* z_dev,
                     void* stream);
/* strided form: vec1 = vec2*alpha + vec3*beta with (start,stride) per vector */
int vclhip_avbv_f64(uint32_t n, double* v1_dev, uint32_t s1, uint32_t i1,
                    const double* v2_dev, uint32_t s2, uint32_t i2, double alpha,
                    const double* v3_dev, uint32_t s3, uint32_t i3, double beta, void* stream);
int vclhip_avbv_f32(uint32_t n, float* v1_dev, uint32_t s1, uint32_t i1, const float* v2_dev, uint32_t s2,
                    uint32_t i2, float alpha, const float* v3_dev, uint32_t s3, uint32_t i3, float beta,
                    void* stream);
/* vec = alpha (vector_assign, linalg/vector_operations.hpp:208-240) */
int vclhip_fill_f64(uint32_t n, double* v_dev, uint32_t start, uint32_t stride, double alpha, void* stream);
int vclhip_fill_f32(uint32_t n, float* v_dev, uint32_t start, uint32_t stride, float alpha, void* stream);
/* z = x / y element-wise (element_op with op_div, linalg/vector_operations.hpp:254-285) */
int vclhip_element_div_f64(uint32_t n, const double* x_dev, const double* y_dev, double* z_dev, void* stream);
int vclhip_element_div_f32(uint32_t n, const float* x_dev, const float* y_dev, float* z_dev, void* stream);
/* convert (linalg/vector_operations.hpp:50-70; host_based :67-83) */
int vclhip_convert_f64_to_f32(uint32_t n, const double* src_dev, float* dst_dev, void* stream);
int vclhip_convert_f32_to_f64(uint32_t n, const float* src_dev, double* dst_dev, void* stream);
/* sum |x_i| (norm_1, linalg/vector_operations.hpp:673-700; host_based :708-733), to the host */
int vclhip_asum_f64(uint32_t n, const double* x_dev, uint32_t x_start, uint32_t x_stride, double* result_host,
                    void* stream);
int vclhip_asum_f32(uint32_t n, const float* x_dev, uint32_t x_start, uint32_t x_stride, float* result_host,
                    void* stream);
/* index of the first largest |x_i| (index_norm_inf, host_based/vector_operations.hpp:891-950); when
 * no entry is above zero the reference returns the vector's start offset, and so does this. */
int vclhip_iamax_f64(uint32_t n, const double* x_dev, uint32_t x_start, uint32_t x_stride, uint32_t* index_host,
                     void* stream);
int vclhip_iamax_f32(uint32_t n, const float* x_dev, uint32_t x_start, uint32_t x_stride, uint32_t* index_host,
                     void* stream);
/* x = alpha x (av with vec1 = vec2, host_based/vector_operations.hpp:88-123) */
int vclhip_scal_f64(uint32_t n, double* x_dev, uint32_t x_start, uint32_t x_stride, double alpha, void* stream);
int vclhip_scal_f32(uint32_t n, float* x_dev, uint32_t x_start, uint32_t x_stride, float alpha, void* stream);
/* y = x (vector assignment) */
int vclhip_copy_f64(uint32_t n, const double* x_dev, uint32_t x_start, uint32_t x_stride, double* y_dev,
                    uint32_t y_start, uint32_t y_stride, void* stream);
int vclhip_copy_f32(uint32_t n, const float* x_dev, uint32_t x_start, uint32_t x_stride, float* y_dev,
                    uint32_t y_start, uint32_t y_stride, void* stream);
/* x <-> y (vector_swap, linalg/vector_operations.hpp:218-245; host_based :302-327) */
int vclhip_swap_f64(uint32_t n, double* x_dev, uint32_t x_start, uint32_t x_stride, double* y_dev,
                    uint32_t y_start, uint32_t y_stride, void* stream);
int vclhip_swap_f32(uint32_t n, float* x_dev, uint32_t x_start, uint32_t x_stride, float* y_dev,
                    uint32_t y_start, uint32_t y_stride, void* stream);
/* x' = c x + s y, y' = c y - s x (plane_rotation, host_based/vector_operations.hpp:1080-1110) */
int vclhip_rot_f64(uint32_t n, double* x_dev, uint32_t x_start, uint32_t x_stride, double* y_dev, uint32_t y_start,
                   uint32_t y_stride, double c, double s, void* stream);
int vclhip_rot_f32(uint32_t n, float* x_dev, uint32_t x_start, uint32_t x_stride, float* y_dev, uint32_t y_start,
                   uint32_t y_stride, float c, float s, void* stream);

/* -------------------------------------------------- pipelined operations ---- */
/* Buffer convention of the reference (cg.hpp:144-172, bicgstab.hpp:114-189): inner products are
 * left in chunks of `chunk` doubles; the host sums each chunk.  These entry points write each
 * completed inner product into the first entry of its chunk and zero the rest of that chunk.
 * All vectors contiguous, length A->n_rows. */
/* linalg::pipelined_cg_vector_update (iterative_operations.hpp:58-87; host_based :378-418):
 * x += alpha p; r -= alpha Ap; p = r + beta p; buf[0..chunk) <- <r,r> */
int vclhip_pipelined_cg_vector_update_f64(uint32_t n, double* x_dev, double alpha, double* p_dev, double* r_dev,
                                          const double* Ap_dev, double beta, double* buf_dev, uint32_t chunk,
                                          void* stream);
/* linalg::pipelined_cg_prod (iterative_operations.hpp:96-122; host_based :427-487):
 * Ap = A p; chunk 1 <- <Ap,Ap>, chunk 2 <- <p,Ap> */
int vclhip_pipelined_cg_prod_csr_f64(const vclhip_csr* A, const double* p_dev, double* Ap_dev, double* buf_dev,
                                     uint32_t chunk, void* stream);
int vclhip_pipelined_cg_prod_sell_f64(const vclhip_sell* A, const double* p_dev, double* Ap_dev, double* buf_dev,
                                      uint32_t chunk, void* stream);
/* linalg::pipelined_bicgstab_update_s (iterative_operations.hpp:133-161; host_based :518-563):
 * alpha = sum(chunk 0)/sum(chunk 3) on the device; s = r - alpha Ap; buf[offset] <- <s,s> */
int vclhip_pipelined_bicgstab_update_s_f64(uint32_t n, double* s_dev, const double* r_dev, const double* Ap_dev,
                                           double* buf_dev, uint32_t chunk, uint32_t offset, void* stream);
/* linalg::pipelined_bicgstab_vector_update (iterative_operations.hpp:170-198; host_based :572-621) */
int vclhip_pipelined_bicgstab_vector_update_f64(uint32_t n, double* x_dev, double alpha, double* p_dev,
                                                double omega, const double* s_dev, double* r_dev,
                                                const double* As_dev, double beta, const double* Ap_dev,
                                                const double* r0star_dev, double* buf_dev, uint32_t chunk,
                                                void* stream);
/* linalg::pipelined_bicgstab_prod (iterative_operations.hpp:207-236; host_based :629-701):
 * Ap = A p; chunk 1 <- <Ap,Ap>, chunk 2 <- <p,Ap>, buf[offset..offset+chunk) <- <Ap,r0*> */
int vclhip_pipelined_bicgstab_prod_csr_f64(const vclhip_csr* A, const double* p_dev, double* Ap_dev,
                                           const double* r0star_dev, double* buf_dev, uint32_t chunk,
                                           uint32_t offset, void* stream);
int vclhip_pipelined_bicgstab_prod_sell_f64(const vclhip_sell* A, const double* p_dev, double* Ap_dev,
                                            const double* r0star_dev, double* buf_dev, uint32_t chunk,
                                            uint32_t offset, void* stream);
/* linalg::pipelined_gmres_prod (iterative_operations.hpp:393-418; host_based :931-937), which is
 * pipelined_cg_prod: Ap = A p; chunk 1 <- <Ap,Ap>, chunk 2 <- <p,Ap> */
int vclhip_pipelined_gmres_prod_csr_f64(const vclhip_csr* A, const double* p_dev, double* Ap_dev, double* buf_dev,
                                        uint32_t chunk, void* stream);
int vclhip_pipelined_gmres_prod_sell_f64(const vclhip_sell* A, const double* p_dev, double* Ap_dev,
                                         double* buf_dev, uint32_t chunk, void* stream);
/* The same fused products for a matrix of any format (coordinate_matrix, ell_matrix,
 * sliced_ell_matrix, hyb_matrix overloads: host_based/iterative_operations.hpp:113-365,
 * :445-518, :629-701). */
int vclhip_pipelined_cg_prod_f64(const vclhip_matrix* A, const double* p_dev, double* Ap_dev, double* buf_dev,
                                 uint32_t chunk, void* stream);
int vclhip_pipelined_bicgstab_prod_f64(const vclhip_matrix* A, const double* p_dev, double* Ap_dev,
                                       const double* r0star_dev, double* buf_dev, uint32_t chunk, uint32_t offset,
                                       void* stream);
/* The Krylov basis of the GMRES operations: vector i at basis_dev + i * ld (ld >= n, the
 * reference's internal_size). */
/* linalg::pipelined_gmres_normalize_vk (iterative_operations.hpp:248-277; host_based :734-779):
 * norm = sqrt(sum buf[chunk .. 2 chunk)); R[offset_in_R] = norm; v_k /= norm;
 * r_dot_vk_buf[chunk_offset .. + chunk) <- <residual, v_k> */
int vclhip_pipelined_gmres_normalize_vk_f64(uint32_t n, double* v_k_dev, const double* residual_dev, double* R_dev,
                                            size_t offset_in_R, const double* buf_dev, double* r_dot_vk_buf_dev,
                                            uint32_t chunk, size_t chunk_offset, void* stream);
/* linalg::pipelined_gmres_gram_schmidt_stage1 (iterative_operations.hpp:286-311; host_based
 * :787-842): chunk j of vi_in_vk_buf <- <v_j, v_k> for j < k, over all n entries */
int vclhip_pipelined_gmres_gram_schmidt_stage1_f64(const double* basis_dev, uint32_t n, size_t ld, uint32_t k,
                                                   double* vi_in_vk_buf_dev, uint32_t chunk, void* stream);
/* linalg::pipelined_gmres_gram_schmidt_stage2 (iterative_operations.hpp:321-349; host_based
 * :851-893): h_j = sum of chunk j of vi_in_vk_buf; v_k -= sum_j h_j v_j;
 * R[j + k krylov_dim] = h_j; buf[chunk .. 2 chunk) <- ||v_k||^2 */
int vclhip_pipelined_gmres_gram_schmidt_stage2_f64(double* basis_dev, uint32_t n, size_t ld, uint32_t k,
                                                   const double* vi_in_vk_buf_dev, double* R_dev,
                                                   uint32_t krylov_dim, double* buf_dev, uint32_t chunk,
                                                   void* stream);
/* linalg::pipelined_gmres_update_result (iterative_operations.hpp:356-389; host_based :898-926):
 * x += c[0] residual + sum_{j=1}^{k-1} c[j] v_{j-1}  (c[0] is applied also when k == 0) */
int vclhip_pipelined_gmres_update_result_f64(double* x_dev, const double* residual_dev, const double* basis_dev,
                                             uint32_t n, size_t ld, const double* coef_dev, uint32_t k,
                                             void* stream);

/* ------------------------------------------------------------- solvers ---- */
/* Solver parameters: the fields of cg_tag (cg.hpp:48-87), bicgstab_tag (bicgstab.hpp:47-90) and
 * mixed_precision_cg_tag (mixed_precision_cg.hpp:47-83), plus launch controls. */
typedef struct vclhip_solver_params {
  double tol;                        /* relative tolerance */
  double abs_tol;                    /* absolute tolerance (0 = off) */
  uint32_t max_iters;
  uint32_t max_iters_before_restart; /* BiCGStab with a preconditioner (bicgstab.hpp:366-370) */
  float inner_tol;                   /* mixed CG: restart when inner <r,r> < inner_tol * initial (squared norms) */
  uint32_t chunk;                    /* iterations enqueued between convergence polls (0 = default 16) */
  int use_graph;                     /* capture one chunk as a hipGraph and replay it */
  int mixed_mode;                    /* mixed CG: 0 = restart as the reference, 1 = reliable update (keeps p) */
  int variant;                       /* BiCGStab with Jacobi: 0 = the reference's recurrences (bicgstab.hpp:447-480),
                                        1 = fully pipelined (fewer passes; rounding differs).
                                        BiCGStab without a preconditioner: 0 = automatic (a small
                                        system runs four kernels per iteration), 1 = six; the
                                        results are the same bit for bit.
                                        CG: 0 = automatic (a small system on a DIA operator runs one
                                        kernel per iteration), 1 = at least two kernels per
                                        iteration; the results are the same bit for bit */
  double* history_host;              /* optional: per-iteration convergence quantity (may be NULL) */
  uint32_t history_len;
  /* optional monitor, called once per iteration with the device solution and the relative residual
   * estimate; returning nonzero stops the solve (cg.hpp:174-175, cg_solver::set_monitor :420-424).
   * A monitor forces one host synchronization per iteration.  x_dev points to the solution vector's
   * element type (double or float). */
  int (*monitor)(const void* x_dev, double rel_residual, void* user);
  void* monitor_user;
  int inner_format;                  /* mixed CG on a CSR matrix: storage of the FP32 inner copy,
                                        VCLHIP_INNER_AUTO (DIA when the matrix has at most 32
                                        diagonals storing at most 25% zeros; else SELL-64 when it
                                        stores at most 25% padding, with 16-bit column offsets when
                                        every |col - row| <= 32767; else CSR), VCLHIP_INNER_CSR,
                                        VCLHIP_INNER_SELL64 (32-bit columns), VCLHIP_INNER_SELL64_D16
                                        or VCLHIP_INNER_DIA.  DIA and SELL-64 sum each row in one
                                        lane in column order and give the same results. */
  uint32_t krylov_dim;               /* GMRES: steps per restart cycle (gmres_tag default 20) */
  int operator_format;               /* CG, BiCGStab, GMRES on a CSR matrix: VCLHIP_OPERATOR_AUTO
                                        (default) solves on a DIA copy in workspace memory when the
                                        matrix has at most 32 diagonals storing at most 25% zeros,
                                        every stored value is nonzero and not NaN, every row's columns
                                        strictly increase, and the copy takes at most a quarter of the
                                        free device memory -- the same products bit for bit (also
                                        for non-finite x) with fewer bytes each; the inner products
                                        fused into them are summed per 256-row tile, which is the CSR
                                        plan's grouping when no row has more than 8 entries (then the
                                        whole solve is the same bit for bit; with longer rows it
                                        agrees to rounding).  A copy of 2^18 rows or more whose
                                        diagonals each hold at most 15 distinct values is read as
                                        one byte per entry into a table of those values (the same
                                        values, so the same products bit for bit);
                                        VCLHIP_OPERATOR_DIA_UNCODED is AUTO with the values read
                                        as stored; VCLHIP_OPERATOR_AS_GIVEN always uses the CSR
                                        matrix; other values are rejected */
  const struct vclhip_dia* dia;      /* optional, with VCLHIP_OPERATOR_AUTO or _DIA_UNCODED: the detection and DIA
                                        copy made once by vclhip_dia_create_* for this CSR matrix,
                                        used instead of detecting and copying in every solve (the
                                        matrix's arrays must be the ones it was made from) */
} vclhip_solver_params;
#define VCLHIP_OPERATOR_AUTO 0
#define VCLHIP_OPERATOR_AS_GIVEN 1
#define VCLHIP_OPERATOR_DIA_UNCODED 2

#define VCLHIP_INNER_AUTO 0
#define VCLHIP_INNER_CSR 1
#define VCLHIP_INNER_SELL64 2
#define VCLHIP_INNER_SELL64_D16 3
#define VCLHIP_INNER_DIA 4
#define VCLHIP_INNER_DIA_CODED 5 /* reported in result.operator_used: the DIA inner copy read through
                                    value tables (2^18 rows or more, at most 15 values a diagonal) */

typedef struct vclhip_solver_result {
  uint32_t iters;     /* tag.iters() */
  double error;       /* tag.error() */
  uint32_t restarts;  /* mixed CG outer corrections / BiCGStab restarts */
  int converged;
  int operator_used;  /* the storage the products ran on: VCLHIP_USED_AS_GIVEN, VCLHIP_USED_DIA or
                         VCLHIP_USED_DIA_CODED (the DIA copy read through its value tables);
                         mixed CG: its FP32 inner copy, VCLHIP_INNER_CSR (the matrix's own format),
                         VCLHIP_INNER_SELL64, VCLHIP_INNER_SELL64_D16, VCLHIP_INNER_DIA or
                         VCLHIP_INNER_DIA_CODED */
  uint32_t kernels_per_iter;  /* CG: kernels launched per iteration after the first (1 in fused mode,
                                 2 in compact mode, else 3); BiCGStab: 4 in compact mode, 6 pipelined,
                                 8 with the reference's Jacobi recurrences; 0 from the other solvers */
} vclhip_solver_result;
#define VCLHIP_USED_AS_GIVEN 0
#define VCLHIP_USED_DIA 1
#define VCLHIP_USED_DIA_CODED 2

/* Fill params with the defaults of cg_tag() / bicgstab_tag() / mixed_precision_cg_tag(). */
void vclhip_solver_params_default(vclhip_solver_params* params);

/* A DIA copy of a CSR matrix kept across solves (no reference counterpart; the solvers otherwise
 * detect the diagonals and copy the matrix in every solve).  vclhip_dia_create_* runs that
 * detection once: *dia always receives an object, holding the copy in device memory when A is a
 * candidate (at most 32 diagonals storing at most 25% zeros, nonzero non-NaN values, strictly
 * increasing columns, at most a quarter of the free device memory) and only the negative answer
 * otherwise.  Pass it as params->dia to solves of this matrix; it stays valid while A's arrays and
 * values are unchanged, and the caller rebuilds it after changing them.  Synchronizes the stream. */
typedef struct vclhip_dia vclhip_dia;
int vclhip_dia_create_f64(const vclhip_csr* A, vclhip_dia** dia, void* stream);
int vclhip_dia_create_f32(const vclhip_csr* A, vclhip_dia** dia, void* stream);
/* 1 when the object holds a DIA copy (the solves then run on it), else 0; nd receives its diagonal count. */
int vclhip_dia_info(const vclhip_dia* dia, int* built, uint32_t* nd);
int vclhip_dia_destroy(vclhip_dia* dia);

#define VCLHIP_PRECOND_NONE 0
#define VCLHIP_PRECOND_JACOBI 1   /* jacobi_precond<compressed_matrix> (jacobi_precond.hpp:103-130) */

/* solve(A, b, cg_tag [, jacobi_precond]) (cg.hpp:338-375).  Without a preconditioner this is the
 * pipelined Chronopoulos-Gear CG of cg.hpp:128-187 with its stopping rule; with Jacobi it is a
 * pipelined preconditioned CG stopping on the M^-1 norm like cg.hpp:283,306-309.
 * x_dev receives the solution (the solver starts from x = 0). */
int vclhip_cg_solve_csr_f64(const vclhip_csr* A, const double* b_dev, double* x_dev, int precond,
                            const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
/* The same solve without any allocation -- no counterpart in the reference; for callers that
 * cannot allocate device or pinned memory (serving loops that must not allocate, harnesses that
 * forbid it).  Its vectors, partial sums and state live in a file-scope scratch array of 32 MiB
 * that each device the library runs on holds until the process ends: n rows need about 24 n bytes
 * plus 257 KiB (32 n with Jacobi, plus 8 bytes per iteration of a requested history), so up to
 * about 1.39 M rows (1.04 M with Jacobi); a larger system returns VCLHIP_ERROR_TOO_LARGE.  The DIA copy of the operator (VCLHIP_OPERATOR_AUTO) is made in what is
 * left of the array, or not at all, unless params->dia brings one.  params->use_graph must be 0.
 * The host waits for the device after every chunk of params->chunk iterations (no pinned memory,
 * no events).  Results equal vclhip_cg_solve_csr_f64's bit for bit.  Calls on one device share its
 * array, so calls that overlap on one device (other host threads, or streams that run concurrently)
 * race; calls made one after another do not, whatever the previous call left there. */
int vclhip_cg_solve_csr_f64_noalloc(const vclhip_csr* A, const double* b_dev, double* x_dev, int precond,
                                    const vclhip_solver_params* params, vclhip_solver_result* result,
                                    void* stream);
int vclhip_cg_solve_sell_f64(const vclhip_sell* A, const double* b_dev, double* x_dev, int precond,
                             const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
/* solve(A, b, bicgstab_tag [, jacobi_precond]) (bicgstab.hpp:495-533).  Pipelined BiCGStab of
 * bicgstab.hpp:97-215; unlike the reference, the last x update is applied before returning, so x
 * matches the residual estimate that stopped the iteration.  With Jacobi: the left-preconditioned
 * iteration of bicgstab.hpp:398-489 (criterion ||M^-1 r|| / ||b||, restarts every
 * max_iters_before_restart iterations or on breakdown), with the reference's recurrences unless
 * params->variant == 1. */
int vclhip_bicgstab_solve_csr_f64(const vclhip_csr* A, const double* b_dev, double* x_dev, int precond,
                                  const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_bicgstab_solve_sell_f64(const vclhip_sell* A, const double* b_dev, double* x_dev, int precond,
                                   const vclhip_solver_params* params, vclhip_solver_result* result,
                                   void* stream);
/* FP32 solves: FP32 matrices and vectors, inner products accumulated in FP64, scalars in FP64
 * (solve(A, b, tag) with compressed_matrix<float> / sliced_ell_matrix<float>). */
int vclhip_cg_solve_csr_f32(const vclhip_csr* A, const float* b_dev, float* x_dev, int precond,
                            const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_cg_solve_sell_f32(const vclhip_sell* A, const float* b_dev, float* x_dev, int precond,
                             const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_bicgstab_solve_csr_f32(const vclhip_csr* A, const float* b_dev, float* x_dev, int precond,
                                  const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_bicgstab_solve_sell_f32(const vclhip_sell* A, const float* b_dev, float* x_dev, int precond,
                                   const vclhip_solver_params* params, vclhip_solver_result* result,
                                   void* stream);

/* Mixed-precision CG: replaces linalg::solve(A, b, mixed_precision_cg_tag)
 * (mixed_precision_cg.hpp:95-193).  A holds FP64 values; the inner iteration runs SpMV, the Jacobi
 * scaling and the vector updates in FP32 with inner products accumulated in FP64, and FP64 is used
 * only for x += x32, r = b - A x and the convergence norm (true residual ||b - A x|| / ||b||).
 * values32: FP32 copy of A's values (nnz floats for CSR, the stored elements for SELL), or NULL to
 * convert inside the call as the reference does (mixed_precision_cg.hpp:127-137).
 * params->mixed_mode 0: the reference's schedule -- FP32 classic CG restarted from the FP64 residual
 * when the FP32 <r,r> has fallen below inner_tol times its value at the last restart (squared
 * norms, mixed_precision_cg.hpp:161); no preconditioner.  1: reliable updates -- FP32 pipelined
 * CG whose residual is replaced by the FP64 one under the same inner_tol test while the search
 * direction is kept (fewer iterations than restarting; precond may be Jacobi).
 * result->iters counts FP32 iterations, result->restarts the FP64 corrections, result->error the
 * last FP64 relative residual. */
int vclhip_mixed_cg_solve_csr(const vclhip_csr* A, const float* values32, const double* b_dev, double* x_dev,
                              int precond, const vclhip_solver_params* params, vclhip_solver_result* result,
                              void* stream);
int vclhip_mixed_cg_solve_sell(const vclhip_sell* A, const float* elements32, const double* b_dev, double* x_dev,
                               int precond, const vclhip_solver_params* params, vclhip_solver_result* result,
                               void* stream);

/* solve(A, b, gmres_tag) without a preconditioner (gmres.hpp:374-383): the pipelined GMRES of
 * gmres.hpp:182-370, restarted every params->krylov_dim steps (1 .. 1024), at most
 * gmres_tag::max_restarts() + 1 cycles for params->max_iters (gmres.hpp:75-81); the convergence
 * test rho_0 / ||b|| < tol on the recomputed residual at the start of a cycle, as the reference.
 * result->iters counts Krylov steps as tag.iters(), result->error is tag.error() (the estimate of
 * the last cycle), result->restarts the cycles run.  Inner products <v_j, v_k> run over all n
 * entries as the reference's device kernels compute them.  params->history_host (optional)
 * receives the estimate after every step; params->monitor is called once per cycle
 * (gmres.hpp:366-367).  FP32: FP32 basis and vectors, inner products accumulated in FP64. */
int vclhip_gmres_solve_csr_f64(const vclhip_csr* A, const double* b_dev, double* x_dev,
                               const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_gmres_solve_sell_f64(const vclhip_sell* A, const double* b_dev, double* x_dev,
                                const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_gmres_solve_csr_f32(const vclhip_csr* A, const float* b_dev, float* x_dev,
                               const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_gmres_solve_sell_f32(const vclhip_sell* A, const float* b_dev, float* x_dev,
                                const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);

/* The solvers for a matrix of any format: solve(A, b, cg_tag / bicgstab_tag / gmres_tag) with
 * A a coordinate_matrix, ell_matrix or hyb_matrix (cg.hpp:204-254, bicgstab.hpp:231-285 and
 * gmres.hpp:384-432 send them to the same pipelined solvers), as well as CSR and SELL.  Same
 * semantics as the per-format entry points above; Jacobi is available for every format. */
int vclhip_cg_solve_f64(const vclhip_matrix* A, const double* b_dev, double* x_dev, int precond,
                        const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_cg_solve_f32(const vclhip_matrix* A, const float* b_dev, float* x_dev, int precond,
                        const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_bicgstab_solve_f64(const vclhip_matrix* A, const double* b_dev, double* x_dev, int precond,
                              const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_bicgstab_solve_f32(const vclhip_matrix* A, const float* b_dev, float* x_dev, int precond,
                              const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_gmres_solve_f64(const vclhip_matrix* A, const double* b_dev, double* x_dev,
                           const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);
int vclhip_gmres_solve_f32(const vclhip_matrix* A, const float* b_dev, float* x_dev,
                           const vclhip_solver_params* params, vclhip_solver_result* result, void* stream);

/* ---------------------------------------------------------------- batched CG */
/* Batched CG: independent SPD systems, the diagonal blocks of one block-diagonal CSR matrix with
 * global column indices; system i is rows [sys_ptr[i], sys_ptr[i+1]).  No counterpart in the
 * reference (its nearest is one solve(A_i, b_i, cg_tag) per system, cg.hpp:128-187, whose
 * recurrences each system follows).  One persistent launch: each workgroup takes systems from a
 * work queue (largest first) and runs CG on one system at a time with p and A p in LDS, so a
 * system has at most 4096 rows (VCLHIP_ERROR_TOO_LARGE otherwise).  A column outside a system's
 * block sets its converged flag to 2. */
typedef struct vclhip_batch_plan vclhip_batch_plan;
/* Launch order and SpMV choice of a batch (host arrays sys_ptr, num_systems+1 entries, and
 * optionally row_ptr).  With row_ptr the plan also holds the row blocks of every system (the CSR
 * kernels' partition, vclhip_csr_plan_blocks), which the staged SpMV reads, and that SpMV is the
 * choice where untuned; the batch matrix passed to the solves must then have these row pointers.
 * Without row_ptr every system takes L lanes per row (about one lane per nonzero of an average
 * row).  Allocates device memory, among it an A p workspace of total rows doubles and, at the
 * first solve or autotune, 2 bytes per matrix entry for the column indices as 16-bit offsets from
 * their system's first row (refilled from A at every solve; the SpMVs read those): one solve at a
 * time per plan. */
int vclhip_batch_plan_create(uint32_t num_systems, const uint32_t* sys_ptr_host, const uint32_t* row_ptr_host,
                             vclhip_batch_plan** plan);
int vclhip_batch_plan_destroy(vclhip_batch_plan* plan);
/* Autotuner: for every size class (floor(log2 rows)) times 1, 2, 4, 8 and 16 lanes per row and,
 * when the plan has row blocks, the staged SpMV on up to 512 of the batch's own systems
 * (12 iterations, 3 repetitions) and keeps the fastest, for the value type given
 * (single_precision = 0: double, 1: float).  A is the batch matrix (values of that type). */
int vclhip_batch_plan_autotune(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                               int single_precision, void* stream);
/* The tuned choice per size class (16 entries): 0..4 = log2(lanes per row), 5 = staged over row
 * blocks, -1 = untuned. */
int vclhip_batch_plan_lanes(const vclhip_batch_plan* plan, int single_precision, int32_t* lanes_log2_by_class);
/* Solves every system: per-system relative tolerance params->tol (and abs_tol) on the recursive
 * residual, params->max_iters; iters/errors/converged are device arrays of num_systems entries
 * (errors: sqrt(<r,r>/||b_i||^2)).  params->use_graph with a non-null stream captures the launch as
 * a hipGraph on first use and replays it while the arguments stay the same.  Asynchronous on the
 * stream. */
int vclhip_batched_cg_solve_f64(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                                const double* b_dev, double* x_dev, const vclhip_solver_params* params,
                                uint32_t* iters, double* errors, uint32_t* converged, void* stream);
int vclhip_batched_cg_solve_f32(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                                const float* b_dev, float* x_dev, const vclhip_solver_params* params,
                                uint32_t* iters, double* errors, uint32_t* converged, void* stream);

/* ---------------------------------------------------------------- MatrixMarket files */
/* io::read_matrix_market_file (io/matrix_market.hpp:102-370) into CSR host arrays.  Coordinate
 * files with a real, pattern or complex field (pattern entries are 1, complex entries keep the real
 * part) and a general or symmetric structure (a symmetric file also sets (col, row)); '%' comment
 * and blank lines are skipped; indices are shifted by index_base and checked; reading stops after
 * the entry count of the size line; an entry given twice keeps its last value.  Columns come out
 * sorted within each row.  *lines_read receives the number of lines read (the reference reader's
 * return value).  The matrix stays allocated until vclhip_mm_free; vclhip_mm_csr copies it into
 * row_ptr (n_rows + 1), col and val (nnz each). */
typedef struct vclhip_mm_matrix vclhip_mm_matrix;
int vclhip_mm_read(const char* path, int64_t index_base, vclhip_mm_matrix** matrix, uint32_t* n_rows,
                   uint32_t* n_cols, uint32_t* nnz, int* symmetric, int64_t* lines_read);
int vclhip_mm_csr(const vclhip_mm_matrix* matrix, uint32_t* row_ptr, uint32_t* col, double* val);
int vclhip_mm_free(vclhip_mm_matrix* matrix);
/* io::write_matrix_market_file (io/matrix_market.hpp:375-433): a "coordinate real general" file,
 * one "row col value" line per entry in row order, values in the default iostream format (%g). */
int vclhip_mm_write(const char* path, int64_t index_base, uint32_t n_rows, uint32_t n_cols, const uint32_t* row_ptr,
                    const uint32_t* col, const double* val);

/* ---------------------------------------------------------------- tools ---- */
/* Project RNG: std::mt19937_64(seed), u = (draw >> 11) * 2^-53; U[a,b] = a + (b-a) u; normals by
 * Box-Muller on consecutive pairs (r = sqrt(-2 ln(1-u1)), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2)).
 * (tools/random.hpp:44-48 is the reference's generator; its sequence is not reproduced.) */
int vclhip_random_uniform(uint64_t seed, uint64_t n, double a, double b, double* out_host);
int vclhip_random_normal(uint64_t seed, uint64_t n, double mean, double sigma, double* out_host);
/* tools::generate_fdm_laplace (tools/matrix_generation.hpp:47-86): 2-D 5-point, row = i + j*px,
 * diagonal 4, neighbours -1.  Arrays: row_ptr n+1, col/val 5n; *nnz receives the count. */
int vclhip_gen_fdm_laplace_2d(uint32_t px, uint32_t py, uint32_t* row_ptr, uint32_t* col, double* val,
                              uint32_t* nnz);
/* 3-D 7-point Laplacian, row = i + nx (j + ny k), diagonal 6, neighbours -1.  col/val 7n. */
int vclhip_gen_laplace_3d(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* row_ptr, uint32_t* col, double* val,
                          uint32_t* nnz);
/* Variable-coefficient diffusion -div(k grad u) on an nx*ny*nz (nz = 1: 2-D, 5-point) cell grid,
 * unit spacing: face coefficient = 2 k_a k_b / (k_a + k_b); a Dirichlet boundary face adds the
 * cell's k to the diagonal.  k_host holds one coefficient per cell (lexicographic). */
int vclhip_gen_diffusion(uint32_t nx, uint32_t ny, uint32_t nz, const double* k_host, uint32_t* row_ptr,
                         uint32_t* col, double* val, uint32_t* nnz);
/* Random nonsymmetric matrix: per row L ~ P(k) ~ k^-alpha on [1, max_len] (inverse CDF) distinct
 * off-diagonal columns drawn uniformly, values N(0,1), diagonal = diag_factor * sum|a_ij|; columns
 * sorted.  First call with NULL arrays returns *nnz; arrays: row_ptr n+1, col/val nnz. */
int vclhip_gen_powerlaw(uint32_t n, uint32_t max_len, double alpha, double diag_factor, uint64_t seed,
                        uint32_t* row_ptr, uint32_t* col, double* val, uint32_t* nnz);
/* Batch of graph Laplacians: system i has n_i ~ U{n_min..n_max} rows, round(n_i*avg_degree/2)
 * undirected edges between distinct uniform endpoints (repeats merged by adding weights), weights
 * U[w_lo, w_hi]; A_i = L_i + shift*I.  Block-diagonal CSR with global column indices; sys_ptr has
 * num_systems+1 entries.  First call with NULL arrays returns *total_rows and *nnz. */
int vclhip_gen_graph_laplacian_batch(uint32_t num_systems, uint32_t n_min, uint32_t n_max, double avg_degree,
                                     double w_lo, double w_hi, double shift, uint64_t seed, uint32_t* sys_ptr,
                                     uint32_t* row_ptr, uint32_t* col, double* val, uint32_t* total_rows,
                                     uint32_t* nnz);

#ifdef __cplusplus
}
#endif

#endif /* VCLHIP_H */
