// ref_shim.cpp -- C entry points over the REAL ViennaCL 1.7.1 host_based
// backend, compiled from the reference headers where they lie (oracle/Makefile
// passes -I<reference root>).  TEST INFRASTRUCTURE ONLY: used to pin the oracle
// restatement (vcl_oracle.c) and as the "reference" CPU baseline in bench.py.
// The output library lives in oracle/_ref/ and is never linked by the product.
//
// Everything here is glue written for this project: it converts raw CSR arrays
// into the reference's own types and calls the reference's public API.

#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/coordinate_matrix.hpp"
#include "viennacl/ell_matrix.hpp"
#include "viennacl/hyb_matrix.hpp"
#include "viennacl/sliced_ell_matrix.hpp"
#include "viennacl/vector.hpp"
#include "viennacl/linalg/prod.hpp"
#include "viennacl/linalg/inner_prod.hpp"
#include "viennacl/linalg/norm_2.hpp"
#include "viennacl/linalg/cg.hpp"
#include "viennacl/linalg/bicgstab.hpp"
#include "viennacl/linalg/mixed_precision_cg.hpp"
#include "viennacl/linalg/gmres.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"
#include "viennacl/linalg/iterative_operations.hpp"
#include "viennacl/tools/matrix_generation.hpp"
#include "viennacl/io/matrix_market.hpp"
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::vector<std::map<unsigned int, double> > stl_matrix;

stl_matrix mm_last;  // the last matrix read by ref_mm_read

void fill_csr(viennacl::compressed_matrix<double>& A, uint32_t n, const uint32_t* rp, const uint32_t* col,
              const double* val)
{
  std::vector<unsigned int> r(rp, rp + n + 1), c(col, col + rp[n]);
  A.set(&r[0], &c[0], val, n, n, rp[n]);
}

stl_matrix to_stl(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val)
{
  stl_matrix m(n);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
      m[i][col[k]] = val[k];
  return m;
}

void fill_sell(viennacl::sliced_ell_matrix<double>& A, uint32_t n, const uint32_t* rp, const uint32_t* col,
               const double* val)
{
  viennacl::copy(to_stl(n, rp, col, val), A);
}

void to_vcl(const double* src, viennacl::vector<double>& v)
{
  viennacl::fast_copy(src, src + v.size(), v.begin());
}

void from_vcl(const viennacl::vector<double>& v, double* dst)
{
  viennacl::fast_copy(v.begin(), v.end(), dst);
}

// Minimal matrix adapter with the clear/resize/operator() interface that
// tools::generate_fdm_laplace (tools/matrix_generation.hpp:47-86) writes into.
struct map_matrix {
  stl_matrix rows;
  struct ref {
    map_matrix* m; std::size_t r, c;
    ref& operator=(double v) { m->rows[r][(unsigned int)c] = v; return *this; }
  };
  void clear() { rows.clear(); }
  void resize(std::size_t r, std::size_t, bool) { rows.assign(r, std::map<unsigned int, double>()); }
  ref operator()(std::size_t r, std::size_t c) { ref x = {this, r, c}; return x; }
};

template <class F>
void with_format(int format, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, F f)
{
  stl_matrix m = to_stl(n, rp, col, val);
  if (format == 2) {
    viennacl::coordinate_matrix<double> A(n, n);
    viennacl::copy(m, A);
    f(A);
  } else if (format == 3) {
    viennacl::ell_matrix<double> A;
    viennacl::copy(m, A);
    f(A);
  } else {
    viennacl::hyb_matrix<double> A;
    viennacl::copy(m, A);
    f(A);
  }
}

}  // namespace

extern "C" {

void ref_set_num_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// tools/matrix_generation.hpp:47-86 -> CSR; returns nnz (arrays sized by caller: n+1, 5n)
uint32_t ref_fdm_laplace(uint32_t px, uint32_t py, uint32_t* rp, uint32_t* col, double* val)
{
  map_matrix m;
  viennacl::tools::generate_fdm_laplace(m, px, py);
  uint32_t k = 0;
  rp[0] = 0;
  for (std::size_t i = 0; i < m.rows.size(); ++i) {
    for (std::map<unsigned int, double>::const_iterator it = m.rows[i].begin(); it != m.rows[i].end(); ++it) {
      col[k] = it->first;
      val[k] = it->second;
      ++k;
    }
    rp[i + 1] = k;
  }
  return k;
}

// y = alpha*A*x + beta*y through linalg::prod_impl (linalg/sparse_matrix_operations.hpp:90-121)
void ref_csr_spmv_f64(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                      const double* x, double alpha, double* y, double beta)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vx(n), vy(n);
  to_vcl(x, vx);
  to_vcl(y, vy);
  viennacl::linalg::prod_impl(A, vx, alpha, vy, beta);
  from_vcl(vy, y);
}

void ref_sell_spmv_f64(uint32_t n, uint32_t C, const uint32_t* rp, const uint32_t* col, const double* val,
                       const double* x, double alpha, double* y, double beta)
{
  viennacl::sliced_ell_matrix<double> A(n, n, C);
  fill_sell(A, n, rp, col, val);
  viennacl::vector<double> vx(n), vy(n);
  to_vcl(x, vx);
  to_vcl(y, vy);
  viennacl::linalg::prod_impl(A, vx, alpha, vy, beta);
  from_vcl(vy, y);
}

// The reference's SELL arrays (sliced_ell_matrix.hpp:140-214), for byte comparison.
void ref_sell_arrays(uint32_t n, uint32_t C, const uint32_t* rp, const uint32_t* col, const double* val,
                     uint32_t* cpb, uint32_t* ci, uint32_t* bs, double* el, uint64_t stored)
{
  viennacl::sliced_ell_matrix<double> A(n, n, C);
  fill_sell(A, n, rp, col, val);
  uint32_t ns = (n + C - 1) / C;
  viennacl::backend::memory_read(A.handle1(), 0, ns * sizeof(uint32_t), cpb);
  viennacl::backend::memory_read(A.handle2(), 0, stored * sizeof(uint32_t), ci);
  viennacl::backend::memory_read(A.handle3(), 0, ns * sizeof(uint32_t), bs);
  viennacl::backend::memory_read(A.handle(), 0, stored * sizeof(double), el);
}

// compressed_matrix::generate_row_block_information (compressed_matrix.hpp:1152-1188)
uint32_t ref_row_blocks(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, uint32_t* out)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  uint32_t nb = (uint32_t)A.blocks1();
  viennacl::backend::memory_read(A.handle3(), 0, (nb + 1) * sizeof(uint32_t), out);
  return nb;
}

void ref_row_info_diag(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, double* out)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> d(n);
  viennacl::linalg::detail::row_info(A, d, viennacl::linalg::detail::SPARSE_ROW_DIAGONAL);
  from_vcl(d, out);
}

double ref_dot_f64(uint32_t n, const double* x, const double* y)
{
  viennacl::vector<double> vx(n), vy(n);
  to_vcl(x, vx);
  to_vcl(y, vy);
  return viennacl::linalg::inner_prod(vx, vy);
}

double ref_norm2_f64(uint32_t n, const double* x)
{
  viennacl::vector<double> vx(n);
  to_vcl(x, vx);
  return viennacl::linalg::norm_2(vx);
}

// linalg::pipelined_cg_vector_update (iterative_operations.hpp:58-87); buf holds 3*256 doubles
void ref_pipelined_cg_vector_update(uint32_t n, double* x, double alpha, double* p, double* r, const double* Ap,
                                    double beta, double* buf)
{
  viennacl::vector<double> vx(n), vp(n), vr(n), vAp(n), vb(768);
  to_vcl(x, vx); to_vcl(p, vp); to_vcl(r, vr); to_vcl(Ap, vAp); to_vcl(buf, vb);
  viennacl::linalg::pipelined_cg_vector_update(vx, alpha, vp, vr, vAp, beta, vb);
  from_vcl(vx, x); from_vcl(vp, p); from_vcl(vr, r); from_vcl(vb, buf);
}

// linalg::pipelined_cg_prod (iterative_operations.hpp:96-122)
void ref_pipelined_cg_prod_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                               const double* p, double* Ap, double* buf)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vp(n), vAp(n), vb(768);
  to_vcl(p, vp); to_vcl(Ap, vAp); to_vcl(buf, vb);
  viennacl::linalg::pipelined_cg_prod(A, vp, vAp, vb);
  from_vcl(vAp, Ap); from_vcl(vb, buf);
}

// linalg::pipelined_bicgstab_* (iterative_operations.hpp:133-236); buf holds 6*256 doubles
void ref_pipelined_bicgstab_prod_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                                     const double* p, double* Ap, const double* r0, double* buf, uint32_t offset)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vp(n), vAp(n), vr0(n), vb(1536);
  to_vcl(p, vp); to_vcl(Ap, vAp); to_vcl(r0, vr0); to_vcl(buf, vb);
  viennacl::linalg::pipelined_bicgstab_prod(A, vp, vAp, vr0, vb, 256, offset);
  from_vcl(vAp, Ap); from_vcl(vb, buf);
}

void ref_pipelined_bicgstab_update_s(uint32_t n, double* s, double* r, const double* Ap, double* buf, uint32_t offset)
{
  viennacl::vector<double> vs(n), vr(n), vAp(n), vb(1536);
  to_vcl(s, vs); to_vcl(r, vr); to_vcl(Ap, vAp); to_vcl(buf, vb);
  viennacl::linalg::pipelined_bicgstab_update_s(vs, vr, vAp, vb, 256, offset);
  from_vcl(vs, s); from_vcl(vb, buf);
}

void ref_pipelined_bicgstab_vector_update(uint32_t n, double* x, double alpha, double* p, double omega,
                                          const double* s, double* r, const double* As, double beta,
                                          const double* Ap, const double* r0, double* buf)
{
  viennacl::vector<double> vx(n), vp(n), vs(n), vr(n), vAs(n), vAp(n), vr0(n), vb(1536);
  to_vcl(x, vx); to_vcl(p, vp); to_vcl(s, vs); to_vcl(r, vr); to_vcl(As, vAs); to_vcl(Ap, vAp);
  to_vcl(r0, vr0); to_vcl(buf, vb);
  viennacl::linalg::pipelined_bicgstab_vector_update(vx, alpha, vp, omega, vs, vr, vAs, beta, vAp, vr0, vb, 256);
  from_vcl(vx, x); from_vcl(vp, p); from_vcl(vr, r); from_vcl(vb, buf);
}

// ---- timing of the reference's own loops (CPU baseline): only the calls into the reference are timed

double ref_bench_csr_spmv(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* x,
                          double* y, uint32_t reps)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vx(n), vy(n);
  to_vcl(x, vx);
  vy = viennacl::linalg::prod(A, vx);  // warm-up
  viennacl::backend::finish();
  double t0 = omp_get_wtime();
  for (uint32_t r = 0; r < reps; ++r)
    vy = viennacl::linalg::prod(A, vx);
  viennacl::backend::finish();
  double t1 = omp_get_wtime();
  from_vcl(vy, y);
  return t1 - t0;
}

// Solve time of cg (kind = 0, jacobi = 0/1), bicgstab (kind = 1) or mixed_precision_cg (kind = 2); iterations
// and error returned.
double ref_bench_solve(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                       double* x, double tol, uint32_t max_iters, int kind, int jacobi, uint32_t* iters, double* error)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vb(n), vx;
  to_vcl(b, vb);
  double t0 = 0, t1 = 0;
  if (kind == 0) {
    viennacl::linalg::cg_tag tag(tol, max_iters);
    if (jacobi) {
      viennacl::linalg::jacobi_precond<viennacl::compressed_matrix<double> > J(A, viennacl::linalg::jacobi_tag());
      t0 = omp_get_wtime();
      vx = viennacl::linalg::solve(A, vb, tag, J);
      t1 = omp_get_wtime();
    } else {
      t0 = omp_get_wtime();
      vx = viennacl::linalg::solve(A, vb, tag);
      t1 = omp_get_wtime();
    }
    *iters = tag.iters();
    *error = tag.error();
  } else if (kind == 2) {
    // mixed_precision_cg.hpp:95-187 with its default inner tolerance: the FP32 copy of A is made
    // inside solve(), as the reference does, so it is part of the timed solve
    viennacl::linalg::mixed_precision_cg_tag tag(tol, max_iters);
    t0 = omp_get_wtime();
    vx = viennacl::linalg::solve(A, vb, tag);
    t1 = omp_get_wtime();
    *iters = tag.iters();
    *error = tag.error();
  } else {
    viennacl::linalg::bicgstab_tag tag(tol, max_iters);
    t0 = omp_get_wtime();
    vx = viennacl::linalg::solve(A, vb, tag);
    t1 = omp_get_wtime();
    *iters = (uint32_t)tag.iters();
    *error = tag.error();
  }
  from_vcl(vx, x);
  return t1 - t0;
}

int ref_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- solvers through linalg::solve (cg.hpp:338-375, bicgstab.hpp:495-533, mixed_precision_cg.hpp:95-193)

void ref_cg_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                double* x, double tol, double abs_tol, uint32_t max_iters, int jacobi,
                uint32_t* iters, double* error)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::cg_tag tag(tol, max_iters);
  tag.abs_tolerance(abs_tol);
  viennacl::vector<double> vx;
  if (jacobi) {
    viennacl::linalg::jacobi_precond<viennacl::compressed_matrix<double> > J(A, viennacl::linalg::jacobi_tag());
    vx = viennacl::linalg::solve(A, vb, tag, J);
  } else {
    vx = viennacl::linalg::solve(A, vb, tag);
  }
  from_vcl(vx, x);
  *iters = tag.iters();
  *error = tag.error();
}

// linalg::pipelined_gmres_* (iterative_operations.hpp:247-418) on main-memory vectors.  The basis
// is one vector of K*ld entries, vector i at offset i*ld (gmres.hpp:194).
void ref_pipelined_gmres_normalize_vk(uint32_t n, double* vk, const double* residual, double* R, uint32_t R_size,
                                      uint32_t offset_in_R, const double* buf, double* r_dot_vk,
                                      uint32_t r_dot_vk_size, uint32_t chunk, uint32_t chunk_offset)
{
  viennacl::vector<double> vv(n), vr(n), vR(R_size), vb(3 * chunk), vd(r_dot_vk_size);
  to_vcl(vk, vv); to_vcl(residual, vr); to_vcl(R, vR); to_vcl(buf, vb); to_vcl(r_dot_vk, vd);
  viennacl::linalg::pipelined_gmres_normalize_vk(vv, vr, vR, offset_in_R, vb, vd, chunk, chunk_offset);
  from_vcl(vv, vk); from_vcl(vR, R); from_vcl(vd, r_dot_vk);
}

void ref_pipelined_gmres_stage1(const double* basis, uint32_t n, uint32_t ld, uint32_t K, uint32_t k,
                                double* vi_in_vk, uint32_t chunk)
{
  viennacl::vector<double> vB(static_cast<size_t>(K) * ld), vo(static_cast<size_t>(K) * chunk);
  to_vcl(basis, vB); to_vcl(vi_in_vk, vo);
  viennacl::linalg::pipelined_gmres_gram_schmidt_stage1(vB, n, ld, k, vo, chunk);
  from_vcl(vo, vi_in_vk);
}

void ref_pipelined_gmres_stage2(double* basis, uint32_t n, uint32_t ld, uint32_t K, uint32_t k,
                                const double* vi_in_vk, double* R, double* buf, uint32_t chunk)
{
  viennacl::vector<double> vB(static_cast<size_t>(K) * ld), vi(static_cast<size_t>(K) * chunk),
      vR(static_cast<size_t>(K) * K), vb(3 * chunk);
  to_vcl(basis, vB); to_vcl(vi_in_vk, vi); to_vcl(R, vR); to_vcl(buf, vb);
  viennacl::linalg::pipelined_gmres_gram_schmidt_stage2(vB, n, ld, k, vi, vR, K, vb, chunk);
  from_vcl(vB, basis); from_vcl(vR, R); from_vcl(vb, buf);
}

void ref_pipelined_gmres_update_result(double* x, const double* residual, const double* basis, uint32_t n,
                                       uint32_t ld, uint32_t K, const double* coef, uint32_t k)
{
  viennacl::vector<double> vx(n), vr(n), vB(static_cast<size_t>(K) * ld), vc(K);
  to_vcl(x, vx); to_vcl(residual, vr); to_vcl(basis, vB); to_vcl(coef, vc);
  viennacl::linalg::pipelined_gmres_update_result(vx, vr, vB, n, ld, vc, k);
  from_vcl(vx, x);
}

// ---- coordinate_matrix / ell_matrix / hyb_matrix, filled by copy() from the STL matrix of the CSR
// arrays.  format: 2 = COO, 3 = ELL, 4 = HYB.

// The arrays copy() writes (coordinate_matrix.hpp:46-99, ell_matrix.hpp:123-171, hyb_matrix.hpp:136-240).
void ref_coo_arrays(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, uint32_t* coords,
                    double* elements)
{
  viennacl::coordinate_matrix<double> A(n, n);
  viennacl::copy(to_stl(n, rp, col, val), A);
  viennacl::backend::memory_read(A.handle12(), 0, 2 * A.nnz() * sizeof(uint32_t), coords);
  viennacl::backend::memory_read(A.handle(), 0, A.nnz() * sizeof(double), elements);
}

void ref_ell_arrays(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, uint32_t* ld,
                    uint32_t* width, uint32_t* coords, double* elements)
{
  viennacl::ell_matrix<double> A;
  viennacl::copy(to_stl(n, rp, col, val), A);
  *ld = (uint32_t)A.internal_size1();
  *width = (uint32_t)A.internal_maxnnz();
  if (!coords) return;
  viennacl::backend::memory_read(A.handle2(), 0, A.internal_nnz() * sizeof(uint32_t), coords);
  viennacl::backend::memory_read(A.handle(), 0, A.internal_nnz() * sizeof(double), elements);
}

void ref_hyb_arrays(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, uint32_t* ld,
                    uint32_t* width, uint32_t* csr_nnz, uint32_t* ell_coords, double* ell_elements,
                    uint32_t* csr_rows, uint32_t* csr_cols, double* csr_elements)
{
  viennacl::hyb_matrix<double> A;
  viennacl::copy(to_stl(n, rp, col, val), A);
  *ld = (uint32_t)A.internal_size1();
  *width = (uint32_t)A.internal_ellnnz();
  *csr_nnz = (uint32_t)A.csr_nnz();
  if (!ell_coords) return;
  const size_t ell = A.internal_size1() * A.internal_ellnnz();
  viennacl::backend::memory_read(A.handle2(), 0, ell * sizeof(uint32_t), ell_coords);
  viennacl::backend::memory_read(A.handle(), 0, ell * sizeof(double), ell_elements);
  viennacl::backend::memory_read(A.handle3(), 0, (n + 1) * sizeof(uint32_t), csr_rows);
  viennacl::backend::memory_read(A.handle4(), 0, A.csr_nnz() * sizeof(uint32_t), csr_cols);
  viennacl::backend::memory_read(A.handle5(), 0, A.csr_nnz() * sizeof(double), csr_elements);
}

// linalg::prod_impl for the format (y = alpha A x + beta y)
void ref_fmt_spmv_f64(int format, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                      const double* x, double alpha, double* y, double beta)
{
  with_format(format, n, rp, col, val, [&](auto& A) {
    viennacl::vector<double> vx(n), vy(n);
    to_vcl(x, vx);
    to_vcl(y, vy);
    viennacl::linalg::prod_impl(A, vx, alpha, vy, beta);
    from_vcl(vy, y);
  });
}

// linalg::pipelined_cg_prod / pipelined_bicgstab_prod for the format (r0 == NULL: CG)
void ref_fmt_pipelined_prod(int format, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                            const double* p, double* Ap, const double* r0, double* buf, uint32_t offset)
{
  with_format(format, n, rp, col, val, [&](auto& A) {
    viennacl::vector<double> vp(n), vAp(n);
    to_vcl(p, vp);
    to_vcl(Ap, vAp);
    if (r0) {
      viennacl::vector<double> vr0(n), vb(1536);
      to_vcl(r0, vr0);
      to_vcl(buf, vb);
      viennacl::linalg::pipelined_bicgstab_prod(A, vp, vAp, vr0, vb, 256, offset);
      from_vcl(vb, buf);
    } else {
      viennacl::vector<double> vb(768);
      to_vcl(buf, vb);
      viennacl::linalg::pipelined_cg_prod(A, vp, vAp, vb);
      from_vcl(vb, buf);
    }
    from_vcl(vAp, Ap);
  });
}

// solve(A, b, cg_tag / bicgstab_tag / gmres_tag) for the format (kind 0 / 1 / 2)
void ref_fmt_solve(int format, int kind, uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val,
                   const double* b, double* x, double tol, uint32_t max_iters, uint32_t krylov_dim, uint32_t* iters,
                   double* error)
{
  with_format(format, n, rp, col, val, [&](auto& A) {
    viennacl::vector<double> vb(n);
    to_vcl(b, vb);
    viennacl::vector<double> vx(n);
    if (kind == 0) {
      viennacl::linalg::cg_tag tag(tol, max_iters);
      vx = viennacl::linalg::solve(A, vb, tag);
      *iters = tag.iters();
      *error = tag.error();
    } else if (kind == 1) {
      viennacl::linalg::bicgstab_tag tag(tol, max_iters);
      vx = viennacl::linalg::solve(A, vb, tag);
      *iters = tag.iters();
      *error = tag.error();
    } else {
      viennacl::linalg::gmres_tag tag(tol, max_iters, krylov_dim);
      vx = viennacl::linalg::solve(A, vb, tag);
      *iters = tag.iters();
      *error = tag.error();
    }
    from_vcl(vx, x);
  });
}

// gmres.hpp:372-378 (compressed_matrix, no preconditioner -> pipelined_solve)
void ref_gmres_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                   double* x, double tol, double abs_tol, uint32_t max_iters, uint32_t krylov_dim, uint32_t* iters,
                   double* error)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::gmres_tag tag(tol, max_iters, krylov_dim);
  tag.abs_tolerance(abs_tol);
  viennacl::vector<double> vx = viennacl::linalg::solve(A, vb, tag);
  from_vcl(vx, x);
  *iters = tag.iters();
  *error = tag.error();
}

void ref_cg_sell(uint32_t n, uint32_t C, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                 double* x, double tol, double abs_tol, uint32_t max_iters, uint32_t* iters, double* error)
{
  viennacl::sliced_ell_matrix<double> A(n, n, C);
  fill_sell(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::cg_tag tag(tol, max_iters);
  tag.abs_tolerance(abs_tol);
  viennacl::vector<double> vx = viennacl::linalg::solve(A, vb, tag);
  from_vcl(vx, x);
  *iters = tag.iters();
  *error = tag.error();
}

void ref_bicgstab_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                      double* x, double tol, double abs_tol, uint32_t max_iters, uint32_t max_before_restart,
                      int jacobi, uint32_t* iters, double* error)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::bicgstab_tag tag(tol, max_iters, max_before_restart);
  tag.abs_tolerance(abs_tol);
  viennacl::vector<double> vx;
  if (jacobi) {
    viennacl::linalg::jacobi_precond<viennacl::compressed_matrix<double> > J(A, viennacl::linalg::jacobi_tag());
    vx = viennacl::linalg::solve(A, vb, tag, J);
  } else {
    vx = viennacl::linalg::solve(A, vb, tag);
  }
  from_vcl(vx, x);
  *iters = (uint32_t)tag.iters();
  *error = tag.error();
}

void ref_bicgstab_sell(uint32_t n, uint32_t C, const uint32_t* rp, const uint32_t* col, const double* val,
                       const double* b, double* x, double tol, double abs_tol, uint32_t max_iters,
                       uint32_t* iters, double* error)
{
  viennacl::sliced_ell_matrix<double> A(n, n, C);
  fill_sell(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::bicgstab_tag tag(tol, max_iters);
  tag.abs_tolerance(abs_tol);
  viennacl::vector<double> vx = viennacl::linalg::solve(A, vb, tag);
  from_vcl(vx, x);
  *iters = (uint32_t)tag.iters();
  *error = tag.error();
}

void ref_mixed_cg_csr(uint32_t n, const uint32_t* rp, const uint32_t* col, const double* val, const double* b,
                      double* x, double tol, uint32_t max_iters, float inner_tol, uint32_t* iters, double* error)
{
  viennacl::compressed_matrix<double> A;
  fill_csr(A, n, rp, col, val);
  viennacl::vector<double> vb(n);
  to_vcl(b, vb);
  viennacl::linalg::mixed_precision_cg_tag tag(tol, max_iters, inner_tol);
  viennacl::vector<double> vx = viennacl::linalg::solve(A, vb, tag);
  from_vcl(vx, x);
  *iters = tag.iters();
  *error = tag.error();
}

// io/matrix_market.hpp:102-326 through its std::vector<std::map> overload (:354-361), the type the
// reference's own tests read files into.  Returns the reader's own result (0 on a parse error, 1 when
// the file cannot be opened, else the number of lines read); the matrix is kept until the next call
// and copied out by ref_mm_csr.
long ref_mm_read(const char* path, long index_base, uint32_t* n_rows, uint32_t* nnz)
{
  mm_last.clear();
  const long r = viennacl::io::read_matrix_market_file(mm_last, path, index_base);
  std::size_t k = 0;
  for (std::size_t i = 0; i < mm_last.size(); ++i) k += mm_last[i].size();
  *n_rows = static_cast<uint32_t>(mm_last.size());
  *nnz = static_cast<uint32_t>(k);
  return r;
}

void ref_mm_csr(uint32_t* rp, uint32_t* col, double* val)
{
  uint32_t k = 0;
  rp[0] = 0;
  for (std::size_t i = 0; i < mm_last.size(); ++i) {
    for (std::map<unsigned int, double>::const_iterator it = mm_last[i].begin(); it != mm_last[i].end(); ++it) {
      col[k] = it->first;
      val[k] = it->second;
      ++k;
    }
    rp[i + 1] = k;
  }
}

}  // extern "C"
