// formats.hip -- SpMV for coordinate_matrix, ell_matrix and hyb_matrix.
//
// COO runs the CSR kernel over the row pointers of its row order, reading the column of entry e at
// coords[2e + 1] (the kernel's column stride 2); its row sums are the reference's sequential
// per-entry accumulation when alpha = 1 and beta = 0, and differ from it only in rounding
// otherwise (the reference adds alpha * value * x entry by entry,
// host_based/sparse_matrix_operations.hpp:1233-1245).  ELL runs one lane per row over the
// column-major ELL part, so that a wavefront reads 64 consecutive entries of each column.  HYB with
// a row-block plan of its CSR part runs the CSR kernel over that part, each row starting from its
// ELL sum (one lane per row again); without a plan, one lane walks the ELL part and then the CSR
// part of its row.
#include "internal.hpp"
#include "spmv_kernels.hpp"

namespace vclhip {
namespace formats_detail {

template <typename T, int CS, class Dev, bool Hyb = false>
int run_csr_like(const Dev& d, uint32_t num_blocks, size_t matrix_bytes, const T* x, uint32_t xs, uint32_t xi,
                 T alpha, T* y, uint32_t ys, uint32_t yi, T beta, void* stream, const char* name)
{
  const uint32_t grid = chunk_grid(num_blocks, spmv_gmax());
  hipStream_t s = static_cast<hipStream_t>(stream);
  T* const np = nullptr;
  const uint32_t* const ns = nullptr;
  if (xs == 0 && xi == 1 && ys == 0 && yi == 1) {
    EpiAxpby<T, true> epi{y, 0, 1, alpha, beta};
    if (stream_loads(matrix_bytes))
      launch_timed(csr_kernel<T, true, EpiAxpby<T, true>, true, CS, Hyb>, dim3(grid), dim3(kBlock), s, d, x, 0u, 1u, epi,
                   np, 0u, ns, 0u);
    else
      launch_timed(csr_kernel<T, true, EpiAxpby<T, true>, false, CS, Hyb>, dim3(grid), dim3(kBlock), s, d, x, 0u, 1u, epi,
                   np, 0u, ns, 0u);
  } else {
    EpiAxpby<T, false> epi{y, ys, yi, alpha, beta};
    launch_timed(csr_kernel<T, false, EpiAxpby<T, false>, false, CS, Hyb>, dim3(grid), dim3(kBlock), s, d, x, xs, xi, epi,
                 np, 0u, ns, 0u);
  }
  return check_launch(name);
}

int check_vectors(const void* x, uint32_t xi, const void* y, uint32_t yi)
{
  if (!x || !y || xi == 0 || yi == 0) return invalid("x/y must be non-NULL with nonzero stride");
  if (x == y) return invalid("x and y must not alias");
  return VCLHIP_SUCCESS;
}

template <typename T>
int coo_spmv(const vclhip_coo* A, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys, uint32_t yi,
             T beta, void* stream)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!A->row_ptr || !A->blocks || (A->nnz && (!A->coords || !A->elements)))
    return invalid("COO arrays must be non-NULL");
  if (A->n_rows >= kPlanRowLimit) return invalid("n_rows must be below 2^28");
  if (A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("num_blocks out of range");
  VCLHIP_TRY(check_vectors(x, xi, y, yi));
  CsrDev<T> d;
  d.row_ptr = A->row_ptr;
  d.col = A->coords ? A->coords + 1 : nullptr;
  d.val = static_cast<const T*>(A->elements);
  d.blocks = A->blocks;
  d.num_blocks = A->num_blocks;
  d.n_rows = A->n_rows;
  return run_csr_like<T, 2>(d, A->num_blocks, static_cast<size_t>(A->nnz) * (sizeof(T) + 8), x, xs, xi, alpha, y, ys,
                            yi, beta, stream, "coo spmv");
}

template <typename T, bool Hyb>
int ell_like_spmv(const EllDev<T>& d, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys, uint32_t yi,
                  T beta, void* stream)
{
  VCLHIP_TRY(check_vectors(x, xi, y, yi));
  const uint32_t grid = chunk_grid((d.n_rows + kBlock - 1) / kBlock, spmv_gmax());
  hipStream_t s = static_cast<hipStream_t>(stream);
  T* const np = nullptr;
  const uint32_t* const ns = nullptr;
  if (xs == 0 && xi == 1 && ys == 0 && yi == 1) {
    EpiAxpby<T, true> epi{y, 0, 1, alpha, beta};
    launch_timed(ell_kernel<T, true, EpiAxpby<T, true>, Hyb>, dim3(grid), dim3(kBlock), s, d, x, 0u, 1u, epi, np, 0u,
                 ns, 0u);
  } else {
    EpiAxpby<T, false> epi{y, ys, yi, alpha, beta};
    launch_timed(ell_kernel<T, false, EpiAxpby<T, false>, Hyb>, dim3(grid), dim3(kBlock), s, d, x, xs, xi, epi, np,
                 0u, ns, 0u);
  }
  return check_launch(Hyb ? "hyb spmv" : "ell spmv");
}

template <typename T>
int ell_spmv(const vclhip_ell* A, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys, uint32_t yi,
             T beta, void* stream)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (A->internal_size1 < A->n_rows) return invalid("internal_size1 below n_rows");
  if (A->internal_maxnnz && (!A->coords || !A->elements)) return invalid("ELL arrays must be non-NULL");
  EllDev<T> d;
  d.coords = A->coords;
  d.val = static_cast<const T*>(A->elements);
  d.n_rows = A->n_rows;
  d.ld = A->internal_size1;
  d.width = A->internal_maxnnz;
  return ell_like_spmv<T, false>(d, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

template <typename T>
int hyb_spmv(const vclhip_hyb* A, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys, uint32_t yi,
             T beta, void* stream)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (A->internal_size1 < A->n_rows) return invalid("internal_size1 below n_rows");
  if ((A->internal_ellnnz && (!A->ell_coords || !A->ell_elements)) || !A->csr_rows || !A->csr_cols ||
      !A->csr_elements)
    return invalid("HYB arrays must be non-NULL");
  if (A->blocks) {
    if (A->n_rows >= kPlanRowLimit) return invalid("n_rows must be below 2^28");
    if (A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("num_blocks out of range");
    VCLHIP_TRY(check_vectors(x, xi, y, yi));
    HybDev<T> h;
    h.row_ptr = A->csr_rows;
    h.col = A->csr_cols;
    h.val = static_cast<const T*>(A->csr_elements);
    h.blocks = A->blocks;
    h.num_blocks = A->num_blocks;
    h.n_rows = A->n_rows;
    h.ell_coords = A->ell_coords;
    h.ell_val = static_cast<const T*>(A->ell_elements);
    h.ell_ld = A->internal_size1;
    h.ell_width = A->internal_ellnnz;
    const size_t ell_bytes = static_cast<size_t>(A->internal_size1) * A->internal_ellnnz * (sizeof(T) + 4);
    return run_csr_like<T, 1, HybDev<T>, true>(h, A->num_blocks, ell_bytes, x, xs, xi, alpha, y, ys, yi, beta,
                                               stream, "hyb spmv");
  }
  EllDev<T> d;
  d.coords = A->ell_coords;
  d.val = static_cast<const T*>(A->ell_elements);
  d.n_rows = A->n_rows;
  d.ld = A->internal_size1;
  d.width = A->internal_ellnnz;
  d.csr_rows = A->csr_rows;
  d.csr_cols = A->csr_cols;
  d.csr_val = static_cast<const T*>(A->csr_elements);
  return ell_like_spmv<T, true>(d, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

}  // namespace formats_detail
}  // namespace vclhip

using namespace vclhip::formats_detail;

extern "C" {

int vclhip_coo_spmv_f64(const vclhip_coo* A, const double* x, uint32_t xs, uint32_t xi, double alpha, double* y,
                        uint32_t ys, uint32_t yi, double beta, void* stream)
{
  return coo_spmv<double>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_coo_spmv_f32(const vclhip_coo* A, const float* x, uint32_t xs, uint32_t xi, float alpha, float* y,
                        uint32_t ys, uint32_t yi, float beta, void* stream)
{
  return coo_spmv<float>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_ell_spmv_f64(const vclhip_ell* A, const double* x, uint32_t xs, uint32_t xi, double alpha, double* y,
                        uint32_t ys, uint32_t yi, double beta, void* stream)
{
  return ell_spmv<double>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_ell_spmv_f32(const vclhip_ell* A, const float* x, uint32_t xs, uint32_t xi, float alpha, float* y,
                        uint32_t ys, uint32_t yi, float beta, void* stream)
{
  return ell_spmv<float>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_hyb_spmv_f64(const vclhip_hyb* A, const double* x, uint32_t xs, uint32_t xi, double alpha, double* y,
                        uint32_t ys, uint32_t yi, double beta, void* stream)
{
  return hyb_spmv<double>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_hyb_spmv_f32(const vclhip_hyb* A, const float* x, uint32_t xs, uint32_t xi, float alpha, float* y,
                        uint32_t ys, uint32_t yi, float beta, void* stream)
{
  return hyb_spmv<float>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

}  // extern "C"
