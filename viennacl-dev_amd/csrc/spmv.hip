// spmv.hip -- SpMV and row_info entry points of libvclhip.
#include <cstdlib>

#include "internal.hpp"
#include "spmv_kernels.hpp"

namespace vclhip {

static uint32_t env_u32(const char* name, uint32_t dflt, uint32_t lo, uint32_t hi)
{
  const char* s = std::getenv(name);
  if (!s || !*s) return dflt;
  unsigned long v = std::strtoul(s, nullptr, 10);
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return static_cast<uint32_t>(v);
}

bool stream_loads(size_t bytes)
{
  static const uint32_t mode = env_u32("VCLHIP_STREAM_LOADS", 2, 0, 2);
  if (mode < 2) return mode == 1;
  return bytes > (size_t(256) << 20);
}

bool dia_stream_loads(size_t working_set)
{
  static const uint32_t mode = env_u32("VCLHIP_DIA_NT", 2, 0, 2);
  if (mode < 2) return mode == 1;
  return working_set > (size_t(256) << 20);
}

bool dia_codes_enabled()
{
  static const uint32_t on = env_u32("VCLHIP_DIA_CODED", 1, 0, 1);
  return on == 1;
}

uint32_t csr_gmax()
{
  static const uint32_t g = env_u32("VCLHIP_CSR_GMAX", 2048, 8, kMaxPartials);
  return g;
}

uint32_t sell_gmax()
{
  static const uint32_t g = env_u32("VCLHIP_SELL_GMAX", 2048, 8, kMaxPartials);
  return g;
}

// Plain SpMV writes no partials, so by default every row block / row tile gets its own workgroup.
uint32_t spmv_gmax()
{
  static const uint32_t g = env_u32("VCLHIP_SPMV_GMAX", 1u << 24, 8, 1u << 24);
  return g;
}

uint32_t vec_gmax()
{
  static const uint32_t g = env_u32("VCLHIP_VEC_GMAX", 2048, 8, kMaxPartials);
  return g;
}

template <typename T>
static int check_csr(const vclhip_csr* A)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!A->row_ptr || !A->blocks || (A->nnz && (!A->col_idx || !A->values)))
    return invalid("CSR arrays must be non-NULL");
  if (A->n_rows >= kPlanRowLimit) return invalid("n_rows must be below 2^28");
  if (A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("num_blocks out of range");
  return VCLHIP_SUCCESS;
}

template <typename T>
static CsrDev<T> csr_dev(const vclhip_csr* A)
{
  CsrDev<T> d;
  d.row_ptr = A->row_ptr;
  d.col = A->col_idx;
  d.val = static_cast<const T*>(A->values);
  d.blocks = A->blocks;
  d.num_blocks = A->num_blocks;
  d.n_rows = A->n_rows;
  return d;
}

template <typename T>
static int check_sell(const vclhip_sell* A)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!A->columns_per_block || !A->column_indices || !A->block_start || !A->elements)
    return invalid("SELL arrays must be non-NULL");
  if (A->C == 0) return invalid("C must be positive");
  if (A->num_slices != (A->n_rows + A->C - 1) / A->C) return invalid("num_slices != ceil(n_rows / C)");
  return VCLHIP_SUCCESS;
}

template <typename T>
static SellDev<T> sell_dev(const vclhip_sell* A)
{
  SellDev<T> d;
  d.cpb = A->columns_per_block;
  d.col = A->column_indices;
  d.start = A->block_start;
  d.val = static_cast<const T*>(A->elements);
  d.n_rows = A->n_rows;
  d.C = A->C;
  return d;
}

template <typename T>
static int csr_spmv(const vclhip_csr* A, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys,
                    uint32_t yi, T beta, void* stream)
{
  VCLHIP_TRY(check_csr<T>(A));
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!x || !y || xi == 0 || yi == 0) return invalid("x/y must be non-NULL with nonzero stride");
  if (x == y) return invalid("x and y must not alias");
  const uint32_t grid = chunk_grid(A->num_blocks, spmv_gmax());
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool nt = stream_loads(static_cast<size_t>(A->nnz) * (sizeof(T) + 4));
  T* const np = nullptr;
  const uint32_t* const ns = nullptr;
  if (xs == 0 && xi == 1 && ys == 0 && yi == 1) {
    EpiAxpby<T, true> epi{y, 0, 1, alpha, beta};
    if (nt)
      launch_timed(csr_kernel<T, true, EpiAxpby<T, true>, true>, dim3(grid), dim3(kBlock), s, csr_dev<T>(A), x, 0u,
                   1u, epi, np, 0u, ns, 0u);
    else
      launch_timed(csr_kernel<T, true, EpiAxpby<T, true>, false>, dim3(grid), dim3(kBlock), s, csr_dev<T>(A), x, 0u,
                   1u, epi, np, 0u, ns, 0u);
  } else {
    EpiAxpby<T, false> epi{y, ys, yi, alpha, beta};
    launch_timed(csr_kernel<T, false, EpiAxpby<T, false>, false>, dim3(grid), dim3(kBlock), s, csr_dev<T>(A), x, xs,
                 xi, epi, np, 0u, ns, 0u);
  }
  return check_launch("csr_kernel");
}

template <typename T>
static int sell_spmv(const vclhip_sell* A, const T* x, uint32_t xs, uint32_t xi, T alpha, T* y, uint32_t ys,
                     uint32_t yi, T beta, void* stream)
{
  VCLHIP_TRY(check_sell<T>(A));
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!x || !y || xi == 0 || yi == 0) return invalid("x/y must be non-NULL with nonzero stride");
  if (x == y) return invalid("x and y must not alias");
  const uint32_t tiles = (A->n_rows + kBlock - 1) / kBlock;
  const uint32_t grid = chunk_grid(tiles, spmv_gmax());
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the stored size is not in the descriptor: nontemporal loads only when asked for
  const bool nt = stream_loads(0);
  T* const np = nullptr;
  const uint32_t* const ns = nullptr;
  if (xs == 0 && xi == 1 && ys == 0 && yi == 1) {
    EpiAxpby<T, true> epi{y, 0, 1, alpha, beta};
    if (nt)
      launch_timed(sell_kernel<T, true, EpiAxpby<T, true>, false, true>, dim3(grid), dim3(kBlock), s, sell_dev<T>(A),
                   x, 0u, 1u, epi, np, 0u, ns, 0u);
    else
      launch_timed(sell_kernel<T, true, EpiAxpby<T, true>, false, false>, dim3(grid), dim3(kBlock), s,
                   sell_dev<T>(A), x, 0u, 1u, epi, np, 0u, ns, 0u);
  } else {
    EpiAxpby<T, false> epi{y, ys, yi, alpha, beta};
    launch_timed(sell_kernel<T, false, EpiAxpby<T, false>, false, false>, dim3(grid), dim3(kBlock), s,
                 sell_dev<T>(A), x, xs, xi, epi, np, 0u, ns, 0u);
  }
  return check_launch("sell_kernel");
}

// host_based/sparse_matrix_operations.hpp:52-98, one lane per row.
template <typename T>
__global__ __launch_bounds__(kBlock) void row_info_kernel(const uint32_t* __restrict__ rp,
                                                          const uint32_t* __restrict__ col,
                                                          const T* __restrict__ val, uint32_t n, T* __restrict__ out,
                                                          int info)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  T value = 0;
  const uint32_t end = rp[row + 1];
  switch (info) {
    case 0:
      for (uint32_t i = rp[row]; i < end; ++i) value = fmax(value, fabs(val[i]));
      break;
    case 1:
      for (uint32_t i = rp[row]; i < end; ++i) value += fabs(val[i]);
      break;
    case 2:
      for (uint32_t i = rp[row]; i < end; ++i) value += val[i] * val[i];
      value = sqrt(value);
      break;
    default:
      for (uint32_t i = rp[row]; i < end; ++i)
        if (col[i] == row) { value = val[i]; break; }
      break;
  }
  out[row] = value;
}

template <typename T>
static int row_info(const vclhip_csr* A, T* out, int info, void* stream)
{
  if (!A) return invalid("matrix is NULL");
  if (A->n_rows == 0) return VCLHIP_SUCCESS;
  if (!A->row_ptr || !out || (A->nnz && (!A->col_idx || !A->values))) return invalid("null array");
  if (info < 0 || info > 3) return invalid("info must be 0..3");
  const uint32_t grid = (A->n_rows + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(row_info_kernel<T>, dim3(grid), dim3(kBlock), 0, static_cast<hipStream_t>(stream), A->row_ptr,
                     A->col_idx, static_cast<const T*>(A->values), A->n_rows, out, info);
  return check_launch("row_info_kernel");
}

// Gradient of a linear solve with respect to the matrix values, on the pattern of A:
// dA[k] = -lambda[row(k)] * x[col(k)].  Walks the row-block plan; a workgroup stages the block's
// row pointers in LDS and every lane finds the row of its entry by bisection, so the col reads and
// dA writes are coalesced whatever the row lengths.
template <typename T>
__global__ __launch_bounds__(kBlock) void csr_adjoint_grad_kernel(CsrDev<T> A, const T* __restrict__ lambda,
                                                                  const T* __restrict__ x, T* __restrict__ dA)
{
  __shared__ uint32_t srp[kCsrWG + 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t lw = xcd_logical_block();
  uint32_t b0, b1;
  chunk_range(A.num_blocks, lw, gridDim.x, b0, b1);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t d0 = A.blocks[b];
    const uint32_t r0 = d0 & (kPlanRowLimit - 1);
    const uint32_t mode = d0 >> 28;
    const uint32_t r1 = A.blocks[b + 1] & (kPlanRowLimit - 1);
    const uint32_t e0 = A.row_ptr[r0];
    const uint32_t e1 = A.row_ptr[r1];
    if (mode == kModeLongRow) {
      const T lam = lambda[r0];
      for (uint32_t e = e0 + tid; e < e1; e += kBlock) dA[e] = -lam * x[A.col[e]];
      continue;
    }
    const uint32_t rows = min(r1 - r0, static_cast<uint32_t>(kCsrWG));
    for (uint32_t i = tid; i <= rows; i += kBlock) srp[i] = A.row_ptr[r0 + i];
    __syncthreads();
    for (uint32_t e = e0 + tid; e < e1; e += kBlock) {
      uint32_t lo = 0, hi = rows;  // srp[lo] <= e < srp[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (srp[mid] <= e)
          lo = mid;
        else
          hi = mid;
      }
      dA[e] = -lambda[r0 + lo] * x[A.col[e]];
    }
    __syncthreads();
  }
}

template <typename T>
static int csr_adjoint_grad(const vclhip_csr* A, const T* lambda, const T* x, T* dA, void* stream)
{
  VCLHIP_TRY(check_csr<T>(A));
  if (A->n_rows == 0 || A->nnz == 0) return VCLHIP_SUCCESS;
  if (!lambda || !x || !dA) return invalid("null vector");
  const uint32_t grid = chunk_grid(A->num_blocks, spmv_gmax());
  hipLaunchKernelGGL(csr_adjoint_grad_kernel<T>, dim3(grid), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                     csr_dev<T>(A), lambda, x, dA);
  return check_launch("csr_adjoint_grad_kernel");
}

}  // namespace vclhip

extern "C" {

int vclhip_csr_spmv_f64(const vclhip_csr* A, const double* x, uint32_t xs, uint32_t xi, double alpha, double* y,
                        uint32_t ys, uint32_t yi, double beta, void* stream)
{
  return vclhip::csr_spmv<double>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_csr_spmv_f32(const vclhip_csr* A, const float* x, uint32_t xs, uint32_t xi, float alpha, float* y,
                        uint32_t ys, uint32_t yi, float beta, void* stream)
{
  return vclhip::csr_spmv<float>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_sell_spmv_f64(const vclhip_sell* A, const double* x, uint32_t xs, uint32_t xi, double alpha, double* y,
                         uint32_t ys, uint32_t yi, double beta, void* stream)
{
  return vclhip::sell_spmv<double>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_sell_spmv_f32(const vclhip_sell* A, const float* x, uint32_t xs, uint32_t xi, float alpha, float* y,
                         uint32_t ys, uint32_t yi, float beta, void* stream)
{
  return vclhip::sell_spmv<float>(A, x, xs, xi, alpha, y, ys, yi, beta, stream);
}

int vclhip_csr_row_info_f64(const vclhip_csr* A, double* out, int info, void* stream)
{
  return vclhip::row_info<double>(A, out, info, stream);
}

int vclhip_csr_row_info_f32(const vclhip_csr* A, float* out, int info, void* stream)
{
  return vclhip::row_info<float>(A, out, info, stream);
}

int vclhip_csr_adjoint_grad_f64(const vclhip_csr* A, const double* lambda, const double* x, double* dA, void* stream)
{
  return vclhip::csr_adjoint_grad<double>(A, lambda, x, dA, stream);
}

int vclhip_csr_adjoint_grad_f32(const vclhip_csr* A, const float* lambda, const float* x, float* dA, void* stream)
{
  return vclhip::csr_adjoint_grad<float>(A, lambda, x, dA, stream);
}

}  // extern "C"
