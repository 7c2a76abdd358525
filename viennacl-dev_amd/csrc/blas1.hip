// blas1.hip -- vector entry points of libvclhip: inner products, norms, axpby (avbv), fill,
// element-wise division and precision conversion.  Counterparts of
// host_based/vector_operations.hpp:67-833 and cuda/vector_operations.hpp (see vclhip.h).
#include <cmath>

#include "internal.hpp"
#include "spmv_kernels.hpp"

namespace vclhip {

template <typename T>
__device__ __forceinline__ T at(const T* __restrict__ v, uint32_t i, uint32_t s, uint32_t inc)
{
  return v[static_cast<size_t>(i) * inc + s];
}

// Per-workgroup partial of sum_i x_i*y_i (accumulated in A).
template <typename T, typename A>
__global__ __launch_bounds__(kBlock) void dot_partial_kernel(uint32_t n, const T* __restrict__ x, uint32_t xs,
                                                             uint32_t xi, const T* __restrict__ y, uint32_t ys,
                                                             uint32_t yi, A* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  A acc[1] = {0};
  if (xi == 1 && yi == 1) {
    const T* xp = x + xs;
    const T* yp = y + ys;
#pragma unroll 4
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) acc[0] += static_cast<A>(xp[i]) * static_cast<A>(yp[i]);
  } else {
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock)
      acc[0] += static_cast<A>(at(x, i, xs, xi)) * static_cast<A>(at(y, i, ys, yi));
  }
  block_store_partials<1, A>(acc, part, 0, lw);
}

template <typename A>
__global__ __launch_bounds__(kBlock) void sum_partials_kernel(const A* __restrict__ part, uint32_t count,
                                                              A* __restrict__ out)
{
  A s = block_sum_partials(part, count);
  if (threadIdx.x == 0) *out = s;
}

// Contiguous sum of x*y with W-wide loads (x, y aligned for the packs).
template <typename T, typename A, int W>
__global__ __launch_bounds__(kBlock) void dot_partial_vec_kernel(uint32_t n, const T* __restrict__ x,
                                                                 const T* __restrict__ y, A* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  A acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> vy = ldp<V>(y, i);
#pragma unroll
    for (int k = 0; k < V; ++k) acc[0] += static_cast<A>(vx.v[k]) * static_cast<A>(vy.v[k]);
  });
  block_store_partials<1, A>(acc, part, 0, lw);
}

template <typename T, int W>
__global__ __launch_bounds__(kBlock) void axpby_kernel(uint32_t n, T a, const T* x, T b, const T* y, T* z)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> vy = ldp<V>(y, i);
    Pack<V, T> vz;
#pragma unroll
    for (int k = 0; k < V; ++k) vz.v[k] = vx.v[k] * a + vy.v[k] * b;
    stp<V>(z, i, vz);
  });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void avbv_kernel(uint32_t n, T* __restrict__ v1, uint32_t s1, uint32_t i1,
                                                      const T* __restrict__ v2, uint32_t s2, uint32_t i2, T alpha,
                                                      const T* __restrict__ v3, uint32_t s3, uint32_t i3, T beta)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock)
    v1[static_cast<size_t>(i) * i1 + s1] = at(v2, i, s2, i2) * alpha + at(v3, i, s3, i3) * beta;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fill_kernel(uint32_t n, T* __restrict__ v, uint32_t s, uint32_t inc, T a)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) v[static_cast<size_t>(i) * inc + s] = a;
}

template <typename T, int W>
__global__ __launch_bounds__(kBlock) void ediv_kernel(uint32_t n, const T* x, const T* y, T* z)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> vy = ldp<V>(y, i);
    Pack<V, T> vz;
#pragma unroll
    for (int k = 0; k < V; ++k) vz.v[k] = vx.v[k] / vy.v[k];
    stp<V>(z, i, vz);
  });
}

// W source elements per step; S and D widths follow (double <-> float: 2 per 16-byte double pack).
template <typename S, typename D, int W>
__global__ __launch_bounds__(kBlock) void convert_kernel(uint32_t n, const S* __restrict__ src, D* __restrict__ dst)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, S> vs = ldp<V>(src, i);
    Pack<V, D> vd;
#pragma unroll
    for (int k = 0; k < V; ++k) vd.v[k] = static_cast<D>(vs.v[k]);
    stp<V>(dst, i, vd);
  });
}

static uint32_t vec_grid(uint32_t n) { return chunk_grid((n + 4 * kBlock - 1) / (4 * kBlock), vec_gmax()); }

// Sum of x*y over n entries into *result_host, through a stream-ordered workspace.
template <typename T, typename A>
static int dot_to_host(uint32_t n, const T* x, uint32_t xs, uint32_t xi, const T* y, uint32_t ys, uint32_t yi,
                       A* result_host, void* stream)
{
  if (!result_host) return invalid("result is NULL");
  *result_host = 0;
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || xi == 0 || yi == 0) return invalid("null vector or zero stride");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t grid = vec_grid(n);
  void* ws = nullptr;
  VCLHIP_HIP(hipMallocAsync(&ws, (grid + 1) * sizeof(A), s));
  A* part = static_cast<A*>(ws);
  if (xi == 1 && yi == 1)
    VCLHIP_LAUNCH_W(aligned16({x + xs, y + ys}), (dot_partial_vec_kernel<T, A, kVecW<T> >),
                    (dot_partial_vec_kernel<T, A, 1>), grid, s, n, x + xs, y + ys, part);
  else
    hipLaunchKernelGGL((dot_partial_kernel<T, A>), dim3(grid), dim3(kBlock), 0, s, n, x, xs, xi, y, ys, yi, part);
  int st = check_launch("dot_partial_kernel");
  if (st == 0) {
    hipLaunchKernelGGL(sum_partials_kernel<A>, dim3(1), dim3(kBlock), 0, s, part, grid, part + grid);
    st = check_launch("sum_partials_kernel");
  }
  if (st == 0) st = hip_status(hipMemcpyAsync(result_host, part + grid, sizeof(A), hipMemcpyDeviceToHost, s), "dot D2H");
  int st2 = hip_status(hipFreeAsync(ws, s), "hipFreeAsync");
  int st3 = hip_status(hipStreamSynchronize(s), "dot sync");
  return st ? st : (st2 ? st2 : st3);
}

}  // namespace vclhip

using namespace vclhip;

extern "C" {

int vclhip_dot_f64(uint32_t n, const double* x, uint32_t xs, uint32_t xi, const double* y, uint32_t ys, uint32_t yi,
                   double* result, void* stream)
{
  return dot_to_host<double, double>(n, x, xs, xi, y, ys, yi, result, stream);
}

int vclhip_dot_f32(uint32_t n, const float* x, uint32_t xs, uint32_t xi, const float* y, uint32_t ys, uint32_t yi,
                   float* result, void* stream)
{
  return dot_to_host<float, float>(n, x, xs, xi, y, ys, yi, result, stream);
}

int vclhip_dot_f32_acc64(uint32_t n, const float* x, const float* y, double* result, void* stream)
{
  return dot_to_host<float, double>(n, x, 0, 1, y, 0, 1, result, stream);
}

int vclhip_norm2_f64(uint32_t n, const double* x, uint32_t xs, uint32_t xi, double* result, void* stream)
{
  VCLHIP_TRY((dot_to_host<double, double>(n, x, xs, xi, x, xs, xi, result, stream)));
  *result = std::sqrt(*result);
  return VCLHIP_SUCCESS;
}

int vclhip_norm2_f32(uint32_t n, const float* x, uint32_t xs, uint32_t xi, float* result, void* stream)
{
  VCLHIP_TRY((dot_to_host<float, float>(n, x, xs, xi, x, xs, xi, result, stream)));
  *result = std::sqrt(*result);
  return VCLHIP_SUCCESS;
}

int vclhip_axpby_f64(uint32_t n, double a, const double* x, double b, const double* y, double* z, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || !z) return invalid("null vector");
  if (aligned16({x, y, z}))
    launch_timed(axpby_kernel<double, 2>, dim3(vec_grid(n)), dim3(kBlock), static_cast<hipStream_t>(stream), n, a, x,
                 b, y, z);
  else
    launch_timed(axpby_kernel<double, 1>, dim3(vec_grid(n)), dim3(kBlock), static_cast<hipStream_t>(stream), n, a, x,
                 b, y, z);
  return check_launch("axpby_kernel");
}

int vclhip_axpby_f32(uint32_t n, float a, const float* x, float b, const float* y, float* z, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || !z) return invalid("null vector");
  VCLHIP_LAUNCH_W(aligned16({x, y, z}), (axpby_kernel<float, 4>), (axpby_kernel<float, 1>), vec_grid(n),
                  static_cast<hipStream_t>(stream), n, a, x, b, y, z);
  return check_launch("axpby_kernel");
}

int vclhip_avbv_f64(uint32_t n, double* v1, uint32_t s1, uint32_t i1, const double* v2, uint32_t s2, uint32_t i2,
                    double alpha, const double* v3, uint32_t s3, uint32_t i3, double beta, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!v1 || !v2 || !v3 || !i1 || !i2 || !i3) return invalid("null vector or zero stride");
  hipLaunchKernelGGL(avbv_kernel<double>, dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, v1,
                     s1, i1, v2, s2, i2, alpha, v3, s3, i3, beta);
  return check_launch("avbv_kernel");
}

int vclhip_fill_f64(uint32_t n, double* v, uint32_t s, uint32_t inc, double a, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!v || !inc) return invalid("null vector or zero stride");
  hipLaunchKernelGGL(fill_kernel<double>, dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, v, s,
                     inc, a);
  return check_launch("fill_kernel");
}

int vclhip_fill_f32(uint32_t n, float* v, uint32_t s, uint32_t inc, float a, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!v || !inc) return invalid("null vector or zero stride");
  hipLaunchKernelGGL(fill_kernel<float>, dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, v, s,
                     inc, a);
  return check_launch("fill_kernel");
}

int vclhip_element_div_f64(uint32_t n, const double* x, const double* y, double* z, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || !z) return invalid("null vector");
  VCLHIP_LAUNCH_W(aligned16({x, y, z}), (ediv_kernel<double, 2>), (ediv_kernel<double, 1>), vec_grid(n),
                  static_cast<hipStream_t>(stream), n, x, y, z);
  return check_launch("ediv_kernel");
}

int vclhip_element_div_f32(uint32_t n, const float* x, const float* y, float* z, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || !z) return invalid("null vector");
  VCLHIP_LAUNCH_W(aligned16({x, y, z}), (ediv_kernel<float, 4>), (ediv_kernel<float, 1>), vec_grid(n),
                  static_cast<hipStream_t>(stream), n, x, y, z);
  return check_launch("ediv_kernel");
}

int vclhip_convert_f64_to_f32(uint32_t n, const double* src, float* dst, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!src || !dst) return invalid("null vector");
  VCLHIP_LAUNCH_W(aligned16({src, dst}), (convert_kernel<double, float, 2>), (convert_kernel<double, float, 1>),
                  vec_grid(n), static_cast<hipStream_t>(stream), n, src, dst);
  return check_launch("convert_kernel");
}

int vclhip_convert_f32_to_f64(uint32_t n, const float* src, double* dst, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!src || !dst) return invalid("null vector");
  VCLHIP_LAUNCH_W(aligned16({src, dst}), (convert_kernel<float, double, 4>), (convert_kernel<float, double, 1>),
                  vec_grid(n), static_cast<hipStream_t>(stream), n, src, dst);
  return check_launch("convert_kernel");
}

}  // extern "C"
