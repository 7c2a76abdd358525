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

// Read-only streaming reference: workgroup lw sums the lw-th contiguous chunk of x with 16-byte
// loads, kStreamU of them per lane in flight (no address depends on a load), and stores one partial
// sum.  It measures the rate at which this device reads HBM; nothing on the solver path uses it.
constexpr int kStreamU = 8;

__global__ __launch_bounds__(kBlock) void stream_read_kernel(uint64_t n, const double* __restrict__ x,
                                                             double* __restrict__ sums)
{
  const uint32_t lw = xcd_logical_block();
  const uint64_t G = gridDim.x;
  const uint64_t packs = n / 2;
  const uint64_t b = packs * lw / G, e = packs * (lw + 1) / G;
  const double2* __restrict__ p = reinterpret_cast<const double2*>(x);
  double s = 0;
  uint64_t i = b + threadIdx.x;
  for (; i + static_cast<uint64_t>(kStreamU - 1) * kBlock < e; i += static_cast<uint64_t>(kStreamU) * kBlock) {
    double2 v[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) v[u] = p[i + static_cast<uint64_t>(u) * kBlock];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) s += v[u].x + v[u].y;
  }
  for (; i < e; i += kBlock) s += p[i].x + p[i].y;
  if (lw == G - 1 && (n & 1) && threadIdx.x == 0) s += x[n - 1];
  s = block_sum(s);
  if (threadIdx.x == 0) sums[lw] = s;
}

// Gather reference: workgroup lw sums x[idx[i]] over the lw-th contiguous chunk of idx, eight
// indices per lane loaded before their gathers, and stores one partial sum.  With a matrix's column
// indices as idx it times the x gathers of a CSR SpMV of that matrix without the values, the row
// pointers, the row sums or the y stores: the floor that random columns set for the SpMV.
// Measurement only; nothing on the solver path uses it.
__global__ __launch_bounds__(kBlock) void gather_read_kernel(uint32_t m, const uint32_t* __restrict__ idx,
                                                             const double* __restrict__ x, uint32_t n,
                                                             double* __restrict__ sums)
{
  // an index at or beyond n reads x[0] instead
  const uint32_t lw = xcd_logical_block();
  const uint64_t G = gridDim.x;
  const uint32_t b = static_cast<uint32_t>(static_cast<uint64_t>(m) * lw / G);
  const uint32_t e = static_cast<uint32_t>(static_cast<uint64_t>(m) * (lw + 1) / G);
  double s = 0;
  uint32_t i = b + threadIdx.x;
  for (; i + static_cast<uint32_t>(kStreamU - 1) * kBlock < e; i += static_cast<uint32_t>(kStreamU) * kBlock) {
    uint32_t c[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) c[u] = idx[i + static_cast<uint32_t>(u) * kBlock];
    double v[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) v[u] = x[c[u] < n ? c[u] : 0u];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) s += v[u];
  }
  for (; i < e; i += kBlock) {
    const uint32_t c = idx[i];
    s += x[c < n ? c : 0u];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) sums[lw] = s;
}

// Latency references (measurement only; nothing on the solver path uses them).  empty_kernel has
// no memory work: its launch-to-launch time is the floor of a kernel boundary.  grid_sync_kernel
// runs `steps` device-wide synchronisations of co-resident workgroups, the price that a persistent
// (single-launch) solver would pay per iteration instead of a kernel boundary.  Per step, every
// workgroup waits for its own stores, makes them visible to the other XCDs (agent-scope release:
// its XCD's L2 written back) and publishes its arrival flag; workgroup 0 polls all flags, then
// publishes a go flag that the others poll, and every workgroup acquires.  Every wait is bounded
// (spin_limit polls); a workgroup that gives up raises the abort flag, which ends all other waits,
// so the kernel always exits.  flags: G arrival flags, then the go flag, then the abort flag.
__global__ __launch_bounds__(kBlock) void empty_kernel() {}

__device__ __forceinline__ bool probe_spin(const uint32_t* word, uint32_t target, uint32_t* abort_flag,
                                           uint32_t spin_limit)
{
  uint32_t spins = 0;
  while (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++spins > spin_limit) {
      __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__global__ __launch_bounds__(kBlock) void grid_sync_kernel(uint32_t steps, uint32_t* flags, uint32_t spin_limit)
{
  __shared__ uint32_t ok;
  const uint32_t G = gridDim.x;
  uint32_t* const go = flags + G;
  uint32_t* const abort_flag = flags + G + 1;
  for (uint32_t k = 1; k <= steps; ++k) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(&flags[blockIdx.x], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {
      bool good = true;
      for (uint32_t i = threadIdx.x; i < G && good; i += 64) good = probe_spin(&flags[i], k, abort_flag, spin_limit);
      good = __all(good);
      if (threadIdx.x == 0) {
        if (good) __hip_atomic_store(go, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = good;
      }
    } else if (blockIdx.x != 0 && threadIdx.x == 0) {
      ok = probe_spin(go, k, abort_flag, spin_limit);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (!ok) return;
  }
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

// sum |x_i| (norm_1_impl, host_based/vector_operations.hpp:708-733)
template <typename T>
__global__ __launch_bounds__(kBlock) void asum_partial_kernel(uint32_t n, const T* __restrict__ x, uint32_t xs,
                                                              uint32_t xi, T* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  T acc[1] = {0};
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) acc[0] += fabs(at(x, i, xs, xi));
  block_store_partials<1, T>(acc, part, 0, lw);
}

// (value, index) of the largest |x_i|, the smallest index among equal values: the first maximum, as
// the reference's strictly-greater scan finds it (index_norm_inf, host_based/vector_operations.hpp:
// 891-950).  NaN entries never win.
template <typename T>
__device__ __forceinline__ void better(T& v, uint32_t& i, T v2, uint32_t i2)
{
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

template <typename T>
__device__ __forceinline__ void block_argmax(T& v, uint32_t& i)
{
  __shared__ T sv[kWavesPerBlock];
  __shared__ uint32_t si[kWavesPerBlock];
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const T v2 = __shfl_xor(v, o, kWave);
    const uint32_t i2 = __shfl_xor(i, o, kWave);
    better(v, i, v2, i2);
  }
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    sv[wave] = v;
    si[wave] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
#pragma unroll
  for (int w = 1; w < kWavesPerBlock; ++w) better(v, i, sv[w], si[w]);
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void iamax_partial_kernel(uint32_t n, const T* __restrict__ x, uint32_t xs,
                                                               uint32_t xi, T* __restrict__ pv,
                                                               uint32_t* __restrict__ pi)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  T v = 0;
  uint32_t idx = 0xFFFFFFFFu;
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    const T a = fabs(at(x, i, xs, xi));
    if (a > v) {
      v = a;
      idx = i;
    }
  }
  block_argmax(v, idx);
  if (threadIdx.x == 0) {
    pv[lw] = v;
    pi[lw] = idx;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void iamax_final_kernel(const T* __restrict__ pv, const uint32_t* __restrict__ pi,
                                                             uint32_t count, uint32_t* __restrict__ out)
{
  T v = 0;
  uint32_t idx = 0xFFFFFFFFu;
  for (uint32_t k = threadIdx.x; k < count; k += kBlock) better(v, idx, pv[k], pi[k]);
  block_argmax(v, idx);
  if (threadIdx.x == 0) *out = idx;
}

// x = alpha x (av with vec1 = vec2, host_based/vector_operations.hpp:88-123)
template <typename T>
__global__ __launch_bounds__(kBlock) void scal_kernel(uint32_t n, T* __restrict__ x, uint32_t xs, uint32_t xi, T a)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    const size_t k = static_cast<size_t>(i) * xi + xs;
    x[k] = x[k] * a;
  }
}

// y = x (kind 0), swap (kind 1, vector_swap :302-327), plane rotation x' = c x + s y, y' = c y - s x
// (kind 2, plane_rotation :1080-1110)
template <typename T, int Kind>
__global__ __launch_bounds__(kBlock) void pair_kernel(uint32_t n, T* __restrict__ x, uint32_t xs, uint32_t xi,
                                                      T* __restrict__ y, uint32_t ys, uint32_t yi, T c, T s)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    const size_t kx = static_cast<size_t>(i) * xi + xs;
    const size_t ky = static_cast<size_t>(i) * yi + ys;
    const T a = x[kx];
    if constexpr (Kind == 0) {
      y[ky] = a;
    } else if constexpr (Kind == 1) {
      x[kx] = y[ky];
      y[ky] = a;
    } else {
      const T b = y[ky];
      x[kx] = c * a + s * b;
      y[ky] = c * b - s * a;
    }
  }
}

template <typename T>
static int asum_to_host(uint32_t n, const T* x, uint32_t xs, uint32_t xi, T* result, void* stream)
{
  if (!result) return invalid("result is NULL");
  *result = 0;
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || xi == 0) return invalid("null vector or zero stride");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t grid = vec_grid(n);
  void* ws = nullptr;
  VCLHIP_HIP(hipMallocAsync(&ws, (grid + 1) * sizeof(T), s));
  T* part = static_cast<T*>(ws);
  hipLaunchKernelGGL(asum_partial_kernel<T>, dim3(grid), dim3(kBlock), 0, s, n, x, xs, xi, part);
  int st = check_launch("asum_partial_kernel");
  if (st == 0) {
    hipLaunchKernelGGL(sum_partials_kernel<T>, dim3(1), dim3(kBlock), 0, s, part, grid, part + grid);
    st = check_launch("sum_partials_kernel");
  }
  if (st == 0) st = hip_status(hipMemcpyAsync(result, part + grid, sizeof(T), hipMemcpyDeviceToHost, s), "asum D2H");
  int st2 = hip_status(hipFreeAsync(ws, s), "hipFreeAsync");
  int st3 = hip_status(hipStreamSynchronize(s), "asum sync");
  return st ? st : (st2 ? st2 : st3);
}

template <typename T>
static int iamax_to_host(uint32_t n, const T* x, uint32_t xs, uint32_t xi, uint32_t* index, void* stream)
{
  if (!index) return invalid("index is NULL");
  *index = xs;
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || xi == 0) return invalid("null vector or zero stride");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t grid = vec_grid(n);
  void* ws = nullptr;
  VCLHIP_HIP(hipMallocAsync(&ws, grid * (sizeof(T) + sizeof(uint32_t)) + 64, s));
  T* pv = static_cast<T*>(ws);
  uint32_t* pi = reinterpret_cast<uint32_t*>(pv + grid);
  uint32_t* out = pi + grid;
  hipLaunchKernelGGL(iamax_partial_kernel<T>, dim3(grid), dim3(kBlock), 0, s, n, x, xs, xi, pv, pi);
  int st = check_launch("iamax_partial_kernel");
  if (st == 0) {
    hipLaunchKernelGGL(iamax_final_kernel<T>, dim3(1), dim3(kBlock), 0, s, pv, pi, grid, out);
    st = check_launch("iamax_final_kernel");
  }
  uint32_t r = 0xFFFFFFFFu;
  if (st == 0) st = hip_status(hipMemcpyAsync(&r, out, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "iamax D2H");
  int st2 = hip_status(hipFreeAsync(ws, s), "hipFreeAsync");
  int st3 = hip_status(hipStreamSynchronize(s), "iamax sync");
  if (st || st2 || st3) return st ? st : (st2 ? st2 : st3);
  *index = (r == 0xFFFFFFFFu) ? xs : r;  // no entry above 0: the reference returns the start offset
  return VCLHIP_SUCCESS;
}

template <typename T, int Kind>
static int pair_op(uint32_t n, T* x, uint32_t xs, uint32_t xi, T* y, uint32_t ys, uint32_t yi, T c, T sn,
                   void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || !y || xi == 0 || yi == 0) return invalid("null vector or zero stride");
  hipLaunchKernelGGL((pair_kernel<T, Kind>), dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n,
                     x, xs, xi, y, ys, yi, c, sn);
  return check_launch("pair_kernel");
}

template <typename T>
static int scal_op(uint32_t n, T* x, uint32_t xs, uint32_t xi, T a, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!x || xi == 0) return invalid("null vector or zero stride");
  hipLaunchKernelGGL(scal_kernel<T>, dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, x, xs,
                     xi, a);
  return check_launch("scal_kernel");
}

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

int vclhip_stream_read_f64(uint64_t n, const double* x, double* sums, uint32_t num_sums, void* stream)
{
  if (!x || !sums) return invalid("null pointer");
  if (num_sums < 8 || num_sums > (1u << 20) || (num_sums & 7u)) return invalid("num_sums: a multiple of 8 in [8, 2^20]");
  if (!aligned16({x})) return invalid("x must be 16-byte aligned");
  launch_timed(stream_read_kernel, dim3(num_sums), dim3(kBlock), static_cast<hipStream_t>(stream), n, x, sums);
  return check_launch("stream_read_kernel");
}

int vclhip_gather_read_f64(uint32_t m, const uint32_t* idx, const double* x, uint32_t n, double* sums,
                           uint32_t num_sums, void* stream)
{
  if (!idx || !x || !sums) return invalid("null pointer");
  if (num_sums < 8 || num_sums > (1u << 20) || (num_sums & 7u)) return invalid("num_sums: a multiple of 8 in [8, 2^20]");
  if (n == 0) return invalid("x must hold at least one element");
  launch_timed(gather_read_kernel, dim3(num_sums), dim3(kBlock), static_cast<hipStream_t>(stream), m, idx, x, n,
               sums);
  return check_launch("gather_read_kernel");
}

int vclhip_empty_launch(uint32_t workgroups, void* stream)
{
  if (workgroups == 0 || workgroups > (1u << 20)) return invalid("workgroups: 1 to 2^20");
  hipLaunchKernelGGL(empty_kernel, dim3(workgroups), dim3(kBlock), 0, static_cast<hipStream_t>(stream));
  return check_launch("empty_kernel");
}

int vclhip_grid_sync_probe(uint32_t workgroups, uint32_t steps, uint32_t* flags, void* stream)
{
  if (!flags) return invalid("null pointer");
  if (workgroups == 0 || workgroups > 4096) return invalid("workgroups: 1 to 4096");
  int per_cu = 0;
  VCLHIP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, grid_sync_kernel, kBlock, 0));
  if (static_cast<uint64_t>(workgroups) > static_cast<uint64_t>(per_cu) * device_cu_count())
    return invalid("workgroups: more than the device holds at once");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t bytes = sizeof(uint32_t) * (static_cast<uint64_t>(workgroups) + 2);
  VCLHIP_TRY(vclhip_fill_zero(flags, bytes, stream));
  // about 10^7 polls of at least 64 cycles each: seconds, far beyond any step of a resident grid
  hipLaunchKernelGGL(grid_sync_kernel, dim3(workgroups), dim3(kBlock), 0, s, steps, flags, 1u << 24);
  return check_launch("grid_sync_kernel");
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

int vclhip_avbv_f32(uint32_t n, float* v1, uint32_t s1, uint32_t i1, const float* v2, uint32_t s2, uint32_t i2,
                    float alpha, const float* v3, uint32_t s3, uint32_t i3, float beta, void* stream)
{
  if (n == 0) return VCLHIP_SUCCESS;
  if (!v1 || !v2 || !v3 || !i1 || !i2 || !i3) return invalid("null vector or zero stride");
  hipLaunchKernelGGL(avbv_kernel<float>, dim3(vec_grid(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, v1,
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

int vclhip_asum_f64(uint32_t n, const double* x, uint32_t xs, uint32_t xi, double* result, void* stream)
{
  return asum_to_host<double>(n, x, xs, xi, result, stream);
}

int vclhip_asum_f32(uint32_t n, const float* x, uint32_t xs, uint32_t xi, float* result, void* stream)
{
  return asum_to_host<float>(n, x, xs, xi, result, stream);
}

int vclhip_iamax_f64(uint32_t n, const double* x, uint32_t xs, uint32_t xi, uint32_t* index, void* stream)
{
  return iamax_to_host<double>(n, x, xs, xi, index, stream);
}

int vclhip_iamax_f32(uint32_t n, const float* x, uint32_t xs, uint32_t xi, uint32_t* index, void* stream)
{
  return iamax_to_host<float>(n, x, xs, xi, index, stream);
}

int vclhip_scal_f64(uint32_t n, double* x, uint32_t xs, uint32_t xi, double alpha, void* stream)
{
  return scal_op<double>(n, x, xs, xi, alpha, stream);
}

int vclhip_scal_f32(uint32_t n, float* x, uint32_t xs, uint32_t xi, float alpha, void* stream)
{
  return scal_op<float>(n, x, xs, xi, alpha, stream);
}

int vclhip_copy_f64(uint32_t n, const double* x, uint32_t xs, uint32_t xi, double* y, uint32_t ys, uint32_t yi,
                    void* stream)
{
  return pair_op<double, 0>(n, const_cast<double*>(x), xs, xi, y, ys, yi, 0.0, 0.0, stream);
}

int vclhip_copy_f32(uint32_t n, const float* x, uint32_t xs, uint32_t xi, float* y, uint32_t ys, uint32_t yi,
                    void* stream)
{
  return pair_op<float, 0>(n, const_cast<float*>(x), xs, xi, y, ys, yi, 0.0f, 0.0f, stream);
}

int vclhip_swap_f64(uint32_t n, double* x, uint32_t xs, uint32_t xi, double* y, uint32_t ys, uint32_t yi,
                    void* stream)
{
  return pair_op<double, 1>(n, x, xs, xi, y, ys, yi, 0.0, 0.0, stream);
}

int vclhip_swap_f32(uint32_t n, float* x, uint32_t xs, uint32_t xi, float* y, uint32_t ys, uint32_t yi, void* stream)
{
  return pair_op<float, 1>(n, x, xs, xi, y, ys, yi, 0.0f, 0.0f, stream);
}

int vclhip_rot_f64(uint32_t n, double* x, uint32_t xs, uint32_t xi, double* y, uint32_t ys, uint32_t yi, double c,
                   double s, void* stream)
{
  return pair_op<double, 2>(n, x, xs, xi, y, ys, yi, c, s, stream);
}

int vclhip_rot_f32(uint32_t n, float* x, uint32_t xs, uint32_t xi, float* y, uint32_t ys, uint32_t yi, float c,
                   float s, void* stream)
{
  return pair_op<float, 2>(n, x, xs, xi, y, ys, yi, c, s, stream);
}

}  // extern "C"
