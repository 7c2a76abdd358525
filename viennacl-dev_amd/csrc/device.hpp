// device.hpp -- device helpers shared by the libvclhip kernels (gfx950, wave64).
//
// Reductions are deterministic: every wavefront reduces with an xor butterfly (all lanes end
// with the same sum, since a+b == b+a), the four wavefronts of a 256-lane workgroup are combined
// in a fixed order, and per-workgroup partials are summed by one workgroup in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <type_traits>

namespace vclhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;           // lanes per workgroup of every kernel here
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kMaxPartials = 4096; // upper bound on workgroups that write partials

// Makes a register value opaque to the compiler at this point: the loads that produce it are
// issued and waited for here, and cannot be sunk into a later conditional use.  The staged SpMV
// loops use it so that every entry's loads are in flight together; without it hipcc moved the
// first entry's column, value and x loads into its conditional LDS store, behind full waits for
// all the other loads (two more memory round trips per row block).
template <typename T>
__device__ __forceinline__ void keep_here(T& v)
{
  asm volatile("" : "+v"(v));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum over aligned groups of L lanes (L a power of two <= 64).
template <typename T>
__device__ __forceinline__ T group_sum(T v, uint32_t L)
{
  for (uint32_t o = L >> 1; o > 0; o >>= 1) v += __shfl_xor(v, static_cast<int>(o), kWave);
  return v;
}

// Workgroup sum of one value per lane; the result is returned to every lane.
template <typename T>
__device__ __forceinline__ T block_sum(T v)
{
  __shared__ T red[kWavesPerBlock];
  v = wave_sum(v);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[wave] = v;
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int w = 1; w < kWavesPerBlock; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Reduces K per-lane accumulators over the workgroup and stores them at part[q*pstride + slot].
template <int K, typename A>
__device__ __forceinline__ void block_store_partials(const A (&acc)[K], A* part, uint32_t pstride, uint32_t slot)
{
  __shared__ A red[K][kWavesPerBlock];
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    A v = wave_sum(acc[q]);
    if ((threadIdx.x & (kWave - 1)) == 0) red[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    A s = red[threadIdx.x][0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) s += red[threadIdx.x][w];
    part[threadIdx.x * pstride + slot] = s;
  }
}

// Sum of part[0..count) by one workgroup in a fixed order; result returned to every lane.
// Lane t adds part[t], part[t+256], ... in that order; the loads of a batch of eight are issued
// together so that the latency is paid once per batch rather than once per partial.
template <typename A>
__device__ __forceinline__ A block_sum_partials(const A* part, uint32_t count)
{
  constexpr uint32_t kBatch = 8;
  A s = 0;
  uint32_t i = threadIdx.x;
  for (; i + (kBatch - 1) * kBlock < count; i += kBatch * kBlock) {
    A v[kBatch];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) v[k] = part[i + k * kBlock];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) s += v[k];
  }
  for (; i < count; i += kBlock) s += part[i];
  return block_sum(s);
}

// The workgroup sum of K per-lane values: a wave64 xor butterfly per wave, then the four waves in
// order (the combination step of block_sum_partials_k).
template <int K, typename A>
__device__ __forceinline__ void block_combine_k(A (&s)[K], A (&out)[K])
{
  __shared__ A red[K][kWavesPerBlock];
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    s[q] = wave_sum(s[q]);
    if ((threadIdx.x & (kWave - 1)) == 0) red[q][wave] = s[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    out[q] = red[q][0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) out[q] += red[q][w];
  }
  __syncthreads();
}

// K block_sum_partials in one pass: lane t adds p[q][t], p[q][t+256], ... in the same order as
// block_sum_partials, and the workgroup combines in the same tree, so the results equal K separate
// calls; the K loads of a step are in flight together and the workgroup synchronizes once.
template <int K, typename A>
__device__ __forceinline__ void block_sum_partials_k(const A* const (&p)[K], const uint32_t (&count)[K], A (&out)[K])
{
  uint32_t nmax = 0;
#pragma unroll
  for (int q = 0; q < K; ++q) nmax = count[q] > nmax ? count[q] : nmax;
  A s[K];
#pragma unroll
  for (int q = 0; q < K; ++q) s[q] = 0;
  // batches of B steps: the B*K loads of a batch are issued before any is added
  constexpr uint32_t B = K <= 2 ? 8 : 4;
  uint32_t i = threadIdx.x;
  for (; i + (B - 1) * kBlock < nmax; i += B * kBlock) {
    A v[B][K];
#pragma unroll
    for (uint32_t b = 0; b < B; ++b)
#pragma unroll
      for (int q = 0; q < K; ++q) v[b][q] = i + b * kBlock < count[q] ? p[q][i + b * kBlock] : A(0);
#pragma unroll
    for (uint32_t b = 0; b < B; ++b)
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (i + b * kBlock < count[q]) s[q] += v[b][q];
  }
  for (; i < nmax; i += kBlock) {
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (i < count[q]) s[q] += p[q][i];
  }
  block_combine_k<K, A>(s, out);
}

template <typename A>
__device__ __forceinline__ void block_sum_partials3(const A* pa, uint32_t na, const A* pb, uint32_t nb,
                                                    const A* pc, uint32_t nc, A& ra, A& rb, A& rc)
{
  const A* const p[3] = {pa, pb, pc};
  const uint32_t count[3] = {na, nb, nc};
  A out[3];
  block_sum_partials_k<3, A>(p, count, out);
  ra = out[0];
  rb = out[1];
  rc = out[2];
}

// [begin, end) of the lw-th of G chunks of n elements with both ends on multiples of W (the last
// chunk also takes the tail), so that W-wide vector accesses of a W*sizeof(T)-aligned array stay
// aligned.
__device__ __forceinline__ void chunk_range_w(uint32_t n, uint32_t W, uint32_t lw, uint32_t G, uint32_t& begin,
                                              uint32_t& end)
{
  const uint32_t packs = n / W;
  begin = static_cast<uint32_t>((static_cast<uint64_t>(packs) * lw) / G) * W;
  end = static_cast<uint32_t>((static_cast<uint64_t>(packs) * (lw + 1)) / G) * W;
  if (lw == G - 1) end = n;
}

// Calls f(i, integral_constant<W>) for the W-element packs of [begin, end) (lane-interleaved, so a
// wavefront touches 64*W consecutive elements per step) and f(i, integral_constant<1>) for the tail.
template <int W, class F>
__device__ __forceinline__ void for_packs(uint32_t begin, uint32_t end, F&& f)
{
  const uint32_t packed_end = begin + ((end - begin) / W) * W;
  for (uint32_t i = begin + W * threadIdx.x; i < packed_end; i += W * kBlock) f(i, std::integral_constant<int, W>());
  if (W > 1)
    for (uint32_t i = packed_end + threadIdx.x; i < end; i += kBlock) f(i, std::integral_constant<int, 1>());
}

// W-wide loads and stores (W = 1, 2, 4 for double as 16-byte halves; 1, 2, 4 for float).
template <int W, typename T>
struct Pack {
  T v[W];
};

template <int W, typename T>
__device__ __forceinline__ Pack<W, T> ldp(const T* __restrict__ p, uint32_t i)
{
  Pack<W, T> r;
  if constexpr (W == 2 && sizeof(T) == 8) {
    const double2 q = *reinterpret_cast<const double2*>(p + i);
    r.v[0] = q.x;
    r.v[1] = q.y;
  } else if constexpr (W == 4 && sizeof(T) == 8) {
    const double2 q0 = *reinterpret_cast<const double2*>(p + i);
    const double2 q1 = *reinterpret_cast<const double2*>(p + i + 2);
    r.v[0] = q0.x;
    r.v[1] = q0.y;
    r.v[2] = q1.x;
    r.v[3] = q1.y;
  } else if constexpr (W == 2 && sizeof(T) == 4) {
    const float2 q = *reinterpret_cast<const float2*>(p + i);
    r.v[0] = q.x;
    r.v[1] = q.y;
  } else if constexpr (W == 4 && sizeof(T) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p + i);
    r.v[0] = q.x;
    r.v[1] = q.y;
    r.v[2] = q.z;
    r.v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) r.v[k] = p[i + k];
  }
  return r;
}

template <int W, typename T>
__device__ __forceinline__ void stp(T* __restrict__ p, uint32_t i, const Pack<W, T>& r)
{
  if constexpr (W == 2 && sizeof(T) == 8) {
    *reinterpret_cast<double2*>(p + i) = make_double2(r.v[0], r.v[1]);
  } else if constexpr (W == 4 && sizeof(T) == 8) {
    *reinterpret_cast<double2*>(p + i) = make_double2(r.v[0], r.v[1]);
    *reinterpret_cast<double2*>(p + i + 2) = make_double2(r.v[2], r.v[3]);
  } else if constexpr (W == 2 && sizeof(T) == 4) {
    *reinterpret_cast<float2*>(p + i) = make_float2(r.v[0], r.v[1]);
  } else if constexpr (W == 4 && sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p + i) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) p[i + k] = r.v[k];
  }
}

// Vector width used for arrays of T: 16-byte packs.
template <typename T>
constexpr int kVecW = 16 / sizeof(T);

// Launches KERNEL_W2 (16-byte packs) when `v2` holds, else KERNEL_W1, with 256-lane workgroups.
#define VCLHIP_LAUNCH_W(v2, KERNEL_W2, KERNEL_W1, grid, stream, ...)                                  \
  do {                                                                                               \
    if (v2)                                                                                          \
      hipLaunchKernelGGL(KERNEL_W2, dim3(grid), dim3(::vclhip::kBlock), 0, stream, __VA_ARGS__);     \
    else                                                                                             \
      hipLaunchKernelGGL(KERNEL_W1, dim3(grid), dim3(::vclhip::kBlock), 0, stream, __VA_ARGS__);     \
  } while (0)

// True when every pointer is aligned for 16-byte packs.
inline bool aligned16(std::initializer_list<const void*> ptrs)
{
  for (const void* p : ptrs)
    if (p && (reinterpret_cast<uintptr_t>(p) & 15u)) return false;
  return true;
}

// Workgroup w is dispatched to XCD (w mod 8).  Renumber so that each XCD owns a contiguous
// range of logical workgroups, which keeps neighbouring rows (and the x entries they gather)
// in one XCD's L2.  gridDim.x must be a multiple of 8.
__device__ __forceinline__ uint32_t xcd_logical_block()
{
  const uint32_t G = gridDim.x;
  const uint32_t w = blockIdx.x;
  return (w & 7u) * (G >> 3) + (w >> 3);
}

// [begin, end) of the lw-th of G near-equal contiguous chunks of `total` items.
__device__ __forceinline__ void chunk_range(uint32_t total, uint32_t lw, uint32_t G, uint32_t& begin, uint32_t& end)
{
  begin = static_cast<uint32_t>((static_cast<uint64_t>(total) * lw) / G);
  end = static_cast<uint32_t>((static_cast<uint64_t>(total) * (lw + 1)) / G);
}

}  // namespace vclhip
