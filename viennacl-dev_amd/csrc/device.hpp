// device.hpp -- device helpers shared by the libvclhip kernels (gfx950, wave64).
//
// Reductions are deterministic: every wavefront reduces with an xor butterfly (all lanes end
// with the same sum, since a+b == b+a), the four wavefronts of a 256-lane workgroup are combined
// in a fixed order, and per-workgroup partials are summed by one workgroup in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vclhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;           // lanes per workgroup of every kernel here
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kMaxPartials = 4096; // upper bound on workgroups that write partials

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
template <typename A>
__device__ __forceinline__ A block_sum_partials(const A* part, uint32_t count)
{
  A s = 0;
  for (uint32_t i = threadIdx.x; i < count; i += kBlock) s += part[i];
  return block_sum(s);
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
