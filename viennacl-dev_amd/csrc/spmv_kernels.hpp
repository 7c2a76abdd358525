// spmv_kernels.hpp -- CSR and SELL-C SpMV kernel templates with a pluggable per-row epilogue.
//
// The epilogue receives each finished row sum; plain SpMV stores alpha*s + beta*y, the Krylov
// kernels also store the row and accumulate the inner products that the solver needs next
// (reference: cuda/iterative_operations.hpp:113-274, 543-607, 895-1074, 1381-1453 fuse the same
// products into their SpMV kernels).  Epilogue interface:
//   using acc_t = ...;  static constexpr int K = number of inner products (0..3);
//   __device__ void operator()(uint32_t row, T sum, acc_t* acc) const;
// Workgroups write their K partial sums to part[q * pstride + logical_block].  A kernel returns
// at once when (*stop & stop_mask) is nonzero, which lets a solver enqueue iterations ahead of
// its convergence test.
//
// Every kernel here is launched with 256 lanes per workgroup and a grid that is a multiple of 8
// (see xcd_logical_block).  Floating-point contraction is disabled for the whole library
// (-ffp-contract=off) so that a row reduced by one lane sums its products in the same order and
// with the same roundings as the reference's sequential host loop.
#pragma once

#include "device.hpp"
#include "plan.hpp"

namespace vclhip {

template <typename T>
struct CsrDev {
  const uint32_t* __restrict__ row_ptr;
  const uint32_t* __restrict__ col;
  const T* __restrict__ val;
  const uint32_t* __restrict__ blocks;
  uint32_t num_blocks;
  uint32_t n_rows;
};

// HYB through the CSR kernel: the CSR part with its row-block plan, plus the ELL part, whose
// entries of a row are summed before the row's CSR entries (the reference's order,
// host_based/sparse_matrix_operations.hpp:1893-1912).
template <typename T>
struct HybDev : CsrDev<T> {
  const uint32_t* __restrict__ ell_coords;
  const T* __restrict__ ell_val;
  uint32_t ell_ld;     // internal_size1
  uint32_t ell_width;  // internal_ellnnz
};

template <typename T>
struct SellDev {
  const uint32_t* __restrict__ cpb;
  const uint32_t* __restrict__ col;
  const uint32_t* __restrict__ start;
  const T* __restrict__ val;
  uint32_t n_rows;
  uint32_t C;
  const int16_t* __restrict__ dcol = nullptr;  // Delta kernels: column - row as 16 bits (col unused)
};

// Loads of the matrix streams: nontemporal (not kept in the caches) when NT holds.
template <bool NT, typename U>
__device__ __forceinline__ U ld_stream(const U* p)
{
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <typename T, bool Contig>
__device__ __forceinline__ T load_x(const T* __restrict__ x, uint32_t c, uint32_t xs, uint32_t xi)
{
  if (Contig) return x[c];
  return x[static_cast<size_t>(c) * xi + xs];
}

// Sum of the ELL entries of one row in column order; zero entries are left out of the sum
// (host_based/sparse_matrix_operations.hpp:1516-1524, :1897-1904).  The value and the column are
// loaded together and x then once, at the column or, for a zero value (whose column may be any
// padding), at x's first element: two dependent round trips instead of three.
constexpr uint32_t kEllGroup = 8;  // ELL entries whose loads a row issues together

template <typename T, bool Contig>
__device__ __forceinline__ T ell_row_sum(const uint32_t* __restrict__ coords, const T* __restrict__ val,
                                         uint32_t ld, uint32_t width, uint32_t row, const T* __restrict__ x,
                                         uint32_t xs, uint32_t xi)
{
  // groups of kEllGroup entries, all loads of a group before its sums (clamped indices, no branch)
  T sum = 0;
  for (uint32_t k0 = 0; k0 < width; k0 += kEllGroup) {
    T v[kEllGroup], xv[kEllGroup];
    uint32_t c[kEllGroup];
#pragma unroll
    for (uint32_t j = 0; j < kEllGroup; ++j) {
      const size_t off = static_cast<size_t>(min(k0 + j, width - 1)) * ld + row;
      v[j] = val[off];
      c[j] = coords[off];
    }
#pragma unroll
    for (uint32_t j = 0; j < kEllGroup; ++j) xv[j] = load_x<T, Contig>(x, (v[j] > 0 || v[j] < 0) ? c[j] : 0u, xs, xi);
#pragma unroll
    for (uint32_t j = 0; j < kEllGroup; ++j)
      if (k0 + j < width && (v[j] > 0 || v[j] < 0)) sum += xv[j] * v[j];
  }
  return sum;
}

// y = alpha*s + beta*y with the reference's beta test (host_based/sparse_matrix_operations.hpp:177-183).
template <typename T, bool Contig>
struct EpiAxpby {
  using acc_t = T;
  static constexpr int K = 0;
  T* __restrict__ y;
  uint32_t ys, yi;
  T alpha, beta;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t*) const
  {
    const size_t idx = Contig ? row : static_cast<size_t>(row) * yi + ys;
    if (beta < 0 || beta > 0)
      y[idx] = alpha * s + beta * y[idx];
    else
      y[idx] = alpha * s;
  }
};

// CSR SpMV over the row-block plan.  Each logical workgroup walks a contiguous chunk of blocks.
// CS is the distance between consecutive column indices: 1 for CSR, 2 for COO, whose (row, col)
// pairs are read through col = coords + 1 with the row pointers of its row order.  Hyb: the matrix
// is a HybDev, the plan covers its CSR part, and each row sum starts from the row's ELL sum.
template <typename T, bool Contig, class Epi, bool NT = false, int CS = 1, bool Hyb = false>
__global__ __launch_bounds__(kBlock) void csr_kernel(std::conditional_t<Hyb, HybDev<T>, CsrDev<T>> A,
                                                     const T* __restrict__ x, uint32_t xs, uint32_t xi, Epi epi,
                                                     typename Epi::acc_t* __restrict__ part, uint32_t pstride,
                                                     const uint32_t* __restrict__ stop, uint32_t stop_mask)
{
  using acc_t = typename Epi::acc_t;
  constexpr int KA = Epi::K > 0 ? Epi::K : 1;
  if (stop && (*stop & stop_mask)) return;

  __shared__ T stage[kCsrNnzCap];
  __shared__ uint32_t srp[kCsrWG + 1];
  __shared__ T spre[Hyb ? kCsrWG : 1];  // ELL sums of the block's rows, for the wavefront pass

  const uint32_t tid = threadIdx.x;
  const uint32_t lw = xcd_logical_block();
  uint32_t b0, b1;
  chunk_range(A.num_blocks, lw, gridDim.x, b0, b1);

  acc_t acc[KA];
#pragma unroll
  for (int q = 0; q < KA; ++q) acc[q] = 0;

  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t d0 = A.blocks[b];
    const uint32_t r0 = d0 & (kPlanRowLimit - 1);
    const uint32_t mode = d0 >> 28;
    const uint32_t r1 = A.blocks[b + 1] & (kPlanRowLimit - 1);
    const uint32_t e0 = A.row_ptr[r0];
    const uint32_t e1 = A.row_ptr[r1];

    if (mode == kModeLongRow) {
      // one row longer than the staging buffer: strided partial sums, then a workgroup reduction
      T s = 0;
      for (uint32_t e = e0 + tid; e < e1; e += kBlock)
        s += A.val[e] * load_x<T, Contig>(x, A.col[static_cast<size_t>(e) * CS], xs, xi);
      s = block_sum(s);
      if (tid == 0) {
        if constexpr (Hyb) s = ell_row_sum<T, Contig>(A.ell_coords, A.ell_val, A.ell_ld, A.ell_width, r0, x, xs, xi) + s;
        epi(r0, s, acc);
      }
      continue;
    }

    // Phase 1: coalesced loads of the block's entries; products into LDS.  Every lane issues all its
    // loads before using any (indices past the block are clamped to its last entry, their products
    // are not stored), so a wavefront has U value/column loads and then U gathers in flight.  With
    // NT the matrix streams are read nontemporally (for matrices larger than the Infinity Cache,
    // whose lines would not be reused anyway).
    const uint32_t cnt = e1 - e0;
    const uint32_t nrows = r1 - r0;
    if (tid < nrows) srp[tid + 1] = A.row_ptr[r0 + tid + 1] - e0;
    T pre = 0;  // Hyb: the ELL sum of row r0 + tid
    if constexpr (Hyb) {
      if (tid < nrows) {
        pre = ell_row_sum<T, Contig>(A.ell_coords, A.ell_val, A.ell_ld, A.ell_width, r0 + tid, x, xs, xi);
        if (mode == kModeMixed) spre[tid] = pre;
      }
    }
    if (cnt > 0) {
      constexpr int U = kCsrNnzCap / kBlock;
      uint32_t c[U];
      T v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t e = e0 + min(tid + k * kBlock, cnt - 1);
        c[k] = ld_stream<NT>(A.col + static_cast<size_t>(e) * CS);
        v[k] = ld_stream<NT>(A.val + e);
      }
      T pr[U];
#pragma unroll
      for (int k = 0; k < U; ++k) pr[k] = load_x<T, Contig>(x, c[k], xs, xi);
#pragma unroll
      for (int k = 0; k < U; ++k) {
        pr[k] = v[k] * pr[k];
        keep_here(pr[k]);
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = tid + k * kBlock;
        if (i < cnt) stage[i] = pr[k];
      }
    }
    if (tid == 0) srp[0] = 0;
    __syncthreads();

    // Phase 2a: each row of at most kShortRow entries is summed by one lane, in order.
    if (tid < nrows) {
      const uint32_t a = srp[tid];
      const uint32_t z = srp[tid + 1];
      if (z - a <= kShortRow) {
        T s = pre;
        for (uint32_t k = a; k < z; ++k) s += stage[k];
        epi(r0 + tid, s, acc);
      }
    }
    // Phase 2b: each longer row is summed by one wavefront (strided partial sums, xor tree); the
    // four wavefronts take the block's long rows in turn.
    if (mode == kModeMixed) {
      const uint32_t wave = tid / kWave;
      const uint32_t lane = tid & (kWave - 1);
      uint32_t turn = 0;
      for (uint32_t base = 0; base < nrows; base += kWave) {
        const uint32_t i = base + lane;
        const bool is_long = i < nrows && srp[i + 1] - srp[i] > kShortRow;
        unsigned long long m = __ballot(is_long);
        while (m) {
          const uint32_t bit = static_cast<uint32_t>(__builtin_ctzll(m));
          m &= m - 1;
          if ((turn++ & (kWavesPerBlock - 1)) != wave) continue;
          const uint32_t row = base + bit;
          const uint32_t a = srp[row];
          const uint32_t z = srp[row + 1];
          T s = 0;
          for (uint32_t k = a + lane; k < z; k += kWave) s += stage[k];
          s = wave_sum(s);
          if constexpr (Hyb) s = spre[row] + s;
          if (lane == 0) epi(r0 + row, s, acc);
        }
      }
    }
    __syncthreads();
  }

  if constexpr (Epi::K > 0) block_store_partials<Epi::K, acc_t>(acc, part, pstride, lw);
}

constexpr uint32_t kSellGroup = 8;  // columns whose loads a SELL row issues together

// SELL-C SpMV, one lane per row; with C = 64 a wavefront is exactly one slice, so all lanes of a
// wavefront run the same number of columns and every load is one contiguous 64-lane access.
// Delta: the columns are stored as 16-bit offsets from the row (the mixed solver's inner copy),
// which cuts the index stream from 4 to 2 bytes per stored entry.
template <typename T, bool Contig, class Epi, bool Delta = false, bool NT = false>
__global__ __launch_bounds__(kBlock) void sell_kernel(SellDev<T> A, const T* __restrict__ x, uint32_t xs,
                                                      uint32_t xi, Epi epi, typename Epi::acc_t* __restrict__ part,
                                                      uint32_t pstride, const uint32_t* __restrict__ stop,
                                                      uint32_t stop_mask)
{
  using acc_t = typename Epi::acc_t;
  constexpr int KA = Epi::K > 0 ? Epi::K : 1;
  if (stop && (*stop & stop_mask)) return;

  const uint32_t tid = threadIdx.x;
  const uint32_t lw = xcd_logical_block();
  const uint32_t tiles = (A.n_rows + kBlock - 1) / kBlock;
  uint32_t t0, t1;
  chunk_range(tiles, lw, gridDim.x, t0, t1);

  acc_t acc[KA];
#pragma unroll
  for (int q = 0; q < KA; ++q) acc[q] = 0;

  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t row = t * kBlock + tid;
    if (row < A.n_rows) {
      const uint32_t s = row / A.C;
      const uint32_t lane = row - s * A.C;
      const uint32_t ncols = A.cpb[s];
      const size_t base = static_cast<size_t>(A.start[s]) + lane;
      const T* __restrict__ ev = A.val + base;
      T sum = 0;
      // Groups of kSellGroup columns: every load of a group is issued before the first sum, the
      // column index clamped so that a short last group re-reads its last column instead of
      // branching (a counted loop left a 7-column slice mostly to the unroll remainder, one column
      // per memory round trip).  Zero values add nothing (host_based/sparse_matrix_operations.hpp:1831-1833).
      for (uint32_t k0 = 0; k0 < ncols; k0 += kSellGroup) {
        T v[kSellGroup], xv[kSellGroup];
        uint32_t c[kSellGroup];
#pragma unroll
        for (uint32_t j = 0; j < kSellGroup; ++j) {
          const size_t off = static_cast<size_t>(min(k0 + j, ncols - 1)) * A.C;
          v[j] = ld_stream<NT>(ev + off);
          if constexpr (Delta)
            c[j] = static_cast<uint32_t>(static_cast<int32_t>(row) + ld_stream<NT>(A.dcol + base + off));
          else
            c[j] = ld_stream<NT>(A.col + base + off);
        }
#pragma unroll
        for (uint32_t j = 0; j < kSellGroup; ++j) xv[j] = load_x<T, Contig>(x, c[j], xs, xi);
#pragma unroll
        for (uint32_t j = 0; j < kSellGroup; ++j)
          if (k0 + j < ncols) sum += (v[j] > 0 || v[j] < 0) ? xv[j] * v[j] : T(0);
      }
      epi(row, sum, acc);
    }
  }

  if constexpr (Epi::K > 0) block_store_partials<Epi::K, acc_t>(acc, part, pstride, lw);
}

constexpr uint32_t kDiaGroup = 8;  // diagonals whose loads a DIA row issues together

// Whether dia_kernel issues a group's loads together (the default) or runs the counted loop: an
// epilogue opts out with `static constexpr bool kDiaGrouped = false`; a coded copy always groups
// (its one-byte loads and table lookups measured 9-12% faster grouped under the BiCGStab epilogues).
template <class E, class = void>
struct dia_grouped : std::true_type {};
template <class E>
struct dia_grouped<E, std::void_t<decltype(E::kDiaGrouped)> > : std::bool_constant<E::kDiaGrouped> {};

// DIA: nd diagonals of n values each, diagonal-major (val[d * n + row] multiplies x[row + off[d]]),
// offsets ascending.  The mixed-precision solver's inner copy of a stencil-like matrix.
// Coded: when every diagonal holds at most kDiaCodes - 1 distinct values, code[d * n + row] is the
// index of val[d * n + row] in table[d * kDiaCodes ...] (index 0 holds the zero of an absent entry),
// one byte per entry instead of the value.
constexpr uint32_t kDiaCodes = 16;
constexpr uint32_t kDiaMaxCoded = 32;  // the most diagonals of a DIA copy (kDiaMax)
template <typename T>
struct DiaDev {
  const T* __restrict__ val;
  const int32_t* __restrict__ off;
  uint32_t n_rows;
  uint32_t nd;
  const uint8_t* __restrict__ code = nullptr;
  const T* __restrict__ table = nullptr;
};

// One lane per row over the diagonals in offset order, i.e. the row's entries in column order, zero
// values skipped as the SELL kernel skips them: the same sums as a row of the SELL copy.  There are
// no column indices to load, and the x reads of a wavefront are 64 consecutive values per diagonal.
// Coded: the workgroup copies the value tables to LDS and each lane reads one byte per entry; the
// value it looks up is the stored one, so the sums are those of the uncoded copy bit for bit.
template <typename T, class Epi, bool NT = false, bool Coded = false>
__global__ __launch_bounds__(kBlock) void dia_kernel(DiaDev<T> A, const T* __restrict__ x, Epi epi,
                                                     typename Epi::acc_t* __restrict__ part, uint32_t pstride,
                                                     const uint32_t* __restrict__ stop, uint32_t stop_mask)
{
  using acc_t = typename Epi::acc_t;
  constexpr int KA = Epi::K > 0 ? Epi::K : 1;
  if (stop && (*stop & stop_mask)) return;
  const uint32_t tid = threadIdx.x;
  __shared__ T stab[Coded ? kDiaMaxCoded * kDiaCodes : 1];
  if constexpr (Coded) {
    for (uint32_t i = tid; i < A.nd * kDiaCodes; i += kBlock) stab[i] = A.table[i];
    __syncthreads();
  }
  // the value of diagonal d in this row: the stored one, or its code's table entry
  auto entry = [&](uint32_t d, uint32_t row) -> T {
    if constexpr (Coded)
      return stab[d * kDiaCodes + A.code[static_cast<size_t>(d) * A.n_rows + row]];
    else
      return ld_stream<NT>(A.val + static_cast<size_t>(d) * A.n_rows + row);
  };
  const uint32_t lw = xcd_logical_block();
  const uint32_t tiles = (A.n_rows + kBlock - 1) / kBlock;
  uint32_t t0, t1;
  chunk_range(tiles, lw, gridDim.x, t0, t1);
  acc_t acc[KA];
#pragma unroll
  for (int q = 0; q < KA; ++q) acc[q] = 0;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t row = t * kBlock + tid;
    if (row < A.n_rows) {
      T sum = 0;
      if constexpr (dia_grouped<Epi>::value || Coded) {
        // Groups of kDiaGroup diagonals: every load of a group is issued before the first sum (the
        // test for the diagonals past A.nd is uniform, a scalar branch), and x is read wherever the
        // column exists, whatever the value.  A counted loop over A.nd runs a 7-diagonal stencil
        // entirely in the compiler's unroll remainder, one diagonal per memory round trip.
        for (uint32_t d0 = 0; d0 < A.nd; d0 += kDiaGroup) {
          T v[kDiaGroup], xv[kDiaGroup];
#pragma unroll
          for (uint32_t j = 0; j < kDiaGroup; ++j) {
            v[j] = T(0);
            xv[j] = T(0);
            if (d0 + j < A.nd) {
              const uint32_t d = d0 + j;
              const uint32_t c = row + static_cast<uint32_t>(A.off[d]);  // out of range (also below 0): no entry
              v[j] = entry(d, row);
              if (c < A.n_rows) xv[j] = x[c];
            }
          }
#pragma unroll
          for (uint32_t j = 0; j < kDiaGroup; ++j)
            if (d0 + j < A.nd) sum += (v[j] > 0 || v[j] < 0) ? xv[j] * v[j] : T(0);
        }
      } else {
        // the counted loop: fewer loads in flight, measured faster under the BiCGStab epilogues
#pragma unroll 8
        for (uint32_t d = 0; d < A.nd; ++d) {
          const uint32_t c = row + static_cast<uint32_t>(A.off[d]);
          const T v = entry(d, row);
          const T xv = c < A.n_rows ? x[c] : T(0);
          sum += (v > 0 || v < 0) ? xv * v : T(0);
        }
      }
      epi(row, sum, acc);
    }
  }
  if constexpr (Epi::K > 0) block_store_partials<Epi::K, acc_t>(acc, part, pstride, lw);
}

// ELL (one slice of height ld, column-major) and HYB (ELL part plus a CSR part per row).
template <typename T>
struct EllDev {
  const uint32_t* __restrict__ coords;
  const T* __restrict__ val;
  uint32_t n_rows;
  uint32_t ld;     // internal_size1
  uint32_t width;  // internal_maxnnz / internal_ellnnz
  const uint32_t* __restrict__ csr_rows = nullptr;  // HYB
  const uint32_t* __restrict__ csr_cols = nullptr;
  const T* __restrict__ csr_val = nullptr;
};

// One lane per row; the lanes of a wavefront read 64 consecutive entries of each ELL column.
// Zero ELL entries are skipped (their x entry is not read), as the reference's host loop does
// (host_based/sparse_matrix_operations.hpp:1516-1524); the HYB CSR part follows in entry order
// (:1906-1912), so every row is summed in the reference's order.  HYB matrices with a plan for
// their CSR part take csr_kernel<..., Hyb = true> instead, which spreads long CSR parts over a
// wavefront.
template <typename T, bool Contig, class Epi, bool Hyb>
__global__ __launch_bounds__(kBlock) void ell_kernel(EllDev<T> A, const T* __restrict__ x, uint32_t xs, uint32_t xi,
                                                     Epi epi, typename Epi::acc_t* __restrict__ part,
                                                     uint32_t pstride, const uint32_t* __restrict__ stop,
                                                     uint32_t stop_mask)
{
  using acc_t = typename Epi::acc_t;
  constexpr int KA = Epi::K > 0 ? Epi::K : 1;
  if (stop && (*stop & stop_mask)) return;

  const uint32_t tid = threadIdx.x;
  const uint32_t lw = xcd_logical_block();
  const uint32_t tiles = (A.n_rows + kBlock - 1) / kBlock;
  uint32_t t0, t1;
  chunk_range(tiles, lw, gridDim.x, t0, t1);

  acc_t acc[KA];
#pragma unroll
  for (int q = 0; q < KA; ++q) acc[q] = 0;

  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t row = t * kBlock + tid;
    if (row < A.n_rows) {
      T sum = ell_row_sum<T, Contig>(A.coords, A.val, A.ld, A.width, row, x, xs, xi);
      if constexpr (Hyb) {
        const uint32_t e1 = A.csr_rows[row + 1];
        for (uint32_t e = A.csr_rows[row]; e < e1; ++e) sum += load_x<T, Contig>(x, A.csr_cols[e], xs, xi) * A.csr_val[e];
      }
      epi(row, sum, acc);
    }
  }

  if constexpr (Epi::K > 0) block_store_partials<Epi::K, acc_t>(acc, part, pstride, lw);
}

// Grid for kernels that take contiguous chunks per logical workgroup: a multiple of 8,
// at most `gmax`, at least 8.
inline uint32_t chunk_grid(uint32_t units, uint32_t gmax)
{
  uint32_t g = units < gmax ? units : gmax;
  g = (g + 7u) & ~7u;
  return g < 8u ? 8u : g;
}

// Tunables read once from the environment (VCLHIP_CSR_GMAX, VCLHIP_SELL_GMAX, VCLHIP_VEC_GMAX).
uint32_t csr_gmax();
// Nontemporal matrix loads for a matrix of `bytes`: VCLHIP_STREAM_LOADS = 0 never, 1 always; by
// default when the matrix is larger than the 256 MiB Infinity Cache.
bool stream_loads(size_t bytes);
// Nontemporal loads of a DIA copy's values in a solver whose working set (the values and four
// vectors) takes `working_set` bytes: VCLHIP_DIA_NT = 0 never, 1 always; by default when it is
// larger than the 256 MiB Infinity Cache (the values then do not push the vectors out of L2).
bool dia_stream_loads(size_t working_set);
uint32_t sell_gmax();
uint32_t vec_gmax();
uint32_t spmv_gmax();

}  // namespace vclhip
