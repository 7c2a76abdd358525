// dia.hpp -- DIA copies of CSR matrices with few diagonals, built on the device at solve time.
//
// A CSR matrix whose entries lie on at most kDiaMax diagonals (stencil operators) is stored as one
// value per row and diagonal, offsets ascending; dia_kernel (spmv_kernels.hpp) then needs no column
// indices and issues every load of a row at once.  With at most kDiaMax diagonals no row has more
// than kDiaMax <= kShortRow entries, so the CSR kernel sums every row in one lane in column order,
// and so does dia_kernel: the products are the same bit for bit.  The inner products fused into
// the product are summed per 256-row tile here and per row block of the CSR plan there; the two
// groupings coincide when no row has more than 8 entries (every block is then a 256-row tile, as
// for the 7-point stencil) and a solve is the same bit for bit; with longer rows it agrees to
// rounding.
// Used for the mixed-precision solver's FP32 inner copy and, unless the caller asks for the
// matrix as given, for the operator of the CG, BiCGStab and GMRES solvers.
#pragma once

#include <vector>

#include "internal.hpp"
#include "solver.hpp"

namespace vclhip {

constexpr uint32_t kDiaMax = 32;    // diagonals of a DIA copy at most
constexpr uint32_t kDiaHash = 128;  // LDS hash slots per workgroup for the offsets of its rows
static_assert(kDiaMax <= kShortRow, "a DIA row must be one the CSR kernel sums in one lane");
static_assert(kDiaMax <= kDiaMaxCoded, "dia_kernel's LDS value tables hold kDiaMaxCoded diagonals");
constexpr uint32_t kDiaDictSlots = kDiaCodes - 1;   // distinct values per diagonal (code 0 is the zero)
constexpr uint32_t kDiaCodedMinRows = 1u << 18;      // a value dictionary only where bytes bound the product
constexpr uint32_t kDiaDictGroups = 256;             // workgroups per diagonal collecting its values

// Marks every offset col - row of the matrix in a bitmap (bit col - row + n).  A workgroup first
// collects the distinct offsets of its 256 rows in an LDS hash table, so that the global bitmap
// sees one atomic OR per distinct offset and workgroup; a workgroup whose rows hold more distinct
// offsets than the table, or a column outside the square matrix, raises *overflow (the matrix is
// then not a DIA candidate).
template <typename I>
__global__ __launch_bounds__(kBlock) void dia_offsets_kernel(const I* __restrict__ rp, const I* __restrict__ col,
                                                             uint32_t n, unsigned long long* __restrict__ bits,
                                                             uint32_t* __restrict__ overflow)
{
  __shared__ uint32_t keys[kDiaHash];
  __shared__ volatile uint32_t full;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kDiaHash; i += kBlock) keys[i] = 0;
  if (tid == 0) full = 0;
  __syncthreads();
  const uint32_t row = blockIdx.x * kBlock + tid;
  if (row < n) {
    const uint32_t e1 = rp[row + 1];
    for (uint32_t e = rp[row]; e < e1 && !full; ++e) {  // once the table is full the answer is known
      const uint32_t c = col[e];
      if (c >= n) {  // not a square-matrix column: no DIA copy
        full = 1;
        break;
      }
      const uint32_t key = c + n - row;  // in [1, 2n - 1]; 0 marks an empty slot
      uint32_t h = (key * 2654435761u) >> 25;
      uint32_t probe = 0;
      // kDiaMax + 1 occupied slots in a row mean more than kDiaMax distinct offsets: not a candidate
      for (; probe <= kDiaMax; ++probe) {
        const uint32_t prev = atomicCAS(&keys[h], 0u, key);
        if (prev == 0u || prev == key) break;
        h = (h + 1) & (kDiaHash - 1);
      }
      if (probe > kDiaMax) full = 1;
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < kDiaHash; i += kBlock) {
    const uint32_t k = keys[i];
    if (k) atomicOr(&bits[k >> 6], 1ull << (k & 63));
  }
  if (tid == 0 && full) *overflow = 1u;
}

// val[d * n + row] = A(row, row + off[d]) converted to T (val zeroed before); off ascending,
// nd <= kDiaMax, and every offset of the matrix is in off.  A row whose columns do not strictly
// increase (unsorted, or an entry given twice) raises *refuse: its sum in column order would not be
// the CSR kernel's, and a repeated entry would be lost.  With `exact`, so does a stored value that
// the DIA kernel would skip (a zero of either sign, or NaN): the CSR kernel multiplies every stored
// entry, so 0 * Inf or a NaN value would reach its sum and not the DIA one.
template <typename S, typename T>
__global__ __launch_bounds__(kBlock) void dia_fill_kernel(const uint32_t* __restrict__ rp,
                                                          const uint32_t* __restrict__ col, const S* __restrict__ a,
                                                          uint32_t n, const int32_t* __restrict__ off, uint32_t nd,
                                                          int exact, T* __restrict__ val, uint32_t* __restrict__ refuse)
{
  __shared__ int32_t so[kDiaMax];
  if (threadIdx.x < nd) so[threadIdx.x] = off[threadIdx.x];
  __syncthreads();
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  const uint32_t e0 = rp[row], e1 = rp[row + 1];
  for (uint32_t e = e0; e < e1; ++e) {
    if (e > e0 && col[e] <= col[e - 1]) *refuse = 1u;
    const S v = a[e];
    if (exact && !(v > S(0) || v < S(0))) *refuse = 1u;
    const int32_t o = static_cast<int32_t>(col[e] - row);
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (so[mid] <= o)
        lo = mid;
      else
        hi = mid;
    }
    val[static_cast<size_t>(lo) * n + row] = static_cast<T>(v);
  }
}

// ---- value dictionaries: one table of at most kDiaDictSlots distinct values per diagonal.
// A value is keyed by its bit pattern (widened to 64 bits); key 0 (+0.0) is an absent entry and
// has code 0.  A set lives in LDS and grows by compare-and-swap: a slot is written once, so a lane
// that finds another key in a slot moves on to the next, and no key is inserted twice.

template <typename T>
__device__ __forceinline__ unsigned long long dia_key(T v)
{
  if constexpr (sizeof(T) == 8)
    return static_cast<unsigned long long>(__double_as_longlong(v));
  else
    return static_cast<unsigned long long>(__float_as_uint(v));
}

template <typename T>
__device__ __forceinline__ T dia_value(unsigned long long k)
{
  if constexpr (sizeof(T) == 8)
    return __longlong_as_double(static_cast<long long>(k));
  else
    return __uint_as_float(static_cast<uint32_t>(k));
}

// Inserts key (nonzero) into slots[0 .. kDiaDictSlots); false when they are full of other keys.
__device__ __forceinline__ bool dia_set_insert(unsigned long long* slots, unsigned long long key)
{
  for (uint32_t j = 0; j < kDiaDictSlots; ++j) {
    const unsigned long long cur = __hip_atomic_load(&slots[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == key) return true;
    if (cur == 0ull) {
      const unsigned long long prev = atomicCAS(&slots[j], 0ull, key);
      if (prev == 0ull || prev == key) return true;
    }
  }
  return false;
}

// Pass 1: workgroup (g, d) collects the distinct values of diagonal d over rows g*kBlock + tid,
// stepping by gridDim.x*kBlock, into sets[(d * gridDim.x + g) * kDiaDictSlots ...]; more than
// kDiaDictSlots of them raise *overflow.
template <typename T>
__global__ __launch_bounds__(kBlock) void dia_dict_collect_kernel(const T* __restrict__ val, uint32_t n,
                                                                  unsigned long long* __restrict__ sets,
                                                                  uint32_t* __restrict__ overflow)
{
  __shared__ unsigned long long slots[kDiaDictSlots];
  __shared__ uint32_t full;
  const uint32_t tid = threadIdx.x, d = blockIdx.y;
  if (tid < kDiaDictSlots) slots[tid] = 0ull;
  if (tid == 0) full = 0u;
  __syncthreads();
  const T* v = val + static_cast<size_t>(d) * n;
  for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
    const uint32_t row = base + tid;
    if (row < n) {
      const unsigned long long key = dia_key(v[row]);
      if (key != 0ull && !dia_set_insert(slots, key)) full = 1u;
    }
    __syncthreads();
    const bool stop_now = full != 0u;  // read by every wave before any wave writes it again
    __syncthreads();
    if (stop_now) break;
  }
  if (tid < kDiaDictSlots) sets[(static_cast<size_t>(d) * gridDim.x + blockIdx.x) * kDiaDictSlots + tid] = slots[tid];
  if (tid == 0 && full) *overflow = 1u;
}

// Pass 2: workgroup d merges the groups' sets of diagonal d into its table (entry 0 the zero of an
// absent entry, unused entries zero) and keeps the keys for pass 3.
template <typename T>
__global__ __launch_bounds__(kBlock) void dia_dict_merge_kernel(const unsigned long long* __restrict__ sets,
                                                                uint32_t groups, T* __restrict__ table,
                                                                unsigned long long* __restrict__ keys,
                                                                uint32_t* __restrict__ overflow)
{
  __shared__ unsigned long long slots[kDiaDictSlots];
  __shared__ uint32_t full;
  const uint32_t tid = threadIdx.x, d = blockIdx.x;
  if (tid < kDiaDictSlots) slots[tid] = 0ull;
  if (tid == 0) full = 0u;
  __syncthreads();
  const unsigned long long* in = sets + static_cast<size_t>(d) * groups * kDiaDictSlots;
  for (uint32_t i = tid; i < groups * kDiaDictSlots; i += kBlock) {
    const unsigned long long key = in[i];
    if (key != 0ull && !dia_set_insert(slots, key)) full = 1u;
  }
  __syncthreads();
  if (tid < kDiaCodes) {
    const unsigned long long k = tid == 0 ? 0ull : slots[tid - 1];
    table[d * kDiaCodes + tid] = dia_value<T>(k);
    keys[d * kDiaCodes + tid] = k;
  }
  if (tid == 0 && full) *overflow = 1u;
}

// Pass 3: code[d * n + row] = the index of the row's value in its diagonal's table.
template <typename T>
__global__ __launch_bounds__(kBlock) void dia_encode_kernel(const T* __restrict__ val, uint32_t n, uint32_t nd,
                                                            const unsigned long long* __restrict__ keys,
                                                            uint8_t* __restrict__ code)
{
  __shared__ unsigned long long sk[kDiaMax * kDiaCodes];
  for (uint32_t i = threadIdx.x; i < nd * kDiaCodes; i += kBlock) sk[i] = keys[i];
  __syncthreads();
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  for (uint32_t d = 0; d < nd; ++d) {
    const unsigned long long key = dia_key(val[static_cast<size_t>(d) * n + row]);
    uint32_t c = 0;
    if (key != 0ull)
      for (uint32_t j = 1; j < kDiaCodes; ++j)
        if (sk[d * kDiaCodes + j] == key) c = j;
    code[static_cast<size_t>(d) * n + row] = static_cast<uint8_t>(c);
  }
}

// Whether DIA copies of kDiaCodedMinRows rows or more get value dictionaries (VCLHIP_DIA_CODED=0
// turns them off for the whole process).
bool dia_codes_enabled();

// Bytes a value dictionary adds behind a copy's values (tables first, 256-byte aligned, then codes).
template <typename T>
inline size_t dia_code_bytes(uint32_t n, uint32_t nd)
{
  return (static_cast<size_t>(nd) * kDiaCodes * sizeof(T) + 255) / 256 * 256 + static_cast<size_t>(nd) * n;
}

// Builds the value dictionary of `dia` into `extra` (dia_code_bytes<T> bytes); *coded tells
// whether every diagonal fit.  Temporaries come from `ws`; without them there is no dictionary.
template <typename T>
static int build_dia_codes(DiaDev<T>& dia, void* extra, Workspace& ws, uint32_t* flag, hipStream_t s, bool* coded)
{
  *coded = false;
  const uint32_t n = dia.n_rows, nd = dia.nd;
  const uint32_t groups = std::min<uint32_t>(kDiaDictGroups, (n + kBlock - 1) / kBlock);
  unsigned long long *sets, *keys;
  if (!ws.try_alloc(&sets, static_cast<size_t>(nd) * groups * kDiaDictSlots) ||
      !ws.try_alloc(&keys, static_cast<size_t>(nd) * kDiaCodes))
    return VCLHIP_SUCCESS;
  T* table = static_cast<T*>(extra);
  uint8_t* code = static_cast<uint8_t*>(extra) + (static_cast<size_t>(nd) * kDiaCodes * sizeof(T) + 255) / 256 * 256;
  VCLHIP_TRY(vclhip_fill_zero(flag, sizeof(uint32_t), s));
  hipLaunchKernelGGL(dia_dict_collect_kernel<T>, dim3(groups, nd), dim3(kBlock), 0, s, dia.val, n, sets, flag);
  VCLHIP_TRY(check_launch("dia_dict_collect_kernel"));
  hipLaunchKernelGGL(dia_dict_merge_kernel<T>, dim3(nd), dim3(kBlock), 0, s, sets, groups, table, keys, flag);
  VCLHIP_TRY(check_launch("dia_dict_merge_kernel"));
  uint32_t hover = 0;
  VCLHIP_HIP(hipMemcpyAsync(&hover, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  if (hover) return VCLHIP_SUCCESS;
  hipLaunchKernelGGL(dia_encode_kernel<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, dia.val, n, nd, keys,
                     code);
  VCLHIP_TRY(check_launch("dia_encode_kernel"));
  dia.code = code;
  dia.table = table;
  *coded = true;
  return VCLHIP_SUCCESS;
}

// Device memory of at most `bytes` taken for an optional copy: a failure leaves *p null and clears
// the HIP error, so that the caller keeps the matrix as given.
struct DiaArrays {
  int32_t* off = nullptr;
  void* val = nullptr;
};

// Builds the DIA copy (values converted from S to T) of a square CSR matrix when it has at most
// kDiaMax distinct diagonals, they store at most `max_fill` times nnz values and the values take at
// most max_bytes; *built tells whether it was made.  Temporaries come from `ws`; the copy's arrays
// from `alloc_out(bytes_off, bytes_val, &arrays)`, which returns false when it cannot provide them
// (then no copy is made).  Every failure of this optional step other than a launch error leaves
// *built false.
template <typename S, typename T, class OutAlloc>
static int build_dia(const vclhip_csr* A, Workspace& ws, hipStream_t s, double max_fill, size_t max_bytes, bool exact,
                     OutAlloc&& alloc_out, DiaDev<T>* out, bool* built)
{
  *built = false;
  const uint32_t n = A->n_rows;
  if (n == 0 || A->n_cols != n || n >= (1u << 30) || !A->row_ptr || !A->col_idx || !A->values)
    return VCLHIP_SUCCESS;
  if (static_cast<uint64_t>(A->nnz) > static_cast<uint64_t>(kDiaMax) * n) return VCLHIP_SUCCESS;
  const size_t words = (2ull * n + 64) / 64;
  unsigned long long* bits;
  uint32_t* overflow;
  if (!ws.try_alloc(&bits, words) || !ws.try_alloc(&overflow, 1)) return VCLHIP_SUCCESS;
  const uint32_t g = (n + kBlock - 1) / kBlock;
  VCLHIP_TRY(vclhip_fill_zero(bits, words * sizeof(unsigned long long), s));
  VCLHIP_TRY(vclhip_fill_zero(overflow, sizeof(uint32_t), s));
  hipLaunchKernelGGL(dia_offsets_kernel<uint32_t>, dim3(g), dim3(kBlock), 0, s, A->row_ptr, A->col_idx, n, bits,
                     overflow);
  VCLHIP_TRY(check_launch("dia_offsets_kernel"));
  std::vector<unsigned long long> hb(words);
  uint32_t hover = 0;
  VCLHIP_HIP(hipMemcpyAsync(hb.data(), bits, words * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipMemcpyAsync(&hover, overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  if (hover) return VCLHIP_SUCCESS;
  std::vector<int32_t> offs;
  for (size_t w = 0; w < words; ++w)
    for (unsigned long long b = hb[w]; b; b &= b - 1) {
      if (offs.size() == kDiaMax) return VCLHIP_SUCCESS;
      const uint64_t k = w * 64 + static_cast<uint64_t>(__builtin_ctzll(b));
      offs.push_back(static_cast<int32_t>(static_cast<int64_t>(k) - n));
    }
  const uint32_t nd = static_cast<uint32_t>(offs.size());
  if (nd == 0 || static_cast<double>(nd) * n > max_fill * static_cast<double>(A->nnz)) return VCLHIP_SUCCESS;
  const size_t vbytes = static_cast<size_t>(nd) * n * sizeof(T);
  if (vbytes > max_bytes) return VCLHIP_SUCCESS;
  // room for a value dictionary behind the values (used when every diagonal's values fit one)
  const size_t vpad = (vbytes + 255) / 256 * 256;
  bool try_codes = n >= kDiaCodedMinRows && dia_codes_enabled() && vpad + dia_code_bytes<T>(n, nd) <= max_bytes;
  DiaArrays arr;
  if (!alloc_out(nd * sizeof(int32_t), try_codes ? vpad + dia_code_bytes<T>(n, nd) : vbytes, &arr)) {
    // without room for a dictionary the values alone still make the copy
    if (!try_codes || !alloc_out(nd * sizeof(int32_t), vbytes, &arr)) return VCLHIP_SUCCESS;
    try_codes = false;
  }
  T* val = static_cast<T*>(arr.val);
  VCLHIP_HIP(hipMemcpyAsync(arr.off, offs.data(), nd * sizeof(int32_t), hipMemcpyHostToDevice, s));
  VCLHIP_TRY(vclhip_fill_zero(val, vbytes, s));
  VCLHIP_TRY(vclhip_fill_zero(overflow, sizeof(uint32_t), s));  // reused as the refusal flag
  hipLaunchKernelGGL((dia_fill_kernel<S, T>), dim3(g), dim3(kBlock), 0, s, A->row_ptr, A->col_idx,
                     static_cast<const S*>(A->values), n, arr.off, nd, exact ? 1 : 0, val, overflow);
  VCLHIP_TRY(check_launch("dia_fill_kernel"));
  VCLHIP_HIP(hipMemcpyAsync(&hover, overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));  // offs lives on the host stack
  if (hover) return VCLHIP_SUCCESS;
  out->val = val;
  out->off = arr.off;
  out->n_rows = n;
  out->nd = nd;
  out->code = nullptr;
  out->table = nullptr;
  if (try_codes) {
    bool coded = false;
    VCLHIP_TRY(build_dia_codes<T>(*out, static_cast<char*>(arr.val) + vpad, ws, overflow, s, &coded));
  }
  *built = true;
  return VCLHIP_SUCCESS;
}

// The copy's arrays as workspace of one solve, released with it.
struct WorkspaceDiaAlloc {
  Workspace& ws;
  bool operator()(size_t off_bytes, size_t val_bytes, DiaArrays* a) const
  {
    char* v = nullptr;
    if (!ws.try_alloc(&a->off, off_bytes / sizeof(int32_t)) || !ws.try_alloc(&v, val_bytes)) return false;
    a->val = v;
    return true;
  }
};

// A quarter of the free device memory: the most an optional copy may take.
inline size_t dia_memory_cap()
{
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return free_b / 4;
}

}  // namespace vclhip

// A DIA copy kept across solves (vclhip_dia_create_*): the outcome of the detection for one CSR
// matrix and, when it is a candidate, the copy in device memory owned by this object.
struct vclhip_dia {
  int single;  // value type: 0 double, 1 float
  vclhip_csr matrix;  // the descriptor it was built from (arrays compared at use)
  bool built;
  vclhip::DiaDev<double> d64;
  vclhip::DiaDev<float> d32;
  void* val;
  int32_t* off;
};

namespace vclhip {

// The solvers' operator: a CSR matrix with few diagonals is replaced by a DIA copy (same products,
// fewer bytes each) unless params->operator_format asks for the matrix as given, or the copy
// would take more than a quarter of the free device memory.  params->dia, when given, is the
// detection and copy made once by vclhip_dia_create_* for this matrix; otherwise the copy is built
// in the solve's workspace.  *used_dia is the operator the solve runs on (VCLHIP_USED_*).
template <typename T>
static int dia_operator(DevOp<T>& op, Workspace& ws, hipStream_t s, const vclhip_solver_params* P, int* used_dia)
{
  *used_dia = 0;
  if (!P) return VCLHIP_SUCCESS;
  if (P->operator_format != VCLHIP_OPERATOR_AUTO && P->operator_format != VCLHIP_OPERATOR_AS_GIVEN &&
      P->operator_format != VCLHIP_OPERATOR_DIA_UNCODED)
    return invalid("operator_format must be VCLHIP_OPERATOR_AUTO, VCLHIP_OPERATOR_AS_GIVEN or "
                   "VCLHIP_OPERATOR_DIA_UNCODED");
  if (P->operator_format == VCLHIP_OPERATOR_AS_GIVEN || op.fmt != OpFormat::Csr || !op.csr_desc)
    return VCLHIP_SUCCESS;
  DiaDev<T> dia{};
  bool built = false;
  if (P->dia) {
    const vclhip_dia* c = P->dia;
    const vclhip_csr* A = op.csr_desc;
    if (c->single != (sizeof(T) == 4 ? 1 : 0)) return invalid("params->dia was built for the other value type");
    if (c->matrix.n_rows != A->n_rows || c->matrix.nnz != A->nnz || c->matrix.row_ptr != A->row_ptr ||
        c->matrix.col_idx != A->col_idx || c->matrix.values != A->values)
      return invalid("params->dia was built for another matrix");
    if (!c->built) return VCLHIP_SUCCESS;
    if constexpr (sizeof(T) == 4)
      dia = c->d32;
    else
      dia = c->d64;
    built = true;
  } else {
    VCLHIP_TRY((build_dia<T, T>(op.csr_desc, ws, s, 1.25, dia_memory_cap(), true, WorkspaceDiaAlloc{ws}, &dia,
                                &built)));
  }
  if (!built) return VCLHIP_SUCCESS;
  if (P->operator_format == VCLHIP_OPERATOR_DIA_UNCODED) {  // the values as stored, no dictionary
    dia.code = nullptr;
    dia.table = nullptr;
  }
  op.fmt = OpFormat::Dia;
  op.dia_nt = dia_stream_loads((static_cast<size_t>(dia.nd) + 4) * dia.n_rows * sizeof(T));
  op.sell = false;
  op.dia = dia;
  op.grid = chunk_grid((op.n + kBlock - 1) / kBlock, sell_gmax());
  *used_dia = dia.code ? VCLHIP_USED_DIA_CODED : VCLHIP_USED_DIA;
  return VCLHIP_SUCCESS;
}

}  // namespace vclhip
