// solver.hpp -- host-side pieces shared by the device-resident solver drivers: the matrix
// wrapper that launches a fused SpMV with a solver's epilogue, the solve stream and the
// stream-ordered workspace.
#pragma once

#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "spmv_kernels.hpp"

namespace vclhip {

// ------------------------------------------------------------------ matrix wrapper

enum class OpFormat { Csr, Sell, Coo, Ell, Hyb, Dia };

template <typename T>
struct DevOp {
  OpFormat fmt = OpFormat::Csr;
  bool sell = false;                   // fmt == Sell (the mixed solver's inner copy is SELL)
  const vclhip_csr* csr_desc = nullptr;  // the caller's CSR descriptor (row_info)
  CsrDev<T> csr{};                     // CSR, and COO through its row pointers (column stride 2)
  SellDev<T> sm{};
  EllDev<T> ell{};                     // ELL and HYB
  uint32_t n = 0;
  uint32_t grid = 8;

  bool nt = false;  // nontemporal matrix loads (matrix larger than the Infinity Cache)

  template <class Epi, bool Delta, bool NT>
  void launch_sell(hipStream_t s, const T* x, const Epi& epi, double* part, const uint32_t* stop,
                   uint32_t mask) const
  {
    hipLaunchKernelGGL((sell_kernel<T, true, Epi, Delta, NT>), dim3(grid), dim3(kBlock), 0, s, sm, x, 0u, 1u, epi,
                       part, kMaxPartials, stop, mask);
  }

  DiaDev<T> dia{};  // a DIA copy (the solvers' operator for stencils, the mixed solver's inner copy)
  bool dia_nt = false;  // its values read nontemporally (dia_stream_loads)
  HybDev<T> hyb{};  // HYB with a plan for its CSR part
  bool hyb_plan = false;

  template <class Epi, int CS>
  void launch_csr(hipStream_t s, const T* x, const Epi& epi, double* part, const uint32_t* stop, uint32_t mask) const
  {
    if (nt)
      hipLaunchKernelGGL((csr_kernel<T, true, Epi, true, CS>), dim3(grid), dim3(kBlock), 0, s, csr, x, 0u, 1u, epi,
                         part, kMaxPartials, stop, mask);
    else
      hipLaunchKernelGGL((csr_kernel<T, true, Epi, false, CS>), dim3(grid), dim3(kBlock), 0, s, csr, x, 0u, 1u, epi,
                         part, kMaxPartials, stop, mask);
  }

  template <class Epi>
  void launch(hipStream_t s, const T* x, const Epi& epi, double* part, const uint32_t* stop, uint32_t mask) const
  {
    switch (fmt) {
      case OpFormat::Sell:
        if (sm.dcol && nt)
          launch_sell<Epi, true, true>(s, x, epi, part, stop, mask);
        else if (sm.dcol)
          launch_sell<Epi, true, false>(s, x, epi, part, stop, mask);
        else if (nt)
          launch_sell<Epi, false, true>(s, x, epi, part, stop, mask);
        else
          launch_sell<Epi, false, false>(s, x, epi, part, stop, mask);
        break;
      case OpFormat::Coo:
        launch_csr<Epi, 2>(s, x, epi, part, stop, mask);
        break;
      case OpFormat::Ell:
        hipLaunchKernelGGL((ell_kernel<T, true, Epi, false>), dim3(grid), dim3(kBlock), 0, s, ell, x, 0u, 1u, epi,
                           part, kMaxPartials, stop, mask);
        break;
      case OpFormat::Dia:
        if (dia.code)
          hipLaunchKernelGGL((dia_kernel<T, Epi, false, true>), dim3(grid), dim3(kBlock), 0, s, dia, x, epi, part,
                             kMaxPartials, stop, mask);
        else if (dia_nt)
          hipLaunchKernelGGL((dia_kernel<T, Epi, true>), dim3(grid), dim3(kBlock), 0, s, dia, x, epi, part, kMaxPartials,
                             stop, mask);
        else
          hipLaunchKernelGGL((dia_kernel<T, Epi>), dim3(grid), dim3(kBlock), 0, s, dia, x, epi, part, kMaxPartials, stop,
                             mask);
        break;
      case OpFormat::Hyb:
        if (hyb_plan)
          hipLaunchKernelGGL((csr_kernel<T, true, Epi, false, 1, true>), dim3(grid), dim3(kBlock), 0, s, hyb, x, 0u,
                             1u, epi, part, kMaxPartials, stop, mask);
        else
          hipLaunchKernelGGL((ell_kernel<T, true, Epi, true>), dim3(grid), dim3(kBlock), 0, s, ell, x, 0u, 1u, epi,
                             part, kMaxPartials, stop, mask);
        break;
      default:
        launch_csr<Epi, 1>(s, x, epi, part, stop, mask);
    }
  }
};

template <typename T>
static int make_op(const vclhip_csr* A, DevOp<T>& op)
{
  if (!A || !A->row_ptr || !A->blocks || !A->col_idx || !A->values) return invalid("CSR matrix arrays");
  if (A->n_rows == 0 || A->n_rows != A->n_cols) return invalid("solvers need a nonempty square matrix");
  if (A->n_rows >= kPlanRowLimit || A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("CSR plan");
  op.fmt = OpFormat::Csr;
  op.sell = false;
  op.csr_desc = A;
  op.csr.row_ptr = A->row_ptr;
  op.csr.col = A->col_idx;
  op.csr.val = static_cast<const T*>(A->values);
  op.csr.blocks = A->blocks;
  op.csr.num_blocks = A->num_blocks;
  op.csr.n_rows = A->n_rows;
  op.n = A->n_rows;
  op.grid = chunk_grid(A->num_blocks, csr_gmax());
  op.nt = stream_loads(static_cast<size_t>(A->nnz) * (sizeof(T) + 4));
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_op(const vclhip_sell* A, DevOp<T>& op)
{
  if (!A || !A->columns_per_block || !A->block_start || !A->elements)
    return invalid("SELL matrix arrays");
  if (A->n_rows == 0 || A->n_rows != A->n_cols || A->C == 0) return invalid("solvers need a nonempty square matrix");
  if (A->num_slices != (A->n_rows + A->C - 1) / A->C) return invalid("num_slices");
  op.fmt = OpFormat::Sell;
  op.sell = true;
  op.sm.cpb = A->columns_per_block;
  op.sm.col = A->column_indices;
  op.sm.start = A->block_start;
  op.sm.val = static_cast<const T*>(A->elements);
  op.sm.n_rows = A->n_rows;
  op.sm.C = A->C;
  op.n = A->n_rows;
  op.grid = chunk_grid((A->n_rows + kBlock - 1) / kBlock, sell_gmax());
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_op(const vclhip_coo* A, DevOp<T>& op)
{
  if (!A || !A->row_ptr || !A->blocks || (A->nnz && (!A->coords || !A->elements))) return invalid("COO matrix arrays");
  if (A->n_rows == 0 || A->n_rows != A->n_cols) return invalid("solvers need a nonempty square matrix");
  if (A->n_rows >= kPlanRowLimit || A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("COO plan");
  op.fmt = OpFormat::Coo;
  op.sell = false;
  op.csr.row_ptr = A->row_ptr;
  op.csr.col = A->coords ? A->coords + 1 : nullptr;
  op.csr.val = static_cast<const T*>(A->elements);
  op.csr.blocks = A->blocks;
  op.csr.num_blocks = A->num_blocks;
  op.csr.n_rows = A->n_rows;
  op.n = A->n_rows;
  op.grid = chunk_grid(A->num_blocks, csr_gmax());
  op.nt = stream_loads(static_cast<size_t>(A->nnz) * (sizeof(T) + 8));
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_ell_op(uint32_t n_rows, uint32_t n_cols, uint32_t ld, uint32_t width, const uint32_t* coords,
                       const void* elements, DevOp<T>& op)
{
  if (n_rows == 0 || n_rows != n_cols) return invalid("solvers need a nonempty square matrix");
  if (ld < n_rows || (width && (!coords || !elements))) return invalid("ELL arrays");
  op.sell = false;
  op.ell.coords = coords;
  op.ell.val = static_cast<const T*>(elements);
  op.ell.n_rows = n_rows;
  op.ell.ld = ld;
  op.ell.width = width;
  op.n = n_rows;
  op.grid = chunk_grid((n_rows + kBlock - 1) / kBlock, sell_gmax());
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_op(const vclhip_ell* A, DevOp<T>& op)
{
  if (!A) return invalid("ELL matrix is NULL");
  VCLHIP_TRY(make_ell_op(A->n_rows, A->n_cols, A->internal_size1, A->internal_maxnnz, A->coords, A->elements, op));
  op.fmt = OpFormat::Ell;
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_op(const vclhip_hyb* A, DevOp<T>& op)
{
  if (!A || !A->csr_rows || !A->csr_cols || !A->csr_elements) return invalid("HYB matrix arrays");
  VCLHIP_TRY(make_ell_op(A->n_rows, A->n_cols, A->internal_size1, A->internal_ellnnz, A->ell_coords,
                         A->ell_elements, op));
  op.fmt = OpFormat::Hyb;
  op.ell.csr_rows = A->csr_rows;
  op.ell.csr_cols = A->csr_cols;
  op.ell.csr_val = static_cast<const T*>(A->csr_elements);
  op.hyb_plan = A->blocks != nullptr;
  if (op.hyb_plan) {
    if (A->n_rows >= kPlanRowLimit || A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("HYB plan");
    op.hyb.row_ptr = A->csr_rows;
    op.hyb.col = A->csr_cols;
    op.hyb.val = static_cast<const T*>(A->csr_elements);
    op.hyb.blocks = A->blocks;
    op.hyb.num_blocks = A->num_blocks;
    op.hyb.n_rows = A->n_rows;
    op.hyb.ell_coords = A->ell_coords;
    op.hyb.ell_val = static_cast<const T*>(A->ell_elements);
    op.hyb.ell_ld = A->internal_size1;
    op.hyb.ell_width = A->internal_ellnnz;
    op.grid = chunk_grid(A->num_blocks, csr_gmax());
  }
  return VCLHIP_SUCCESS;
}

template <typename T>
static int make_op(const vclhip_matrix* A, DevOp<T>& op)
{
  if (!A || !A->desc) return invalid("matrix is NULL");
  switch (A->format) {
    case VCLHIP_FORMAT_CSR: return make_op(static_cast<const vclhip_csr*>(A->desc), op);
    case VCLHIP_FORMAT_SELL: return make_op(static_cast<const vclhip_sell*>(A->desc), op);
    case VCLHIP_FORMAT_COO: return make_op(static_cast<const vclhip_coo*>(A->desc), op);
    case VCLHIP_FORMAT_ELL: return make_op(static_cast<const vclhip_ell*>(A->desc), op);
    case VCLHIP_FORMAT_HYB: return make_op(static_cast<const vclhip_hyb*>(A->desc), op);
    default: return invalid("unknown matrix format");
  }
}

inline uint32_t vec_grid_n(uint32_t n) { return chunk_grid((n + 4 * kBlock - 1) / (4 * kBlock), vec_gmax()); }

// Vector grid of the device-resident solvers: vec_grid_n's chunks of 1024 elements, but where that
// gives fewer workgroups than CUs, chunks down to 256 elements over up to one workgroup per CU, so
// that a small system's vector updates use the whole device and not a quarter of it.
inline uint32_t solver_vec_grid(uint32_t n)
{
  uint32_t units = (n + 4 * kBlock - 1) / (4 * kBlock);
  const uint32_t cus = static_cast<uint32_t>(device_cu_count());
  if (units < cus) units = std::min((n + kBlock - 1) / kBlock, cus);
  return chunk_grid(units, vec_gmax());
}

// ------------------------------------------------------------------ solve stream

// Graph capture needs a real stream.  When the caller passes the null stream and asks for graphs,
// the solve runs on a private non-blocking stream, ordered after all earlier device work.
class SolveStream {
 public:
  SolveStream(void* user, bool need_capture)
  {
    s_ = static_cast<hipStream_t>(user);
    if (!s_ && need_capture) {
      status_ = hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (!status_) status_ = hip_status(hipStreamCreateWithFlags(&own_, hipStreamNonBlocking), "hipStreamCreate");
      s_ = own_;
    }
  }
  ~SolveStream()
  {
    if (own_) {
      (void)hipStreamSynchronize(own_);
      (void)hipStreamDestroy(own_);
    }
  }
  hipStream_t get() const { return s_; }
  int status() const { return status_; }

 private:
  hipStream_t s_ = nullptr;
  hipStream_t own_ = nullptr;
  int status_ = VCLHIP_SUCCESS;
};

// ------------------------------------------------------------------ device workspace

// Stream-ordered allocations released when the object goes out of scope; or, made over an arena
// (a caller's device buffer), 256-byte aligned pieces of that buffer and no allocation at all.
class Workspace {
 public:
  explicit Workspace(hipStream_t s) : s_(s) {}
  Workspace(hipStream_t s, void* arena, size_t bytes) : s_(s), arena_(static_cast<char*>(arena)), cap_(bytes) {}
  ~Workspace()
  {
    for (void* p : ptrs_) (void)hipFreeAsync(p, s_);
  }
  template <typename U>
  int alloc(U** p, size_t count)
  {
    if (arena_) {
      if (!take(p, count)) {
        set_error("scratch memory: %zu more bytes needed, %zu of %zu left", count * sizeof(U), cap_ - used_, cap_);
        return VCLHIP_ERROR_TOO_LARGE;
      }
      return VCLHIP_SUCCESS;
    }
    void* v = nullptr;
    VCLHIP_HIP(hipMallocAsync(&v, count * sizeof(U) + 256, s_));
    ptrs_.push_back(v);
    *p = static_cast<U*>(v);
    return VCLHIP_SUCCESS;
  }
  // For optional buffers: false (and the HIP error cleared) when the allocation fails.
  template <typename U>
  bool try_alloc(U** p, size_t count)
  {
    if (arena_) return take(p, count);
    void* v = nullptr;
    if (hipMallocAsync(&v, count * sizeof(U) + 256, s_) != hipSuccess || !v) {
      (void)hipGetLastError();
      *p = nullptr;
      return false;
    }
    ptrs_.push_back(v);
    *p = static_cast<U*>(v);
    return true;
  }
  // Bytes still free in the arena (no limit without one).
  size_t arena_left() const { return arena_ ? cap_ - used_ : ~size_t(0); }
  // The arena bytes that alloc(count elements of `elem` bytes) takes: like a device allocation here,
  // 256 bytes beyond the request, rounded up to keep the next piece 256-byte aligned.
  static size_t arena_bytes(size_t elem, size_t count) { return (elem * count + 256 + 255) & ~size_t(255); }

 private:
  template <typename U>
  bool take(U** p, size_t count)
  {
    const size_t bytes = arena_bytes(sizeof(U), count);
    if (bytes > cap_ - used_) {
      *p = nullptr;
      return false;
    }
    *p = reinterpret_cast<U*>(arena_ + used_);
    used_ += bytes;
    return true;
  }

  hipStream_t s_;
  std::vector<void*> ptrs_;
  char* arena_ = nullptr;  // 256-byte aligned
  size_t cap_ = 0, used_ = 0;
};

}  // namespace vclhip
