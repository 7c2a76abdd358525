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

template <typename T>
struct DevOp {
  bool sell = false;
  CsrDev<T> csr{};
  SellDev<T> sm{};
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

  template <class Epi>
  void launch(hipStream_t s, const T* x, const Epi& epi, double* part, const uint32_t* stop, uint32_t mask) const
  {
    if (sell) {
      if (sm.dcol && nt)
        launch_sell<Epi, true, true>(s, x, epi, part, stop, mask);
      else if (sm.dcol)
        launch_sell<Epi, true, false>(s, x, epi, part, stop, mask);
      else if (nt)
        launch_sell<Epi, false, true>(s, x, epi, part, stop, mask);
      else
        launch_sell<Epi, false, false>(s, x, epi, part, stop, mask);
    } else if (nt)
      hipLaunchKernelGGL((csr_kernel<T, true, Epi, true>), dim3(grid), dim3(kBlock), 0, s, csr, x, 0u, 1u, epi, part,
                         kMaxPartials, stop, mask);
    else
      hipLaunchKernelGGL((csr_kernel<T, true, Epi>), dim3(grid), dim3(kBlock), 0, s, csr, x, 0u, 1u, epi, part,
                         kMaxPartials, stop, mask);
  }
};

template <typename T>
static int make_op(const vclhip_csr* A, DevOp<T>& op)
{
  if (!A || !A->row_ptr || !A->blocks || !A->col_idx || !A->values) return invalid("CSR matrix arrays");
  if (A->n_rows == 0 || A->n_rows != A->n_cols) return invalid("solvers need a nonempty square matrix");
  if (A->n_rows >= kPlanRowLimit || A->num_blocks == 0 || A->num_blocks > A->n_rows) return invalid("CSR plan");
  op.sell = false;
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

inline uint32_t vec_grid_n(uint32_t n) { return chunk_grid((n + 4 * kBlock - 1) / (4 * kBlock), vec_gmax()); }

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

// Stream-ordered allocations released when the object goes out of scope.
class Workspace {
 public:
  explicit Workspace(hipStream_t s) : s_(s) {}
  ~Workspace()
  {
    for (void* p : ptrs_) (void)hipFreeAsync(p, s_);
  }
  template <typename U>
  int alloc(U** p, size_t count)
  {
    void* v = nullptr;
    VCLHIP_HIP(hipMallocAsync(&v, count * sizeof(U) + 256, s_));
    ptrs_.push_back(v);
    *p = static_cast<U*>(v);
    return VCLHIP_SUCCESS;
  }

 private:
  hipStream_t s_;
  std::vector<void*> ptrs_;
};

}  // namespace vclhip
