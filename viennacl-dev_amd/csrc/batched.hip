// batched.hip -- batched CG: thousands of small independent SPD systems in one persistent launch.
//
// The systems are the diagonal blocks of one block-diagonal CSR matrix (global column indices) and
// sys_ptr gives their row ranges.  Each workgroup of the persistent kernel takes the next system
// from a work queue (largest first), runs the whole CG for it, and writes the per-system result
// (iterations, estimated error, convergence flag).  Inside a workgroup the search direction p lives
// in LDS, x and r in registers (16 rows per lane, rows <= 4096), and A p in a workspace of the plan
// (global memory, read back by the same workgroup).  The SpMV of a system is one of two kinds,
// chosen per size class by an autotuner that times the candidates on the batch itself: L lanes per
// row (L = 1..16), each group of L lanes walking its rows; or staged over the system's row blocks
// (the CSR kernel's scheme: a block's up to 2048 entries read with coalesced loads, the products
// a_ij p_j put in LDS, the rows summed from there), which keeps many more loads in flight.  The
// recurrences are the pipelined CG of cg.hpp:128-187, one system at a time.
//
// The SpMVs read the column indices as 16-bit offsets from the system's first row (a system has at
// most 4096 rows), written by one pass over the caller's 32-bit indices at the start of every solve:
// the matrix stream, re-read in every iteration, shrinks from 12 to 10 bytes per entry in FP64 and
// from 8 to 6 in FP32, for 6 bytes per entry once.  A column outside the system's block becomes
// 0xFFFF, which no row of the system reaches (the invalid-system flag, as before).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "device.hpp"
#include "plan.hpp"

struct vclhip_batch_plan {
  uint32_t num_systems = 0;
  uint32_t total_rows = 0;
  uint32_t max_rows = 0;
  std::vector<uint32_t> sys_ptr;      // host copy
  std::vector<uint32_t> order;        // launch order: largest first
  uint32_t* order_dev = nullptr;
  uint8_t* lanes_dev[2] = {nullptr, nullptr};  // SpMV choice per system, for f64 and f32
  uint32_t* blocks_dev = nullptr;     // row blocks of every system (staged SpMV), global rows
  uint32_t* sys_blk_dev = nullptr;    // system s owns blocks sys_blk[s] .. sys_blk[s+1]-1
  double* ap_dev = nullptr;           // A p workspace, total_rows entries
  uint16_t* col16_dev = nullptr;      // column offsets within each system, col16_cap entries
  size_t col16_cap = 0;
  uint32_t* counter_dev = nullptr;    // work queue head
  int lanes_log2[16][2];              // per size class (log2 rows) and value type (0 f64, 1 f32); -1 = untuned
                                      // 0..4: 2^l lanes per row, 5: staged over row blocks
  int device = 0;
  // cached graph of the last graph-mode launch
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<uintptr_t> gkey;
};

namespace vclhip {

constexpr int kBatchRowsPerLane = 16;
constexpr uint32_t kBatchMaxRows = kBlock * kBatchRowsPerLane;  // 4096
constexpr int kLaneChoices = 5;                                   // 1, 2, 4, 8, 16 lanes per row
constexpr int kStaged = kLaneChoices;                             // the staged SpMV's choice number

template <typename T>
struct BatchDev {
  const uint32_t* sys_ptr;
  const uint32_t* row_ptr;
  const uint16_t* col16;  // column - first row of the system, 0xFFFF outside its block
  const T* val;
  const T* b;
  T* x;
  const uint32_t* order;
  const uint8_t* lanes;
  const uint32_t* blocks;   // staged SpMV plan (null: not available)
  const uint32_t* sys_blk;
  T* ap;                    // A p, total_rows entries
  uint32_t* counter;
  uint32_t num_systems;
  uint32_t max_iters;
  int force_lanes;  // >= 0: every system uses 2^force_lanes lanes (autotuning)
  double tol, abs_tol;
  uint32_t* iters;
  double* err;
  uint32_t* conv;
};

// Sum over the workgroup of three per-lane values, returned to every lane in a fixed order.
__device__ __forceinline__ void batch_reduce3(double& a, double& b, double& c, double* red)
{
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[wave * 3 + 0] = a;
    red[wave * 3 + 1] = b;
    red[wave * 3 + 2] = c;
  }
  __syncthreads();
  a = red[0];
  b = red[1];
  c = red[2];
#pragma unroll
  for (int w = 1; w < kWavesPerBlock; ++w) {
    a += red[w * 3 + 0];
    b += red[w * 3 + 1];
    c += red[w * 3 + 2];
  }
}

// Ap = A p for one system (L lanes per row, p from LDS), with the lane partials of <Ap,Ap>, <p,Ap>.
template <typename T, int L>
__device__ __forceinline__ void batch_spmv(const BatchDev<T>& B, uint32_t r0, uint32_t n, const T* __restrict__ p,
                                           T* __restrict__ ap, double& ApAp, double& pAp, bool& bad)
{
  constexpr uint32_t G = kBlock / L;
  const uint32_t g = threadIdx.x / L;
  const uint32_t j = threadIdx.x % L;
  for (uint32_t i = g; i < n; i += G) {
    const uint32_t a = B.row_ptr[r0 + i], e = B.row_ptr[r0 + i + 1];
    T s = 0;
    for (uint32_t k = a + j; k < e; k += L) {
      const uint32_t c = B.col16[k];
      if (c < n)
        s += B.val[k] * p[c];
      else
        bad = true;
    }
    if constexpr (L > 1) {
#pragma unroll
      for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    }
    if (j == 0) {
      ap[i] = s;
      ApAp += static_cast<double>(s * s);
      pAp += static_cast<double>(p[i] * s);
    }
  }
}

// Ap = A p for system s (rows rs .. rs+n-1) over its row blocks, with the lane partials of <Ap,Ap>,
// <p,Ap>.  Per block: coalesced loads of the block's entries (each lane issues all its loads
// before using any), the products a_ij p_j into LDS, then each row of at most kShortRow entries
// summed in order by one lane and each longer row by a wavefront; a row longer than the staging
// buffer is summed by the whole workgroup straight from memory.
template <typename T>
__device__ __forceinline__ void batch_spmv_staged(const BatchDev<T>& B, uint32_t s, uint32_t rs, uint32_t n,
                                                  const T* __restrict__ p, T* __restrict__ ap, double& ApAp,
                                                  double& pAp, bool& bad, T* __restrict__ stage,
                                                  uint32_t* __restrict__ srp)
{
  const uint32_t tid = threadIdx.x;
  const uint32_t b1 = B.sys_blk[s + 1];
  for (uint32_t b = B.sys_blk[s]; b < b1; ++b) {
    const uint32_t d0 = B.blocks[b];
    const uint32_t r0 = d0 & (kPlanRowLimit - 1);  // global rows
    const uint32_t mode = d0 >> 28;
    const uint32_t r1 = B.blocks[b + 1] & (kPlanRowLimit - 1);
    const uint32_t e0 = B.row_ptr[r0];
    const uint32_t e1 = B.row_ptr[r1];
    if (mode == kModeLongRow) {
      T sum = 0;
      for (uint32_t e = e0 + tid; e < e1; e += kBlock) {
        const uint32_t c = B.col16[e];
        if (c < n)
          sum += B.val[e] * p[c];
        else
          bad = true;
      }
      sum = block_sum(sum);
      if (tid == 0) {
        const uint32_t i = r0 - rs;
        ap[i] = sum;
        ApAp += static_cast<double>(sum * sum);
        pAp += static_cast<double>(p[i] * sum);
      }
      continue;
    }
    const uint32_t cnt = e1 - e0;
    const uint32_t nrows = r1 - r0;
    if (cnt > kCsrNnzCap || nrows > kCsrWG) {  // the matrix is not the one the plan was made for
      bad = true;
      continue;
    }
    if (tid < nrows) srp[tid + 1] = B.row_ptr[r0 + tid + 1] - e0;
    if (tid == 0) srp[0] = 0;
    if (cnt > 0) {
      constexpr int U = kCsrNnzCap / kBlock;
      uint32_t c[U];
      T v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t e = e0 + min(tid + k * kBlock, cnt - 1);
        c[k] = B.col16[e];
        v[k] = B.val[e];
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {  // every entry's loads in flight before the first conditional store
        keep_here(c[k]);
        keep_here(v[k]);
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = tid + k * kBlock;
        if (i < cnt) {
          T prod = 0;
          if (c[k] < n)
            prod = v[k] * p[c[k]];
          else
            bad = true;
          stage[i] = prod;
        }
      }
    }
    __syncthreads();
    if (tid < nrows) {
      // clamped to the block: row pointers that disagree with the plan cannot read past the stage
      const uint32_t z = min(srp[tid + 1], cnt);
      const uint32_t a = min(srp[tid], z);
      if (z - a <= kShortRow) {
        T sum = 0;
        for (uint32_t k = a; k < z; ++k) sum += stage[k];
        const uint32_t i = r0 - rs + tid;
        ap[i] = sum;
        ApAp += static_cast<double>(sum * sum);
        pAp += static_cast<double>(p[i] * sum);
      }
    }
    if (mode == kModeMixed) {
      const uint32_t wave = tid / kWave;
      const uint32_t lane = tid & (kWave - 1);
      uint32_t turn = 0;
      for (uint32_t base = 0; base < nrows; base += kWave) {
        const uint32_t i = base + lane;
        const bool is_long = i < nrows && min(srp[i + 1], cnt) - min(srp[i], min(srp[i + 1], cnt)) > kShortRow;
        unsigned long long m = __ballot(is_long);
        while (m) {
          const uint32_t bit = static_cast<uint32_t>(__builtin_ctzll(m));
          m &= m - 1;
          if ((turn++ & (kWavesPerBlock - 1)) != wave) continue;
          const uint32_t row = base + bit;
          const uint32_t z = min(srp[row + 1], cnt);
          T sum = 0;
          for (uint32_t k = min(srp[row], z) + lane; k < z; k += kWave) sum += stage[k];
          sum = wave_sum(sum);
          if (lane == 0) {
            const uint32_t li = r0 - rs + row;
            ap[li] = sum;
            ApAp += static_cast<double>(sum * sum);
            pAp += static_cast<double>(p[li] * sum);
          }
        }
      }
    }
    __syncthreads();  // stage and srp are reused by the next block
  }
}

// The SpMV of choice l (0..4: 2^l lanes per row, kStaged: over the row blocks).
template <typename T, int L>
__device__ __forceinline__ void batch_spmv_any(const BatchDev<T>& B, uint32_t s, uint32_t r0, uint32_t n,
                                               const T* __restrict__ p, T* __restrict__ ap, double& ApAp,
                                               double& pAp, bool& bad, T* stage, uint32_t* srp)
{
  if constexpr (L == 0)
    batch_spmv_staged<T>(B, s, r0, n, p, ap, ApAp, pAp, bad, stage, srp);
  else
    batch_spmv<T, L>(B, r0, n, p, ap, ApAp, pAp, bad);
}

// Pipelined CG (cg.hpp:128-187) of system s by the whole workgroup; L = 0 selects the staged SpMV.
template <typename T, int L>
__device__ void batch_solve_system(const BatchDev<T>& B, uint32_t s, T* __restrict__ p, T* __restrict__ stage,
                                   uint32_t* __restrict__ srp, double* red)
{
  const uint32_t r0 = B.sys_ptr[s];
  const uint32_t n = B.sys_ptr[s + 1] - r0;
  const uint32_t t = threadIdx.x;
  T* const ap = B.ap + r0;
  T xr[kBatchRowsPerLane], rr[kBatchRowsPerLane];
  double bb = 0, ApAp = 0, pAp = 0;
  bool bad = false;
#pragma unroll
  for (int k = 0; k < kBatchRowsPerLane; ++k) {
    const uint32_t i = t + k * kBlock;
    xr[k] = 0;
    rr[k] = 0;
    if (i < n) {
      const T v = B.b[r0 + i];
      rr[k] = v;
      p[i] = v;
      bb += static_cast<double>(v * v);
    }
  }
  __syncthreads();
  batch_spmv_any<T, L>(B, s, r0, n, p, ap, ApAp, pAp, bad, stage, srp);
  batch_reduce3(bb, ApAp, pAp, red);
  const double nrm = sqrt(bb);  // norm_2 then squared, cg.hpp:149
  const double nrs = nrm * nrm;
  uint32_t it = 0;
  double rho = nrs;
  bool converged = false;
  if (nrs <= B.abs_tol * B.abs_tol) {
    converged = true;
  } else {
    double alpha = nrs / pAp;
    const double nap = sqrt(ApAp);
    double beta = (alpha * alpha * nap * nap - nrs) / nrs;
    while (it < B.max_iters) {
      __syncthreads();  // every lane is done with red, p and ap of the previous step
      const T ta = static_cast<T>(alpha), tb = static_cast<T>(beta);
      double part_rr = 0;
#pragma unroll
      for (int k = 0; k < kBatchRowsPerLane; ++k) {
        const uint32_t i = t + k * kBlock;
        if (i < n) {
          T pv = p[i];
          xr[k] += ta * pv;
          rr[k] -= ta * ap[i];
          pv = rr[k] + tb * pv;
          p[i] = pv;
          part_rr += static_cast<double>(rr[k] * rr[k]);
        }
      }
      __syncthreads();
      double part_ApAp = 0, part_pAp = 0;
      batch_spmv_any<T, L>(B, s, r0, n, p, ap, part_ApAp, part_pAp, bad, stage, srp);
      batch_reduce3(part_rr, part_ApAp, part_pAp, red);
      ++it;
      rho = part_rr;
      if (fabs(rho / nrs) < B.tol * B.tol || fabs(rho) < B.abs_tol * B.abs_tol) {
        converged = true;
        break;
      }
      alpha = rho / part_pAp;
      beta = (alpha * alpha * part_ApAp - rho) / rho;
    }
  }
#pragma unroll
  for (int k = 0; k < kBatchRowsPerLane; ++k) {
    const uint32_t i = t + k * kBlock;
    if (i < n) B.x[r0 + i] = xr[k];
  }
  // a column outside the system's diagonal block marks the system as invalid (conv = 2)
  const bool any_bad = __syncthreads_or(bad);
  if (t == 0) {
    B.iters[s] = it;
    B.err[s] = nrs > 0 ? sqrt(fabs(rho) / nrs) : 0.0;
    B.conv[s] = any_bad ? 2u : (converged ? 1u : 0u);
  }
  __syncthreads();  // p, stage, srp and red are reused by the next system
}

// Persistent kernel: each workgroup takes systems from the queue until it is empty.  The register
// budget matches what LDS allows: three FP64 workgroups per CU (50 KB each; 17% faster than two),
// four FP32 workgroups (16% faster than three; LDS would allow six, registers would spill).
template <typename T>
__global__ __launch_bounds__(kBlock, (sizeof(T) == 4 ? 4 : 3)) void batched_cg_kernel(BatchDev<T> B)
{
  __shared__ T p[kBatchMaxRows];
  __shared__ T stage[kCsrNnzCap];
  __shared__ uint32_t srp[kCsrWG + 1];
  __shared__ double red[3 * kWavesPerBlock];
  __shared__ uint32_t next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(B.counter, 1u);
    __syncthreads();
    const uint32_t q = next;
    __syncthreads();
    if (q >= B.num_systems) return;  // the queue is drained for every lane of the workgroup
    const uint32_t s = B.order[q];
    const int l2 = B.force_lanes >= 0 ? B.force_lanes : static_cast<int>(B.lanes[s]);
    switch (l2) {
      case 0: batch_solve_system<T, 1>(B, s, p, stage, srp, red); break;
      case 1: batch_solve_system<T, 2>(B, s, p, stage, srp, red); break;
      case 2: batch_solve_system<T, 4>(B, s, p, stage, srp, red); break;
      case 3: batch_solve_system<T, 8>(B, s, p, stage, srp, red); break;
      case 4: batch_solve_system<T, 16>(B, s, p, stage, srp, red); break;
      default:
        if (B.blocks)
          batch_solve_system<T, 0>(B, s, p, stage, srp, red);
        else
          batch_solve_system<T, 16>(B, s, p, stage, srp, red);
        break;
    }
  }
}

__global__ void batch_reset_kernel(uint32_t* counter)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// col16[e] = col[e] - (first row of e's system), 0xFFFF where that is outside the system; entries
// at or beyond `cap` (the buffer's size) are not written, whatever the row pointers say.
__global__ __launch_bounds__(kBlock) void batch_local_cols_kernel(const uint32_t* __restrict__ sys_ptr,
                                                                  const uint32_t* __restrict__ row_ptr,
                                                                  const uint32_t* __restrict__ col,
                                                                  uint16_t* __restrict__ col16, uint32_t num_systems,
                                                                  uint32_t cap)
{
  for (uint32_t s = blockIdx.x; s < num_systems; s += gridDim.x) {
    const uint32_t rs = sys_ptr[s], re = sys_ptr[s + 1];
    const uint32_t n = re - rs;
    const uint32_t e0 = row_ptr[rs];
    const uint32_t e1 = min(row_ptr[re], cap);
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
      const uint32_t c = col[e] - rs;  // unsigned: a column left of the block wraps to a large value
      col16[e] = c < n ? static_cast<uint16_t>(c) : static_cast<uint16_t>(0xFFFFu);
    }
  }
}

static int size_class(uint32_t n)
{
  int c = 0;
  while ((2u << c) <= n && c < 15) ++c;
  return c;  // floor(log2 n)
}

template <typename T>
static int persistent_grid(int* grid)
{
  int per_cu = 0;
  VCLHIP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_cg_kernel<T>, kBlock, 0));
  if (per_cu < 1) per_cu = 1;
  *grid = per_cu * device_cu_count();
  return VCLHIP_SUCCESS;
}

// Makes room for A's column offsets (outside any capture: it may allocate).
static int reserve_cols(vclhip_batch_plan* plan, const vclhip_csr* A)
{
  const size_t need = std::max<size_t>(A->nnz, 1);
  if (plan->col16_dev && plan->col16_cap >= need) return VCLHIP_SUCCESS;
  if (plan->col16_dev) (void)hipFree(plan->col16_dev);
  plan->col16_dev = nullptr;
  plan->col16_cap = 0;
  VCLHIP_HIP(hipMalloc(&plan->col16_dev, need * sizeof(uint16_t)));
  plan->col16_cap = need;
  return VCLHIP_SUCCESS;
}

static int local_cols(const vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr, hipStream_t s)
{
  const uint32_t grid = std::min<uint32_t>(plan->num_systems, 4096u);
  hipLaunchKernelGGL(batch_local_cols_kernel, dim3(grid), dim3(kBlock), 0, s, sys_ptr, A->row_ptr, A->col_idx,
                     plan->col16_dev, plan->num_systems, static_cast<uint32_t>(std::min<size_t>(plan->col16_cap, A->nnz)));
  return check_launch("batch_local_cols_kernel");
}

template <typename T>
static int launch_batched(const BatchDev<T>& B, uint32_t grid, hipStream_t s)
{
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, B.counter);
  VCLHIP_TRY(check_launch("batch_reset_kernel"));
  hipLaunchKernelGGL(batched_cg_kernel<T>, dim3(grid), dim3(kBlock), 0, s, B);
  return check_launch("batched_cg_kernel");
}

template <typename T>
static int check_batch(const vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr, const T* b,
                       const T* x, const uint32_t* iters, const double* err, const uint32_t* conv)
{
  if (!plan || !A || !sys_ptr || !b || !x || !iters || !err || !conv) return invalid("null argument");
  if (!A->row_ptr || !A->col_idx || !A->values) return invalid("CSR matrix arrays");
  if (A->n_rows != plan->total_rows || A->n_cols != plan->total_rows)
    return invalid("matrix size differs from the plan's total rows");
  return VCLHIP_SUCCESS;
}

template <typename T>
static BatchDev<T> make_batch(const vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr,
                              const T* b, T* x, const vclhip_solver_params* P, uint32_t* iters, double* err,
                              uint32_t* conv)
{
  BatchDev<T> B;
  B.sys_ptr = sys_ptr;
  B.row_ptr = A->row_ptr;
  B.col16 = plan->col16_dev;
  B.val = static_cast<const T*>(A->values);
  B.b = b;
  B.x = x;
  B.order = plan->order_dev;
  B.lanes = plan->lanes_dev[sizeof(T) == 4 ? 1 : 0];
  B.blocks = plan->blocks_dev;
  B.sys_blk = plan->sys_blk_dev;
  B.ap = reinterpret_cast<T*>(plan->ap_dev);
  B.counter = plan->counter_dev;
  B.num_systems = plan->num_systems;
  B.max_iters = P->max_iters;
  B.force_lanes = -1;
  B.tol = P->tol;
  B.abs_tol = P->abs_tol;
  B.iters = iters;
  B.err = err;
  B.conv = conv;
  return B;
}

// Times every lanes-per-row choice on the systems of each size class (at most 512 of them, a fixed
// number of iterations) and keeps the fastest.
template <typename T>
static int autotune(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr, hipStream_t s, int ti)
{
  const uint32_t ns = plan->num_systems;
  std::vector<std::vector<uint32_t> > cls(16);
  for (uint32_t i = 0; i < ns; ++i) cls[size_class(plan->sys_ptr[i + 1] - plan->sys_ptr[i])].push_back(i);
  T *b = nullptr, *x = nullptr;
  uint32_t *iters = nullptr, *conv = nullptr, *sub = nullptr;
  double* err = nullptr;
  int status = VCLHIP_SUCCESS;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    for (void* q : {static_cast<void*>(b), static_cast<void*>(x), static_cast<void*>(iters), static_cast<void*>(conv),
                    static_cast<void*>(sub), static_cast<void*>(err)})
      if (q) (void)hipFree(q);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  const size_t N = plan->total_rows;
  status = hip_status(hipMalloc(&b, N * sizeof(T)), "hipMalloc");
  if (!status) status = hip_status(hipMalloc(&x, N * sizeof(T)), "hipMalloc");
  if (!status) status = hip_status(hipMalloc(&iters, ns * sizeof(uint32_t)), "hipMalloc");
  if (!status) status = hip_status(hipMalloc(&conv, ns * sizeof(uint32_t)), "hipMalloc");
  if (!status) status = hip_status(hipMalloc(&sub, ns * sizeof(uint32_t)), "hipMalloc");
  if (!status) status = hip_status(hipMalloc(&err, ns * sizeof(double)), "hipMalloc");
  if (!status) status = hip_status(hipEventCreate(&e0), "hipEventCreate");
  if (!status) status = hip_status(hipEventCreate(&e1), "hipEventCreate");
  if (!status) {
    // right-hand side of ones: a fixed amount of work per iteration, independent of convergence
    std::vector<T> ones(N, T(1));
    status = hip_status(hipMemcpyAsync(b, ones.data(), N * sizeof(T), hipMemcpyHostToDevice, s), "b upload");
    if (!status) status = hip_status(hipStreamSynchronize(s), "sync");
  }
  int grid = 0;
  if (!status) status = persistent_grid<T>(&grid);
  if (!status) status = reserve_cols(plan, A);
  if (!status) status = local_cols(plan, A, sys_ptr, s);
  vclhip_solver_params P;
  vclhip_solver_params_default(&P);
  P.tol = 0;
  P.abs_tol = 0;
  P.max_iters = 12;
  for (int c = 0; c < 16 && !status; ++c) {
    if (cls[c].empty()) continue;
    std::vector<uint32_t> pick(cls[c].begin(), cls[c].begin() + std::min<size_t>(cls[c].size(), 512));
    status = hip_status(hipMemcpyAsync(sub, pick.data(), pick.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s),
                        "subset upload");
    if (status) break;
    BatchDev<T> B = make_batch<T>(plan, A, sys_ptr, b, x, &P, iters, err, conv);
    B.order = sub;
    B.num_systems = static_cast<uint32_t>(pick.size());
    float best = 0;
    int best_l = 0;
    const int choices = plan->blocks_dev ? kLaneChoices + 1 : kLaneChoices;
    for (int l = 0; l < choices && !status; ++l) {
      B.force_lanes = l;
      float ms_total = 0;
      for (int rep = 0; rep < 3 && !status; ++rep) {
        status = launch_batched(B, static_cast<uint32_t>(std::min<size_t>(grid, pick.size())), s);
        if (!status) status = hip_status(hipEventRecord(e0, s), "hipEventRecord");
        if (!status) status = launch_batched(B, static_cast<uint32_t>(std::min<size_t>(grid, pick.size())), s);
        if (!status) status = hip_status(hipEventRecord(e1, s), "hipEventRecord");
        if (!status) status = hip_status(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        if (!status) status = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (rep > 0) ms_total += ms;  // the first repetition warms up
      }
      if (!status && (l == 0 || ms_total < best)) {
        best = ms_total;
        best_l = l;
      }
    }
    plan->lanes_log2[c][ti] = best_l;
  }
  cleanup();
  return status;
}

// Per-system SpMV choice from the class table; where untuned, the staged SpMV when the plan has row
// blocks, else about one lane per nonzero of an average row, capped at 16.
static int upload_lanes(vclhip_batch_plan* plan, int ti, const uint32_t* row_ptr_host)
{
  if (!plan->lanes_dev[ti]) VCLHIP_HIP(hipMalloc(&plan->lanes_dev[ti], plan->num_systems));
  std::vector<uint8_t> lanes(plan->num_systems);
  for (uint32_t i = 0; i < plan->num_systems; ++i) {
    const uint32_t n = plan->sys_ptr[i + 1] - plan->sys_ptr[i];
    int l = plan->lanes_log2[size_class(n)][ti];
    if (l < 0 && plan->blocks_dev) l = kStaged;
    if (l < 0) {
      double avg = 8;
      if (row_ptr_host && n)
        avg = double(row_ptr_host[plan->sys_ptr[i + 1]] - row_ptr_host[plan->sys_ptr[i]]) / n;
      l = 0;
      while (l < kLaneChoices - 1 && (2 << l) <= avg) ++l;
    }
    lanes[i] = static_cast<uint8_t>(l);
  }
  VCLHIP_HIP(hipMemcpy(plan->lanes_dev[ti], lanes.data(), lanes.size(), hipMemcpyHostToDevice));
  return VCLHIP_SUCCESS;
}

static void drop_graph(vclhip_batch_plan* plan)
{
  if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  plan->gexec = nullptr;
  plan->graph = nullptr;
  plan->gkey.clear();
}

template <typename T>
static int batched_solve(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr, const T* b, T* x,
                         const vclhip_solver_params* P, uint32_t* iters, double* err, uint32_t* conv, void* stream)
{
  VCLHIP_TRY(check_batch<T>(plan, A, sys_ptr, b, x, iters, err, conv));
  if (!P) return invalid("params");
  int dev = 0;
  VCLHIP_HIP(hipGetDevice(&dev));
  if (dev != plan->device) return invalid("the plan was created on another device");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int grid = 0;
  VCLHIP_TRY(persistent_grid<T>(&grid));
  const uint32_t g = static_cast<uint32_t>(std::min<int64_t>(grid, plan->num_systems));
  const uint16_t* const col16_before = plan->col16_dev;
  VCLHIP_TRY(reserve_cols(plan, A));
  if (plan->col16_dev != col16_before) drop_graph(plan);
  const BatchDev<T> B = make_batch<T>(plan, A, sys_ptr, b, x, P, iters, err, conv);
  if (!P->use_graph || !s) {
    VCLHIP_TRY(local_cols(plan, A, sys_ptr, s));
    return launch_batched(B, g, s);
  }
  // graph mode: capture once per argument set, replay afterwards
  std::vector<uintptr_t> key = {reinterpret_cast<uintptr_t>(A->row_ptr), reinterpret_cast<uintptr_t>(A->col_idx),
                                reinterpret_cast<uintptr_t>(A->values), reinterpret_cast<uintptr_t>(sys_ptr),
                                reinterpret_cast<uintptr_t>(b), reinterpret_cast<uintptr_t>(x),
                                reinterpret_cast<uintptr_t>(iters), reinterpret_cast<uintptr_t>(err),
                                reinterpret_cast<uintptr_t>(conv), reinterpret_cast<uintptr_t>(plan->col16_dev),
                                static_cast<uintptr_t>(A->nnz), sizeof(T), P->max_iters};
  uint64_t tb, ab;
  std::memcpy(&tb, &P->tol, 8);
  std::memcpy(&ab, &P->abs_tol, 8);
  key.push_back(static_cast<uintptr_t>(tb));
  key.push_back(static_cast<uintptr_t>(ab));
  if (!plan->gexec || plan->gkey != key) {
    drop_graph(plan);
    VCLHIP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = local_cols(plan, A, sys_ptr, s);
    if (!st) st = launch_batched(B, g, s);
    hipGraph_t graph = nullptr;
    const int st2 = hip_status(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
    if (st || st2) {
      if (graph) (void)hipGraphDestroy(graph);
      return st ? st : st2;
    }
    plan->graph = graph;
    VCLHIP_HIP(hipGraphInstantiate(&plan->gexec, plan->graph, nullptr, nullptr, 0));
    plan->gkey = key;
  }
  VCLHIP_HIP(hipGraphLaunch(plan->gexec, s));
  return VCLHIP_SUCCESS;
}

}  // namespace vclhip

using namespace vclhip;

extern "C" {

int vclhip_batch_plan_create(uint32_t num_systems, const uint32_t* sys_ptr_host, const uint32_t* row_ptr_host,
                             vclhip_batch_plan** plan_out)
{
  if (!plan_out || !sys_ptr_host || num_systems == 0) return invalid("batch plan arguments");
  *plan_out = nullptr;
  if (sys_ptr_host[0] != 0) return invalid("sys_ptr[0] must be 0");
  uint32_t max_rows = 0;
  for (uint32_t i = 0; i < num_systems; ++i) {
    if (sys_ptr_host[i + 1] < sys_ptr_host[i]) return invalid("sys_ptr must be nondecreasing");
    max_rows = std::max(max_rows, sys_ptr_host[i + 1] - sys_ptr_host[i]);
  }
  if (max_rows > kBatchMaxRows) {
    set_error("batched CG holds a system in one workgroup: at most %u rows per system (got %u)", kBatchMaxRows,
              max_rows);
    return VCLHIP_ERROR_TOO_LARGE;
  }
  vclhip_batch_plan* plan = new vclhip_batch_plan();
  plan->num_systems = num_systems;
  plan->total_rows = sys_ptr_host[num_systems];
  plan->max_rows = max_rows;
  plan->sys_ptr.assign(sys_ptr_host, sys_ptr_host + num_systems + 1);
  for (auto& r : plan->lanes_log2) r[0] = r[1] = -1;
  plan->order.resize(num_systems);
  for (uint32_t i = 0; i < num_systems; ++i) plan->order[i] = i;
  std::stable_sort(plan->order.begin(), plan->order.end(), [&](uint32_t a, uint32_t b) {
    return plan->sys_ptr[a + 1] - plan->sys_ptr[a] > plan->sys_ptr[b + 1] - plan->sys_ptr[b];
  });
  int st = hip_status(hipGetDevice(&plan->device), "hipGetDevice");
  if (!st) st = hip_status(hipMalloc(&plan->ap_dev, std::max<size_t>(plan->total_rows, 1) * sizeof(double)), "hipMalloc");
  if (!st && row_ptr_host) {
    for (uint32_t r = 0; r < plan->total_rows && !st; ++r)
      if (row_ptr_host[r + 1] < row_ptr_host[r]) {
        set_error("invalid argument: row_ptr must be nondecreasing (row %u)", r);
        st = VCLHIP_ERROR_INVALID_VALUE;
      }
  }
  if (!st && row_ptr_host && plan->total_rows < kPlanRowLimit) {
    // row blocks of each system (the CSR kernels' partition of its rows), rows made global
    std::vector<uint32_t> blocks, sys_blk(num_systems + 1, 0), local(max_rows + 1);
    for (uint32_t i = 0; i < num_systems; ++i) {
      const uint32_t a = sys_ptr_host[i], n = sys_ptr_host[i + 1] - a;
      const uint32_t nb = n ? plan_csr_blocks(row_ptr_host + a, n, local.data()) : 0;
      for (uint32_t k = 0; k < nb; ++k) blocks.push_back(((local[k] & (kPlanRowLimit - 1)) + a) | (local[k] & ~(kPlanRowLimit - 1)));
      sys_blk[i + 1] = static_cast<uint32_t>(blocks.size());
    }
    blocks.push_back(plan->total_rows);
    st = hip_status(hipMalloc(&plan->blocks_dev, blocks.size() * sizeof(uint32_t)), "hipMalloc");
    if (!st) st = hip_status(hipMalloc(&plan->sys_blk_dev, sys_blk.size() * sizeof(uint32_t)), "hipMalloc");
    if (!st)
      st = hip_status(hipMemcpy(plan->blocks_dev, blocks.data(), blocks.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice), "blocks upload");
    if (!st)
      st = hip_status(hipMemcpy(plan->sys_blk_dev, sys_blk.data(), sys_blk.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice), "blocks upload");
  }
  if (!st) st = hip_status(hipMalloc(&plan->order_dev, num_systems * sizeof(uint32_t)), "hipMalloc");
  if (!st) st = hip_status(hipMalloc(&plan->counter_dev, sizeof(uint32_t)), "hipMalloc");
  if (!st)
    st = hip_status(hipMemcpy(plan->order_dev, plan->order.data(), num_systems * sizeof(uint32_t),
                              hipMemcpyHostToDevice), "order upload");
  if (!st) st = upload_lanes(plan, 0, row_ptr_host);
  if (!st) st = upload_lanes(plan, 1, row_ptr_host);
  if (st) {
    vclhip_batch_plan_destroy(plan);
    return st;
  }
  *plan_out = plan;
  return VCLHIP_SUCCESS;
}

int vclhip_batch_plan_destroy(vclhip_batch_plan* plan)
{
  if (!plan) return VCLHIP_SUCCESS;
  drop_graph(plan);
  if (plan->order_dev) (void)hipFree(plan->order_dev);
  for (auto* l : plan->lanes_dev)
    if (l) (void)hipFree(l);
  if (plan->counter_dev) (void)hipFree(plan->counter_dev);
  for (void* q : {static_cast<void*>(plan->blocks_dev), static_cast<void*>(plan->sys_blk_dev),
                  static_cast<void*>(plan->ap_dev), static_cast<void*>(plan->col16_dev)})
    if (q) (void)hipFree(q);
  delete plan;
  return VCLHIP_SUCCESS;
}

int vclhip_batch_plan_autotune(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                               int single_precision, void* stream)
{
  if (!plan || !A || !sys_ptr_dev) return invalid("null argument");
  if (A->n_rows != plan->total_rows) return invalid("matrix size differs from the plan's total rows");
  hipStream_t s = static_cast<hipStream_t>(stream);
  drop_graph(plan);
  const int ti = single_precision ? 1 : 0;
  if (single_precision)
    VCLHIP_TRY(autotune<float>(plan, A, sys_ptr_dev, s, ti));
  else
    VCLHIP_TRY(autotune<double>(plan, A, sys_ptr_dev, s, ti));
  return upload_lanes(plan, ti, nullptr);
}

int vclhip_batch_plan_lanes(const vclhip_batch_plan* plan, int single_precision, int32_t* lanes_log2_by_class)
{
  if (!plan || !lanes_log2_by_class) return invalid("null argument");
  for (int c = 0; c < 16; ++c) lanes_log2_by_class[c] = plan->lanes_log2[c][single_precision ? 1 : 0];
  return VCLHIP_SUCCESS;
}

int vclhip_batched_cg_solve_f64(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                                const double* b, double* x, const vclhip_solver_params* params, uint32_t* iters,
                                double* errors, uint32_t* converged, void* stream)
{
  return batched_solve<double>(plan, A, sys_ptr_dev, b, x, params, iters, errors, converged, stream);
}

int vclhip_batched_cg_solve_f32(vclhip_batch_plan* plan, const vclhip_csr* A, const uint32_t* sys_ptr_dev,
                                const float* b, float* x, const vclhip_solver_params* params, uint32_t* iters,
                                double* errors, uint32_t* converged, void* stream)
{
  return batched_solve<float>(plan, A, sys_ptr_dev, b, x, params, iters, errors, converged, stream);
}

}  // extern "C"
