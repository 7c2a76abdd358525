// gmres.hip -- pipelined GMRES (gmres.hpp:182-370) as a device-resident solver, and the reference's
// pipelined GMRES operations (iterative_operations.hpp:247-418) as entry points.
//
// The method is Walker's "simpler GMRES" with classical Gram-Schmidt, as the reference runs it:
// v_0 = A r / ||A r|| with r = (b - A x) / rho_0, then v_k = A v_{k-1} made orthogonal to
// v_0 .. v_{k-1} and normalized; R holds <v_i, v_k> above the diagonal and ||v_k|| on it, and
// xi_k = <r, v_k>.  After krylov_dim steps the error estimator, the back substitution and the
// update x += rho_0 (eta_0 r + sum eta_j v_{j-1}) close the cycle, and the residual is computed
// again from scratch.  The reference copies R and the <r, v_k> partials to the host once per cycle
// (gmres.hpp:290-300) and runs the estimator there; here a single-workgroup kernel does that work,
// and the host only polls a state record once per cycle.
//
// Storage.  A basis vector is stored as the Gram-Schmidt pass leaves it, before the division by
// sigma_k = ||v_k||: the division moves into the next SpMV's epilogue (A v~_k / sigma_k) and into
// the coefficients that read v~_k (h_j / sigma_j in Gram-Schmidt, coef_j / sigma_{j-1} in the
// update), and <r, v_k> is accumulated by the Gram-Schmidt pass itself.  In the same way r is kept
// as b - A x and rho_0 is folded into the first SpMV and the coefficients.  A step k >= 1 thus
// costs one SpMV, one pass that reads k + 1 vectors (<v_j, v_k>, j < k) and one that reads k + 2
// and writes one (v_k -= sum_j h_j v_j with ||v_k||^2 and <r, v_k>); the reference's sequence has
// a further read-modify-write pass per step (pipelined_gmres_normalize_vk).
//
// The inner products <v_j, v_k> run over all n entries, as the reference's CUDA and OpenCL kernels
// compute them (cuda/iterative_operations.hpp:1709).  Its OpenMP host kernel leaves out the last
// entries (host_based/iterative_operations.hpp:820-822: the range ends at ((n-1)/T)*T for T
// threads), which changes the iteration counts that backend reports.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "solver.hpp"
#include "dia.hpp"
#include "spmv_kernels.hpp"

namespace vclhip {
namespace gmres_detail {

constexpr uint32_t kGmDone = 1u;
constexpr int kTile = 16;               // basis vectors read by one Gram-Schmidt or update pass
constexpr uint32_t kMaxKrylov = 1024;   // largest krylov_dim accepted

struct GState {
  double norm_rhs;  // ||b||
  double rho0;      // ||b - A x|| at the start of the current cycle
  double rho;       // error estimator of the current cycle (gmres.hpp:227, :321)
  double error;     // tag.error(): |rho rho_0 / ||b||| of the last finished cycle
  double tol, abs_tol;
  double* history;  // per Krylov step: the estimate |rho rho_0 / ||b|||
  uint32_t history_len;
  uint32_t iters;       // tag.iters()
  uint32_t cycles;      // cycles started
  uint32_t max_cycles;  // tag.max_restarts() + 1
  uint32_t k_used;      // basis size used by the last update
  uint32_t flags;
  uint32_t converged;
};

// Partial-sum slots (kMaxPartials doubles each): j < K holds <v_j, v_k>; K holds ||.||^2 and
// K + 1 holds <r, v_k>.

__device__ __forceinline__ bool stopped(const uint32_t* stop) { return stop && (*stop & kGmDone); }

// Stores the workgroup sums of acc[0 .. cnt) (cnt <= J) at part[q * pstride + slot].
template <int J>
__device__ __forceinline__ void store_partials_n(const double (&acc)[J], double* part, uint32_t pstride,
                                                 uint32_t slot, uint32_t cnt)
{
  __shared__ double red[J][kWavesPerBlock];
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < J; ++q) {
    if (static_cast<uint32_t>(q) < cnt) {
      const double v = wave_sum(acc[q]);
      if ((threadIdx.x & (kWave - 1)) == 0) red[q][wave] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    double s = red[threadIdx.x][0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) s += red[threadIdx.x][w];
    part[threadIdx.x * pstride + slot] = s;
  }
}

// ------------------------------------------------------------------ SpMV epilogues

// v_0 = A r / rho_0 (gmres.hpp:228, :237-239); ||v_0||^2 and <b - A x, v_0>
template <typename T>
struct EpiFirst {
  using acc_t = double;
  static constexpr int K = 2;
  T* __restrict__ v;
  const T* __restrict__ r;
  const GState* __restrict__ st;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t* acc) const
  {
    const T y = s / static_cast<T>(st->rho0);
    v[row] = y;
    acc[0] += static_cast<double>(y * y);
    acc[1] += static_cast<double>(r[row] * y);
  }
};

// v_k = A v~_{k-1} / sigma_{k-1}  (gmres.hpp:253-256 with the division of normalize_vk moved here)
template <typename T>
struct EpiNext {
  using acc_t = double;
  static constexpr int K = 0;
  T* __restrict__ v;
  const double* __restrict__ sigma;  // &sigma[k - 1]
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t*) const
  {
    v[row] = s / static_cast<T>(*sigma);
  }
};

// r = b - A x; ||r||^2  (gmres.hpp:217-221)
template <typename T>
struct EpiResidual {
  using acc_t = double;
  static constexpr int K = 1;
  const T* __restrict__ b;
  T* __restrict__ r;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t* acc) const
  {
    const T v = b[row] - s;
    r[row] = v;
    acc[0] += static_cast<double>(v * v);
  }
};

// ------------------------------------------------------------------ vector kernels

// x = 0; r = b; partial ||b||^2  (gmres.hpp:197-205)
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void gm_init_kernel(uint32_t n, const T* __restrict__ b, T* __restrict__ x,
                                                         T* __restrict__ r, double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int P = decltype(w)::value;
    const Pack<P, T> vb = ldp<P>(b, i);
    Pack<P, T> zero;
#pragma unroll
    for (int e = 0; e < P; ++e) {
      zero.v[e] = 0;
      acc[0] += static_cast<double>(vb.v[e] * vb.v[e]);
    }
    stp<P>(x, i, zero);
    stp<P>(r, i, vb);
  });
  block_store_partials<1, double>(acc, part, kMaxPartials, lw);
}

// Gram-Schmidt stage 1: partials of <v_{j0+j}, v_k> for j < nj (host_based :787-842, over all n entries).
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void gm_dots_kernel(uint32_t n, const T* __restrict__ V, size_t ld, uint32_t k,
                                                         uint32_t j0, uint32_t nj, const uint32_t* __restrict__ stop,
                                                         double* __restrict__ part)
{
  if (stopped(stop)) return;
  const T* __restrict__ vk = V + k * ld;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[kTile];
#pragma unroll
  for (int j = 0; j < kTile; ++j) acc[j] = 0;
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int P = decltype(w)::value;
    const Pack<P, T> a = ldp<P>(vk, i);
    Pack<P, T> q[kTile];
#pragma unroll
    for (int j = 0; j < kTile; ++j)
      if (static_cast<uint32_t>(j) < nj) q[j] = ldp<P>(V + (j0 + j) * ld, i);
#pragma unroll
    for (int j = 0; j < kTile; ++j)
      if (static_cast<uint32_t>(j) < nj)
#pragma unroll
        for (int e = 0; e < P; ++e) acc[j] += static_cast<double>(q[j].v[e] * a.v[e]);
  });
  store_partials_n<kTile>(acc, part + j0 * kMaxPartials, kMaxPartials, lw, nj);
}

// Gram-Schmidt stage 2 (host_based :851-893): v_k -= sum_{j<nj} c[j0+j] v_{j0+j}, in j order.  The
// last pass (`last`) also accumulates ||v_k||^2 and, when r is given, <r, v_k> into part[0] and
// part[kMaxPartials].
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void gm_orth_kernel(uint32_t n, T* __restrict__ V, size_t ld, uint32_t k,
                                                         uint32_t j0, uint32_t nj, uint32_t last,
                                                         const T* __restrict__ r, const double* __restrict__ c,
                                                         const uint32_t* __restrict__ stop,
                                                         double* __restrict__ part)
{
  if (stopped(stop)) return;
  T* __restrict__ vk = V + k * ld;
  T cj[kTile];
#pragma unroll
  for (int j = 0; j < kTile; ++j) cj[j] = static_cast<uint32_t>(j) < nj ? static_cast<T>(c[j0 + j]) : T(0);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[2] = {0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int P = decltype(w)::value;
    Pack<P, T> a = ldp<P>(vk, i);
    Pack<P, T> q[kTile];
#pragma unroll
    for (int j = 0; j < kTile; ++j)
      if (static_cast<uint32_t>(j) < nj) q[j] = ldp<P>(V + (j0 + j) * ld, i);
    Pack<P, T> vr;
    if (last && r) vr = ldp<P>(r, i);
#pragma unroll
    for (int j = 0; j < kTile; ++j)
      if (static_cast<uint32_t>(j) < nj)
#pragma unroll
        for (int e = 0; e < P; ++e) a.v[e] -= cj[j] * q[j].v[e];
    if (last) {
#pragma unroll
      for (int e = 0; e < P; ++e) {
        acc[0] += static_cast<double>(a.v[e] * a.v[e]);
        if (r) acc[1] += static_cast<double>(vr.v[e] * a.v[e]);
      }
    }
    stp<P>(vk, i, a);
  });
  if (last) block_store_partials<2, double>(acc, part, kMaxPartials, lw);
}

// x += c[0] r + sum_{j=1}^{k-1} c[j] v_{j-1}  (host_based :898-926).  With a state record, k is the
// basis size of the cycle that just ended.
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void gm_update_kernel(uint32_t n, T* __restrict__ x, const T* __restrict__ r,
                                                           const T* __restrict__ V, size_t ld,
                                                           const double* __restrict__ c,
                                                           const GState* __restrict__ st, uint32_t k_arg)
{
  uint32_t k = k_arg;
  if (st) {
    if (st->flags & kGmDone) return;
    k = st->k_used;
  }
  const T c0 = static_cast<T>(c[0]);
  const uint32_t lw = xcd_logical_block();
  uint32_t b0, b1;
  chunk_range_w(n, W, lw, gridDim.x, b0, b1);
  for_packs<W>(b0, b1, [&](uint32_t i, auto w) {
    constexpr int P = decltype(w)::value;
    Pack<P, T> vx = ldp<P>(x, i);
    const Pack<P, T> vr = ldp<P>(r, i);
#pragma unroll
    for (int e = 0; e < P; ++e) vx.v[e] += c0 * vr.v[e];
    for (uint32_t j0 = 1; j0 < k; j0 += kTile) {
      const uint32_t nj = min(static_cast<uint32_t>(kTile), k - j0);
      Pack<P, T> q[kTile];
#pragma unroll
      for (int j = 0; j < kTile; ++j)
        if (static_cast<uint32_t>(j) < nj) q[j] = ldp<P>(V + (j0 - 1 + j) * ld, i);
#pragma unroll
      for (int j = 0; j < kTile; ++j)
        if (static_cast<uint32_t>(j) < nj) {
          const T cj = static_cast<T>(c[j0 + j]);
#pragma unroll
          for (int e = 0; e < P; ++e) vx.v[e] += cj * q[j].v[e];
        }
    }
    stp<P>(x, i, vx);
  });
}

// ------------------------------------------------------------------ scalar kernels (one workgroup)

// Start of a cycle (gmres.hpp:212-232): rho_0 = ||r||, the trivial-right-hand-side and convergence
// tests, rho = 1.
__global__ __launch_bounds__(kBlock) void gm_cycle_start_kernel(GState* __restrict__ st,
                                                                const double* __restrict__ part, uint32_t g,
                                                                uint32_t first)
{
  if (st->flags & kGmDone) return;
  const double rr = block_sum_partials(part, g);
  if (threadIdx.x != 0) return;
  const double rho0 = sqrt(rr);
  if (first) st->norm_rhs = rho0;
  st->rho0 = rho0;
  if (st->cycles >= st->max_cycles) {
    st->flags |= kGmDone;
    return;
  }
  if (rho0 <= st->abs_tol) {
    st->flags |= kGmDone;
    st->converged = 1;
    return;
  }
  st->rho = 1.0;
  if (rho0 / st->norm_rhs < st->tol || rho0 < st->abs_tol) {
    st->flags |= kGmDone;
    st->converged = 1;
    return;
  }
  st->cycles += 1;
}

// h_j = <v_j, v_k> for j = blockIdx.x (stage 2 step 1, host_based :867-870).  The partials hold
// <v~_j, v_k> with the stored, unnormalized v~_j = sigma_j v_j, so h_j = <v~_j, v_k> / sigma_j;
// R[j + k K] = h_j, and the coefficient on the stored vector is c_j = h_j / sigma_j.
__global__ __launch_bounds__(kBlock) void gm_h_kernel(const GState* __restrict__ st, const double* __restrict__ part,
                                                      uint32_t g, uint32_t k, uint32_t K, double* __restrict__ R,
                                                      const double* __restrict__ sigma, double* __restrict__ c)
{
  if (st->flags & kGmDone) return;
  const uint32_t j = blockIdx.x;
  const double d = block_sum_partials(part + j * kMaxPartials, g);
  if (threadIdx.x == 0) {
    const double h = d / sigma[j];
    R[j + k * K] = h;
    c[j] = h / sigma[j];
  }
}

// ||v_k|| and xi_k = <r/rho_0, v_k/||v_k||>  (normalize_vk, host_based :752-776): R[k + k K],
// sigma_k, xi_k.
__global__ __launch_bounds__(kBlock) void gm_norm_kernel(const GState* __restrict__ st,
                                                         const double* __restrict__ part, uint32_t g, uint32_t k,
                                                         uint32_t K, double* __restrict__ R,
                                                         double* __restrict__ sigma, double* __restrict__ xi)
{
  if (st->flags & kGmDone) return;
  const double nsq = block_sum_partials(part, g);
  const double rv = block_sum_partials(part + kMaxPartials, g);
  if (threadIdx.x != 0) return;
  const double s = sqrt(nsq);
  R[k + k * K] = s;
  sigma[k] = s;
  xi[k] = rv / st->rho0 / s;
}

__device__ __forceinline__ void gm_record(GState* st, uint32_t iter, double v)
{
  if (st->history && iter >= 1 && iter <= st->history_len) st->history[iter - 1] = v;
}

// End of a cycle (gmres.hpp:291-352): premature-convergence check, error estimator, back
// substitution and the update coefficients.  coef persists across cycles: the update always
// applies coef[0] (host_based :920), which keeps the previous cycle's value when no step is used.
__global__ __launch_bounds__(kWave) void gm_cycle_end_kernel(GState* __restrict__ st, const double* __restrict__ R,
                                                             const double* __restrict__ xi,
                                                             const double* __restrict__ sigma,
                                                             double* __restrict__ coef, double* __restrict__ upd,
                                                             uint32_t K)
{
  __shared__ double eta[kMaxKrylov];
  if (threadIdx.x != 0 || (st->flags & kGmDone)) return;
  uint32_t k = K;
  const uint32_t full = K;
  const double tol = st->tol;
  for (uint32_t i = 0; i < K; ++i)
    if (fabs(R[i + i * K]) < tol * R[0]) {
      k = i;
      break;
    }
  double rho = st->rho;
  uint32_t iters = st->iters;
  const double rho0 = st->rho0;
  const double nb = st->norm_rhs;
  for (uint32_t i = 0; i < k; ++i) {
    ++iters;
    const double v = xi[i];
    if (v >= rho || v <= -rho) {
      gm_record(st, iters, fabs(rho * rho0 / nb));
      k = i;
      break;
    }
    rho *= sin(acos(v / rho));
    gm_record(st, iters, fabs(rho * rho0 / nb));
  }
  for (uint32_t i = 0; i < k; ++i) eta[i] = xi[i];
  for (int i2 = static_cast<int>(k) - 1; i2 > -1; --i2) {
    const uint32_t i = static_cast<uint32_t>(i2);
    double e = eta[i];
    for (uint32_t j = i + 1; j < k; ++j) e -= R[i + j * full] * eta[j];
    eta[i] = e / R[i + i * full];
  }
  for (uint32_t i = 0; i < k; ++i) coef[i] = rho0 * eta[i];
  upd[0] = coef[0] / rho0;
  for (uint32_t j = 1; j < k; ++j) upd[j] = coef[j] / sigma[j - 1];
  st->rho = rho;
  st->iters = iters;
  st->k_used = k;
  st->error = fabs(rho * rho0 / nb);
}

// ------------------------------------------------------------------ kernels of the operation entry points

// norm = sqrt(sum buf[chunk .. 2 chunk)) in order; R[offset] = norm  (host_based :752-756)
__global__ void gm_op_norm_kernel(const double* __restrict__ buf, uint32_t chunk, double* __restrict__ R,
                                  size_t offset, double* __restrict__ norm)
{
  if (threadIdx.x != 0) return;
  double s = 0;
  for (uint32_t i = 0; i < chunk; ++i) s += buf[i + chunk];
  s = sqrt(s);
  R[offset] = s;
  *norm = s;
}

// v_k /= norm; partial <r, v_k>  (host_based :761-776)
template <int W>
__global__ __launch_bounds__(kBlock) void gm_op_normalize_kernel(uint32_t n, double* __restrict__ v,
                                                                 const double* __restrict__ r,
                                                                 const double* __restrict__ norm,
                                                                 double* __restrict__ part)
{
  const double nv = *norm;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int P = decltype(w)::value;
    Pack<P, double> a = ldp<P>(v, i);
    const Pack<P, double> vr = ldp<P>(r, i);
#pragma unroll
    for (int e = 0; e < P; ++e) {
      a.v[e] = a.v[e] / nv;
      acc[0] += vr.v[e] * a.v[e];
    }
    stp<P>(v, i, a);
  });
  block_store_partials<1, double>(acc, part, kMaxPartials, lw);
}

// buf[dest + q * qstride] = sum of slot q's g partials, the other chunk - 1 entries of that chunk zero.
__global__ __launch_bounds__(kBlock) void gm_op_chunks_kernel(const double* __restrict__ part, uint32_t g,
                                                              double* __restrict__ buf, size_t dest,
                                                              size_t qstride, uint32_t chunk)
{
  const uint32_t q = blockIdx.x;
  const double v = block_sum_partials(part + q * kMaxPartials, g);
  double* out = buf + dest + q * qstride;
  for (uint32_t i = threadIdx.x; i < chunk; i += kBlock) out[i] = (i == 0) ? v : 0.0;
}

// h_j = sum_c buf[j chunk + c] in order; R[j + k krylov_dim] = h_j  (host_based :867-870, :889-891)
__global__ void gm_op_h_kernel(const double* __restrict__ buf, uint32_t chunk, uint32_t k, uint32_t krylov_dim,
                               double* __restrict__ R, double* __restrict__ h)
{
  for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
    double s = 0;
    for (uint32_t c = 0; c < chunk; ++c) s += buf[j * chunk + c];
    h[j] = s;
    R[j + static_cast<size_t>(k) * krylov_dim] = s;
  }
}

// ------------------------------------------------------------------ host side

template <typename T>
struct GmLaunch {
  uint32_t n;
  size_t ld;
  uint32_t g;
  bool wide;  // 16-byte packs on every vector
};

template <typename T>
void launch_dots(const GmLaunch<T>& L, hipStream_t s, const T* V, uint32_t k, uint32_t j0, uint32_t nj,
                 const uint32_t* stop, double* part)
{
  if (L.wide)
    hipLaunchKernelGGL((gm_dots_kernel<T, kVecW<T> >), dim3(L.g), dim3(kBlock), 0, s, L.n, V, L.ld, k, j0, nj, stop,
                       part);
  else
    hipLaunchKernelGGL((gm_dots_kernel<T, 1>), dim3(L.g), dim3(kBlock), 0, s, L.n, V, L.ld, k, j0, nj, stop, part);
}

template <typename T>
void launch_orth(const GmLaunch<T>& L, hipStream_t s, T* V, uint32_t k, uint32_t j0, uint32_t nj, bool last,
                 const T* r, const double* c, const uint32_t* stop, double* part)
{
  if (L.wide)
    hipLaunchKernelGGL((gm_orth_kernel<T, kVecW<T> >), dim3(L.g), dim3(kBlock), 0, s, L.n, V, L.ld, k, j0, nj,
                       last ? 1u : 0u, r, c, stop, part);
  else
    hipLaunchKernelGGL((gm_orth_kernel<T, 1>), dim3(L.g), dim3(kBlock), 0, s, L.n, V, L.ld, k, j0, nj,
                       last ? 1u : 0u, r, c, stop, part);
}

template <typename T>
void launch_update(const GmLaunch<T>& L, hipStream_t s, T* x, const T* r, const T* V, const double* c,
                   const GState* st, uint32_t k)
{
  if (L.wide)
    hipLaunchKernelGGL((gm_update_kernel<T, kVecW<T> >), dim3(L.g), dim3(kBlock), 0, s, L.n, x, r, V, L.ld, c, st, k);
  else
    hipLaunchKernelGGL((gm_update_kernel<T, 1>), dim3(L.g), dim3(kBlock), 0, s, L.n, x, r, V, L.ld, c, st, k);
}

struct PinnedG {
  GState* h = nullptr;
  ~PinnedG()
  {
    if (h) (void)hipHostFree(h);
  }
};

struct GraphG {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphG()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct EventsG {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~EventsG()
  {
    for (auto ev : e)
      if (ev) (void)hipEventDestroy(ev);
  }
};

template <typename T, class M>
int gmres_solve(const M* Ain, const T* b, T* x, const vclhip_solver_params* P, vclhip_solver_result* Res,
                void* stream)
{
  if (!P || !Res || !b || !x) return invalid("null argument");
  const uint32_t K = P->krylov_dim;
  if (K == 0 || K > kMaxKrylov) return invalid("krylov_dim must be in [1, 1024]");
  DevOp<T> op;
  VCLHIP_TRY(make_op(Ain, op));
  SolveStream ss(stream, P->use_graph != 0);
  VCLHIP_TRY(ss.status());
  hipStream_t s = ss.get();
  const uint32_t n = op.n;
  const size_t ld = (static_cast<size_t>(n) + 63) / 64 * 64;  // 512-byte aligned basis vectors
  uint32_t max_restarts = P->max_iters / K;                   // gmres_tag::max_restarts (gmres.hpp:75-81)
  if (max_restarts > 0 && max_restarts * K == P->max_iters) max_restarts -= 1;
  const uint32_t max_cycles = max_restarts + 1;

  Workspace ws(s);
  int used_dia = 0;
  VCLHIP_TRY(dia_operator(op, ws, s, P, &used_dia));
  T *V, *r;
  double *R, *sigma, *xi, *coef, *upd, *c, *part, *hist = nullptr;
  GState* st;
  VCLHIP_TRY(ws.alloc(&V, ld * K));
  VCLHIP_TRY(ws.alloc(&r, n));
  VCLHIP_TRY(ws.alloc(&R, static_cast<size_t>(K) * K));
  VCLHIP_TRY(ws.alloc(&sigma, K));
  VCLHIP_TRY(ws.alloc(&xi, K));
  VCLHIP_TRY(ws.alloc(&coef, K));
  VCLHIP_TRY(ws.alloc(&upd, K));
  VCLHIP_TRY(ws.alloc(&c, K));
  VCLHIP_TRY(ws.alloc(&part, static_cast<size_t>(K + 2) * kMaxPartials));
  VCLHIP_TRY(ws.alloc(&st, 1));
  const uint32_t hist_len = (P->history_host && P->history_len) ? P->history_len : 0;
  if (hist_len) VCLHIP_TRY(ws.alloc(&hist, hist_len));
  VCLHIP_HIP(hipMemsetAsync(coef, 0, K * sizeof(double), s));
  VCLHIP_HIP(hipMemsetAsync(R, 0, static_cast<size_t>(K) * K * sizeof(double), s));

  GState h;
  std::memset(&h, 0, sizeof(h));
  h.tol = P->tol;
  h.abs_tol = P->abs_tol;
  h.history = hist;
  h.history_len = hist_len;
  h.max_cycles = max_cycles;
  VCLHIP_HIP(hipMemcpyAsync(st, &h, sizeof(GState), hipMemcpyHostToDevice, s));
  VCLHIP_HIP(hipStreamSynchronize(s));

  GmLaunch<T> L{n, ld, solver_vec_grid(n), aligned16({b, x})};
  const uint32_t* stop = &st->flags;
  double* pnorm = part + static_cast<size_t>(K) * kMaxPartials;

  auto cycle_body = [&]() -> int {
    for (uint32_t k = 0; k < K; ++k) {
      T* vk = V + k * ld;
      if (k == 0) {
        op.launch(s, r, EpiFirst<T>{vk, r, st}, pnorm, stop, kGmDone);
      } else {
        op.launch(s, V + (k - 1) * ld, EpiNext<T>{vk, sigma + (k - 1)}, pnorm, stop, kGmDone);
        for (uint32_t j0 = 0; j0 < k; j0 += kTile)
          launch_dots<T>(L, s, V, k, j0, std::min<uint32_t>(kTile, k - j0), stop, part);
        hipLaunchKernelGGL(gm_h_kernel, dim3(k), dim3(kBlock), 0, s, st, part, L.g, k, K, R, sigma, c);
        for (uint32_t j0 = 0; j0 < k; j0 += kTile) {
          const uint32_t nj = std::min<uint32_t>(kTile, k - j0);
          launch_orth<T>(L, s, V, k, j0, nj, j0 + nj == k, r, c, stop, pnorm);
        }
      }
      hipLaunchKernelGGL(gm_norm_kernel, dim3(1), dim3(kBlock), 0, s, st, pnorm, k == 0 ? op.grid : L.g, k, K, R,
                         sigma, xi);
      VCLHIP_TRY(check_launch("gmres step"));
    }
    hipLaunchKernelGGL(gm_cycle_end_kernel, dim3(1), dim3(kWave), 0, s, st, R, xi, sigma, coef, upd, K);
    launch_update<T>(L, s, x, r, V, upd, st, 0u);
    return check_launch("gmres cycle end");
  };
  auto restart = [&]() -> int {
    op.launch(s, x, EpiResidual<T>{b, r}, pnorm, stop, kGmDone);
    hipLaunchKernelGGL(gm_cycle_start_kernel, dim3(1), dim3(kBlock), 0, s, st, pnorm, op.grid, 0u);
    return check_launch("gmres restart");
  };

  // cycle 0: x = 0, r = b
  if (L.wide)
    hipLaunchKernelGGL((gm_init_kernel<T, kVecW<T> >), dim3(L.g), dim3(kBlock), 0, s, n, b, x, r, pnorm);
  else
    hipLaunchKernelGGL((gm_init_kernel<T, 1>), dim3(L.g), dim3(kBlock), 0, s, n, b, x, r, pnorm);
  hipLaunchKernelGGL(gm_cycle_start_kernel, dim3(1), dim3(kBlock), 0, s, st, pnorm, L.g, 1u);
  VCLHIP_TRY(check_launch("gmres start"));
  VCLHIP_TRY(cycle_body());

  PinnedG pin;
  VCLHIP_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin.h), 2 * sizeof(GState), hipHostMallocDefault));
  EventsG ev;
  VCLHIP_HIP(hipEventCreateWithFlags(&ev.e[0], hipEventDisableTiming));
  VCLHIP_HIP(hipEventCreateWithFlags(&ev.e[1], hipEventDisableTiming));
  GraphG gr;
  const bool monitor = P->monitor != nullptr;
  const bool want_graph = P->use_graph && !monitor;

  int status = VCLHIP_SUCCESS;
  int pending = -1, slot = 0;
  for (uint32_t cyc = 0; cyc < max_cycles && !status; ++cyc) {
    if (cyc > 0) {
      if (want_graph) {
        if (!gr.exec) {
          status = hip_status(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
          if (status) break;
          int st2 = restart();
          if (!st2) st2 = cycle_body();
          status = hip_status(hipStreamEndCapture(s, &gr.graph), "hipStreamEndCapture");
          if (st2) status = st2;
          if (!status) status = hip_status(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0),
                                           "hipGraphInstantiate");
          if (status) break;
        }
        status = hip_status(hipGraphLaunch(gr.exec, s), "hipGraphLaunch");
      } else {
        status = restart();
        if (!status) status = cycle_body();
      }
      if (status) break;
    }
    status = hip_status(hipMemcpyAsync(&pin.h[slot], st, sizeof(GState), hipMemcpyDeviceToHost, s), "state D2H");
    if (!status) status = hip_status(hipEventRecord(ev.e[slot], s), "hipEventRecord");
    if (status) break;
    bool finished = false;
    if (monitor) {
      status = hip_status(hipEventSynchronize(ev.e[slot]), "hipEventSynchronize");
      if (status) break;
      const GState& g = pin.h[slot];
      if (g.flags & kGmDone) {
        finished = true;
      } else if (P->monitor(x, g.error, P->monitor_user)) {  // gmres.hpp:366-367
        const uint32_t done = kGmDone;
        status = hip_status(hipMemcpyAsync(&st->flags, &done, sizeof(uint32_t), hipMemcpyHostToDevice, s),
                            "stop flag");
        if (!status) status = hip_status(hipStreamSynchronize(s), "stop flag");
        finished = true;
      }
    } else if (pending >= 0) {
      status = hip_status(hipEventSynchronize(ev.e[pending]), "hipEventSynchronize");
      if (status) break;
      if (pin.h[pending].flags & kGmDone) finished = true;
    }
    if (finished) break;
    pending = slot;
    slot ^= 1;
  }
  const int st_sync = hip_status(hipStreamSynchronize(s), "gmres sync");
  if (!status) status = st_sync;
  if (status) return status;
  GState fin;
  VCLHIP_HIP(hipMemcpy(&fin, st, sizeof(GState), hipMemcpyDeviceToHost));
  Res->iters = fin.iters;
  Res->error = fin.error;
  Res->restarts = fin.cycles;
  Res->operator_used = used_dia;
  Res->converged = static_cast<int>(fin.converged);
  Res->kernels_per_iter = 0;
  if (hist && fin.iters) {
    VCLHIP_HIP(hipMemcpyAsync(P->history_host, hist, std::min(hist_len, fin.iters) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    VCLHIP_HIP(hipStreamSynchronize(s));
  }
  return VCLHIP_SUCCESS;
}

bool basis_ok(const void* p, size_t ld) { return aligned16({p}) && (ld % 2 == 0); }

}  // namespace gmres_detail
}  // namespace vclhip

using namespace vclhip;
using namespace vclhip::gmres_detail;

extern "C" {

int vclhip_gmres_solve_csr_f64(const vclhip_csr* A, const double* b, double* x, const vclhip_solver_params* P,
                               vclhip_solver_result* R, void* stream)
{
  return gmres_solve<double>(A, b, x, P, R, stream);
}

int vclhip_gmres_solve_sell_f64(const vclhip_sell* A, const double* b, double* x, const vclhip_solver_params* P,
                                vclhip_solver_result* R, void* stream)
{
  return gmres_solve<double>(A, b, x, P, R, stream);
}

int vclhip_gmres_solve_csr_f32(const vclhip_csr* A, const float* b, float* x, const vclhip_solver_params* P,
                               vclhip_solver_result* R, void* stream)
{
  return gmres_solve<float>(A, b, x, P, R, stream);
}

int vclhip_gmres_solve_sell_f32(const vclhip_sell* A, const float* b, float* x, const vclhip_solver_params* P,
                                vclhip_solver_result* R, void* stream)
{
  return gmres_solve<float>(A, b, x, P, R, stream);
}

int vclhip_gmres_solve_f64(const vclhip_matrix* A, const double* b, double* x, const vclhip_solver_params* P,
                           vclhip_solver_result* R, void* stream)
{
  return gmres_solve<double>(A, b, x, P, R, stream);
}

int vclhip_gmres_solve_f32(const vclhip_matrix* A, const float* b, float* x, const vclhip_solver_params* P,
                           vclhip_solver_result* R, void* stream)
{
  return gmres_solve<float>(A, b, x, P, R, stream);
}

int vclhip_pipelined_gmres_normalize_vk_f64(uint32_t n, double* v_k, const double* residual, double* R,
                                            size_t offset_in_R, const double* inner_prod_buffer,
                                            double* r_dot_vk_buffer, uint32_t chunk, size_t chunk_offset,
                                            void* stream)
{
  if (!v_k || !residual || !R || !inner_prod_buffer || !r_dot_vk_buffer || chunk == 0) return invalid("null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double *part, *norm;
  VCLHIP_TRY(ws.alloc(&part, kMaxPartials));
  VCLHIP_TRY(ws.alloc(&norm, 1));
  hipLaunchKernelGGL(gm_op_norm_kernel, dim3(1), dim3(kWave), 0, s, inner_prod_buffer, chunk, R, offset_in_R, norm);
  const uint32_t g = vec_grid_n(n);
  if (aligned16({v_k, residual}))
    hipLaunchKernelGGL((gm_op_normalize_kernel<2>), dim3(g), dim3(kBlock), 0, s, n, v_k, residual, norm, part);
  else
    hipLaunchKernelGGL((gm_op_normalize_kernel<1>), dim3(g), dim3(kBlock), 0, s, n, v_k, residual, norm, part);
  hipLaunchKernelGGL(gm_op_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, g, r_dot_vk_buffer, chunk_offset,
                     size_t(0), chunk);
  return check_launch("pipelined_gmres_normalize_vk");
}

int vclhip_pipelined_gmres_gram_schmidt_stage1_f64(const double* basis, uint32_t n, size_t ld, uint32_t k,
                                                   double* vi_in_vk_buffer, uint32_t chunk, void* stream)
{
  if (!basis || !vi_in_vk_buffer || chunk == 0) return invalid("null argument");
  if (ld < n) return invalid("internal size smaller than the vector size");
  if (k == 0) return VCLHIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double* part;
  VCLHIP_TRY(ws.alloc(&part, static_cast<size_t>(k) * kMaxPartials));
  GmLaunch<double> L{n, ld, vec_grid_n(n), basis_ok(basis, ld)};
  for (uint32_t j0 = 0; j0 < k; j0 += kTile)
    launch_dots<double>(L, s, basis, k, j0, std::min<uint32_t>(kTile, k - j0), nullptr, part);
  hipLaunchKernelGGL(gm_op_chunks_kernel, dim3(k), dim3(kBlock), 0, s, part, L.g, vi_in_vk_buffer, size_t(0),
                     static_cast<size_t>(chunk), chunk);
  return check_launch("pipelined_gmres_gram_schmidt_stage1");
}

int vclhip_pipelined_gmres_gram_schmidt_stage2_f64(double* basis, uint32_t n, size_t ld, uint32_t k,
                                                   const double* vi_in_vk_buffer, double* R, uint32_t krylov_dim,
                                                   double* inner_prod_buffer, uint32_t chunk, void* stream)
{
  if (!basis || !vi_in_vk_buffer || !R || !inner_prod_buffer || chunk == 0) return invalid("null argument");
  if (ld < n) return invalid("internal size smaller than the vector size");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double *part, *h;
  VCLHIP_TRY(ws.alloc(&part, 2 * kMaxPartials));
  VCLHIP_TRY(ws.alloc(&h, std::max<uint32_t>(k, 1)));
  if (k) hipLaunchKernelGGL(gm_op_h_kernel, dim3(1), dim3(kWave), 0, s, vi_in_vk_buffer, chunk, k, krylov_dim, R, h);
  GmLaunch<double> L{n, ld, vec_grid_n(n), basis_ok(basis, ld)};
  if (k == 0) {
    launch_orth<double>(L, s, basis, 0, 0, 0, true, nullptr, h, nullptr, part);
  } else {
    for (uint32_t j0 = 0; j0 < k; j0 += kTile) {
      const uint32_t nj = std::min<uint32_t>(kTile, k - j0);
      launch_orth<double>(L, s, basis, k, j0, nj, j0 + nj == k, nullptr, h, nullptr, part);
    }
  }
  hipLaunchKernelGGL(gm_op_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, L.g, inner_prod_buffer,
                     static_cast<size_t>(chunk), size_t(0), chunk);
  return check_launch("pipelined_gmres_gram_schmidt_stage2");
}

int vclhip_pipelined_gmres_update_result_f64(double* result, const double* residual, const double* basis, uint32_t n,
                                             size_t ld, const double* coefficients, uint32_t k, void* stream)
{
  if (!result || !residual || !coefficients || (k > 1 && !basis)) return invalid("null argument");
  if (k > 1 && ld < n) return invalid("internal size smaller than the vector size");
  hipStream_t s = static_cast<hipStream_t>(stream);
  GmLaunch<double> L{n, ld, vec_grid_n(n), aligned16({result, residual}) && (k <= 1 || basis_ok(basis, ld))};
  launch_update<double>(L, s, result, residual, basis, coefficients, nullptr, k);
  return check_launch("pipelined_gmres_update_result");
}

}  // extern "C"
