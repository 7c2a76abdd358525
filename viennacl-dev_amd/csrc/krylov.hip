// krylov.hip -- fused pipelined CG / BiCGStab kernels, the reference's pipelined operations
// (iterative_operations.hpp) as entry points, and the device-resident solver drivers.
//
// Solver design.  Each iteration is a short fixed sequence of kernels; all scalars (alpha, beta,
// omega, inner products, iteration count, convergence flags) live in a device-resident state
// record written only by single-workgroup "scalar" kernels.  The host enqueues iterations in
// chunks without waiting, polls the state of the previous chunk while the next one runs, and
// every kernel returns at once after the state says the solve has ended.  There is therefore no
// per-iteration device-to-host copy (the reference copies its partial buffer back every
// iteration, cg.hpp:168 and bicgstab.hpp:180).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.hpp"
#include "solver.hpp"
#include "spmv_kernels.hpp"
#include "dia.hpp"

namespace vclhip {

// ------------------------------------------------------------------ solver state

enum : uint32_t {
  kFlagDone = 1u,       // every kernel of the solve returns at once
  kFlagFinishing = 2u,  // BiCGStab: the final vector update is pending, everything else stops
  // Mixed-precision CG.  An iteration that asks for an FP64 correction sets Pause (the remaining
  // iterations of the chunk return at once) and clears NoPause (the correction kernels at the end of
  // the chunk run).  The correction clears TailIdle while it refreshes the FP32 state.
  kFlagPause = 4u,
  kFlagNoPause = 8u,
  kFlagTailIdle = 16u,
};

constexpr uint32_t kSlots = 8;  // partial-sum slots of kMaxPartials doubles each
// Compact CG: at most this many partials read by all update workgroups together per iteration
// (1 Mi: the 512^2 diffusion operator, 256 update workgroups reading 2304 partials each, runs 9%
// faster than with three kernels; at n = 1 M the reads would cost more than the kernel saved).
constexpr uint64_t kCompactPartialReads = 1ull << 20;
// Fused CG (one kernel per iteration): every workgroup of the grid reduces all partials, so the
// bound is on grid x partials.  Measured (CG on 2-D and 3-D Laplacians, one kernel against two,
// with vector chunks of 1024 elements): 8.2 against 10.6 us per iteration at 147 K reads (256^2),
// 11.0 against 12.0 at 518 K (350^2), even at 555-832 K (50^3, 392^2), 17.6 against 14.6 at 1.4 M
// (450^2).
constexpr uint64_t kFusedPartialReads = 1ull << 19;

struct KState {
  double alpha, beta, omega;
  double rr;        // CG: <r,r> or <r,M^-1 r>;  BiCGStab: <r,r0*>
  double nrs;       // CG: norm of the right-hand side squared;  BiCGStab: ||b||
  double res;       // BiCGStab residual-norm estimate
  double ap_r0;     // BiCGStab <Ap,r0*> of the current iteration
  double tol, abs_tol;
  double* history;  // device array or null
  uint32_t history_len;
  uint32_t iter, max_iters;
  uint32_t flags;
  uint32_t converged;
  uint32_t final_mode;        // BiCGStab: 0 full update, 1 converged (x only), 2 last iteration (full)
  uint32_t rr_from_partials;  // BiCGStab: <r,r0*> comes from the previous vector update's partials
  uint32_t breakdown;
  // mixed-precision CG
  double rr2;       // FP32 recursive <r,r>
  double rr_ref;    // <r,r> after the last FP64 correction (reliable update)
  float inner, initial;  // reference schedule: FP32 <r,r> and its value at the last restart
  float inner_tol;
  uint32_t updates;      // FP64 corrections done
  uint32_t iter2[2];     // CG, compact mode: iteration count, read from [k & 1] and written to the other
};

__host__ __device__ __forceinline__ double* slot_ptr(double* part, uint32_t slot) { return part + slot * kMaxPartials; }
__host__ __device__ __forceinline__ const double* slot_ptr(const double* part, uint32_t slot)
{
  return part + slot * kMaxPartials;
}

__device__ __forceinline__ void record_history(KState* st, uint32_t iter, double v)
{
  if (st->history && iter >= 1 && iter <= st->history_len) st->history[iter - 1] = v;
}

// ------------------------------------------------------------------ epilogues

// CG: Ap = A p; <Ap,Ap> (or <Ap, D^-1 Ap>) and <p,Ap>  (host_based/iterative_operations.hpp:83-97)
template <typename T, bool Jac>
struct EpiCg {
  using acc_t = double;
  static constexpr int K = 2;
  T* __restrict__ Ap;
  const T* __restrict__ p;
  const T* __restrict__ d;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t* acc) const
  {
    Ap[row] = s;
    if (Jac)
      acc[0] += static_cast<double>(s * (s / d[row]));
    else
      acc[0] += static_cast<double>(s * s);
    acc[1] += static_cast<double>(p[row] * s);
  }
};

// BiCGStab: t = A p (left-preconditioned: t = D^-1 A p); Full: <t,t>, <p,t>, <t,r0*>; else <t,r0*> only.
template <typename T, bool Jac, bool Full>
struct EpiBicg {
  using acc_t = double;
  static constexpr int K = Full ? 3 : 1;
  static constexpr bool kDiaGrouped = false;  // uncoded 128^3 DIA operator: 38 ms counted, 41-43 ms grouped
  T* __restrict__ out;
  const T* __restrict__ p;
  const T* __restrict__ r0;
  const T* __restrict__ d;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t* acc) const
  {
    const T t = Jac ? s / d[row] : s;
    out[row] = t;
    if (Full) {
      acc[0] += static_cast<double>(t * t);
      acc[1] += static_cast<double>(p[row] * t);
      acc[2] += static_cast<double>(t * r0[row]);
    } else {
      acc[0] += static_cast<double>(t * r0[row]);
    }
  }
};

// BiCGStab restart (bicgstab.hpp:434-445): r = M^-1 (b - A x); p = r0* = r; <r,r>
template <typename T, bool Jac>
struct EpiRestart {
  using acc_t = double;
  static constexpr int K = 1;
  const T* __restrict__ b;
  const T* __restrict__ d;
  T* __restrict__ r;
  T* __restrict__ p;
  T* __restrict__ r0;
  __device__ __forceinline__ void operator()(uint32_t row, T s, acc_t* acc) const
  {
    T v = b[row] - s;
    if (Jac) v = v / d[row];
    r[row] = v;
    p[row] = v;
    r0[row] = v;
    acc[0] += static_cast<double>(v * v);
  }
};

// ------------------------------------------------------------------ CG kernels

// x = 0; r = b; p = M^-1 b; partial <b, M^-1 b>  (cg.hpp:138-149, :268-279)
template <typename T, bool Jac, int W>
__global__ __launch_bounds__(kBlock) void cg_init_kernel(uint32_t n, const T* __restrict__ b, T* __restrict__ x,
                                                         T* __restrict__ r, T* __restrict__ p,
                                                         const T* __restrict__ d, double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vb = ldp<V>(b, i);
    Pack<V, T> z = vb, zero;
    if constexpr (Jac) {
      const Pack<V, T> vd = ldp<V>(d, i);
#pragma unroll
      for (int k = 0; k < V; ++k) z.v[k] = vb.v[k] / vd.v[k];
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
      zero.v[k] = 0;
      acc[0] += static_cast<double>(vb.v[k] * z.v[k]);
    }
    stp<V>(x, i, zero);
    stp<V>(r, i, vb);
    stp<V>(p, i, z);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

// cg.hpp:149-156 (pipelined start) / :279-287 (preconditioned start)
template <bool Jac>
__global__ __launch_bounds__(kBlock) void cg_setup_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                          uint32_t g_vec, uint32_t g_prod)
{
  double bb, ApAp, pAp;
  block_sum_partials3(slot_ptr(part, 0), g_vec, slot_ptr(part, 1), g_prod, slot_ptr(part, 2), g_prod, bb, ApAp, pAp);
  if (threadIdx.x != 0) return;
  double nrs;
  if (Jac) {
    nrs = bb;
  } else {
    const double nrm = sqrt(bb);  // norm_2 then squared, cg.hpp:149
    nrs = nrm * nrm;
  }
  st->nrs = nrs;
  st->rr = nrs;
  st->iter = 0;
  const double at2 = st->abs_tol * st->abs_tol;
  if ((Jac ? fabs(nrs) : nrs) <= at2) {
    st->flags = kFlagDone;
    st->converged = 1;
    return;
  }
  const double alpha = nrs / pAp;
  double beta;
  if (Jac) {
    beta = (alpha * alpha * ApAp - nrs) / nrs;
  } else {
    const double nap = sqrt(ApAp);
    beta = (alpha * alpha * nap * nap - nrs) / nrs;
  }
  st->alpha = alpha;
  st->beta = beta;
  if (st->max_iters == 0) st->flags = kFlagDone;
}

// x += alpha p; r -= alpha Ap; z = M^-1 r; p = z + beta p; partial <r,z>  (host_based/iterative_operations.hpp:402-415)
template <typename T, bool Jac, int W>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(uint32_t n, T* __restrict__ x, T* __restrict__ p,
                                                           T* __restrict__ r, const T* __restrict__ Ap,
                                                           const T* __restrict__ d, const KState* __restrict__ st,
                                                           double* __restrict__ part)
{
  if (st->flags & kFlagDone) return;
  const T alpha = static_cast<T>(st->alpha);
  const T beta = static_cast<T>(st->beta);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vp = ldp<V>(p, i);
    Pack<V, T> vr = ldp<V>(r, i);
    Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> va = ldp<V>(Ap, i);
    Pack<V, T> vd;
    if constexpr (Jac) vd = ldp<V>(d, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += alpha * vp.v[k];
      vr.v[k] -= alpha * va.v[k];
      T z = vr.v[k];
      if constexpr (Jac) z = vr.v[k] / vd.v[k];
      vp.v[k] = z + beta * vp.v[k];
      acc[0] += static_cast<double>(vr.v[k] * z);
    }
    stp<V>(x, i, vx);
    stp<V>(p, i, vp);
    stp<V>(r, i, vr);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

// cg.hpp:170-180: reduce, test, next alpha/beta
__global__ __launch_bounds__(kBlock) void cg_iter_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                         uint32_t g_vec, uint32_t g_prod)
{
  if (st->flags & kFlagDone) return;
  double rr, ApAp, pAp;
  block_sum_partials3(slot_ptr(part, 0), g_vec, slot_ptr(part, 1), g_prod, slot_ptr(part, 2), g_prod, rr, ApAp, pAp);
  if (threadIdx.x != 0) return;
  const uint32_t iter = st->iter + 1;
  st->iter = iter;
  st->rr = rr;
  record_history(st, iter, rr);
  if (fabs(rr / st->nrs) < st->tol * st->tol || fabs(rr) < st->abs_tol * st->abs_tol) {
    st->flags |= kFlagDone;
    st->converged = 1;
    return;
  }
  if (iter >= st->max_iters) {
    st->flags |= kFlagDone;
    return;
  }
  const double alpha = rr / pAp;
  st->alpha = alpha;
  st->beta = (alpha * alpha * ApAp - rr) / rr;
}

// Compact CG step for small systems, one kernel instead of two: the scalar step of iteration k
// (cg.hpp:170-180; every workgroup reduces the few partials itself, in the same fixed order, so
// all of them take the same decision and the same alpha, beta) followed by the vector update of
// iteration k + 1.  The <r,z> partials of update k are in slot 0 or 3 by the parity of k and the
// update writes the other one; likewise the iteration count, so that no workgroup reads what
// another one writes in this kernel.  Workgroup 0 stores the state.
template <typename T, bool Jac, int W>
__global__ __launch_bounds__(kBlock) void cg_step_kernel(uint32_t n, T* __restrict__ x, T* __restrict__ p,
                                                         T* __restrict__ r, const T* __restrict__ Ap,
                                                         const T* __restrict__ d, KState* __restrict__ st,
                                                         double* __restrict__ part, uint32_t par, uint32_t g_vec,
                                                         uint32_t g_prod)
{
  if (st->flags & kFlagDone) return;
  double rr, ApAp, pAp;
  block_sum_partials3(slot_ptr(part, par ? 3u : 0u), g_vec, slot_ptr(part, 1), g_prod, slot_ptr(part, 2), g_prod,
                      rr, ApAp, pAp);
  const uint32_t iter = st->iter2[par] + 1;
  const bool conv = fabs(rr / st->nrs) < st->tol * st->tol || fabs(rr) < st->abs_tol * st->abs_tol;
  const bool done = conv || iter >= st->max_iters;
  const double alpha = rr / pAp;
  const double beta = (alpha * alpha * ApAp - rr) / rr;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = iter;
    st->iter2[par ^ 1u] = iter;
    st->rr = rr;
    record_history(st, iter, rr);
    if (done) {
      if (conv) st->converged = 1;
      st->flags |= kFlagDone;
    } else {
      st->alpha = alpha;
      st->beta = beta;
    }
  }
  if (done) return;
  const T ta = static_cast<T>(alpha);
  const T tb = static_cast<T>(beta);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vp = ldp<V>(p, i);
    Pack<V, T> vr = ldp<V>(r, i);
    Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> va = ldp<V>(Ap, i);
    Pack<V, T> vd;
    if constexpr (Jac) vd = ldp<V>(d, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += ta * vp.v[k];
      vr.v[k] -= ta * va.v[k];
      T z = vr.v[k];
      if constexpr (Jac) z = vr.v[k] / vd.v[k];
      vp.v[k] = z + tb * vp.v[k];
      acc[0] += static_cast<double>(vr.v[k] * z);
    }
    stp<V>(x, i, vx);
    stp<V>(p, i, vp);
    stp<V>(r, i, vr);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, par ? 0u : 3u), kMaxPartials, lw);
}

// One kernel per CG iteration for small systems on a DIA operator: the scalar step of iteration k
// (as cg_step_kernel), the vector update of iteration k + 1 on this workgroup's vector chunk, and the
// product A p_{k+1} with its inner products on this workgroup's product chunk.  The product needs
// p_{k+1} of rows that other workgroups update in this same kernel, so it forms them itself from the
// previous iteration's r, A p and p with the update's own expressions (the same values bit for bit);
// r, p and A p are kept twice, read from the `in` copies and written to the `out` ones.  Workgroup
// lw < g_vec takes the update kernel's chunk lw and lw < g_prod the DIA product kernel's chunk lw,
// so every partial sum is the one of the two-kernel compact mode and the solve is the same bit for
// bit.  Partial slots by the parity of k: <r,z> of update k in 0 / 3 (update k + 1 writes the
// other), <Ap,Ap> and <p,Ap> of product k in 1, 2 / 4, 5 (product k + 1 writes the other pair).
//
// The loads that do not depend on this iteration's scalars -- the partial sums, the first
// kFusedPacks packs of the lane's share of the update and the first tile of the product (the
// values of up to kDiaGroup diagonals and the r, A p, p entries they gather) -- are all issued
// before the partial sums are reduced, so that they are in flight together: on a system held in the
// caches the iteration is then about one memory round trip long instead of four.
constexpr int kFusedPacks = 2;
constexpr uint32_t kFusedPartials = 4;

template <typename T, bool Jac, int W>
__global__ __launch_bounds__(kBlock) void cg_fused_kernel(uint32_t n, T* __restrict__ x, const T* __restrict__ r_in,
                                                          T* __restrict__ r_out, const T* __restrict__ p_in,
                                                          T* __restrict__ p_out, const T* __restrict__ Ap_in,
                                                          T* __restrict__ Ap_out, const T* __restrict__ d, DiaDev<T> A,
                                                          KState* __restrict__ st, double* __restrict__ part,
                                                          uint32_t par, uint32_t g_vec, uint32_t g_prod)
{
  // the state is read with everything else and tested once the loads are in flight
  const uint32_t flags = st->flags;
  const uint32_t iter_prev = st->iter2[par];
  const uint32_t max_iters = st->max_iters;
  const double nrs = st->nrs, tol = st->tol, abs_tol = st->abs_tol;
  const uint32_t tid = threadIdx.x;
  const uint32_t lw = xcd_logical_block();

  // ---- loads issued ahead of the scalar step: first the partial sums it reduces (up to
  // kFusedPartials per slot and lane; lane t takes entries t, t + 256, ... as block_sum_partials_k)
  const double* const ps[3] = {slot_ptr(part, par ? 3u : 0u), slot_ptr(part, par ? 4u : 1u),
                               slot_ptr(part, par ? 5u : 2u)};
  const uint32_t pc[3] = {g_vec, g_prod, g_prod};
  double pv[kFusedPartials][3];
#pragma unroll
  for (uint32_t b = 0; b < kFusedPartials; ++b)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      // unconditional (clamped) loads, so that no sum is placed next to its load and waits for it
      const uint32_t i = tid + b * kBlock;
      pv[b][q] = ps[q][i < pc[q] ? i : pc[q] - 1];
    }
  const bool has_vec = lw < g_vec;
  uint32_t c0 = 0, c1 = 0;
  if (has_vec) chunk_range_w(n, W, lw, g_vec, c0, c1);
  const uint32_t packed_end = c0 + ((c1 - c0) / W) * W;
  Pack<W, T> fp[kFusedPacks], fr[kFusedPacks], fx[kFusedPacks], fa[kFusedPacks], fd[kFusedPacks];
#pragma unroll
  for (int k = 0; k < kFusedPacks; ++k) {
    const uint32_t i = c0 + W * tid + static_cast<uint32_t>(k) * W * kBlock;
    if (has_vec && i < packed_end) {
      fp[k] = ldp<W>(p_in, i);
      fr[k] = ldp<W>(r_in, i);
      fx[k] = ldp<W>(x, i);
      fa[k] = ldp<W>(Ap_in, i);
      if constexpr (Jac) fd[k] = ldp<W>(d, i);
    }
  }
  const bool has_prod = lw < g_prod;
  const uint32_t tiles = (n + kBlock - 1) / kBlock;
  uint32_t t0 = 0, t1 = 0;
  if (has_prod) chunk_range(tiles, lw, g_prod, t0, t1);
  const uint32_t row0 = t0 * kBlock + tid;
  const bool first_row = has_prod && t0 < t1 && row0 < n;
  T gv[kDiaGroup], gr[kDiaGroup], ga[kDiaGroup], gp[kDiaGroup], gd[kDiaGroup];
  T own_r = 0, own_a = 0, own_p = 0, own_d = 1;
  if (first_row) {
#pragma unroll
    for (uint32_t j = 0; j < kDiaGroup; ++j) {
      gv[j] = T(0);
      gr[j] = ga[j] = gp[j] = T(0);
      gd[j] = T(1);
      if (j < A.nd) {
        const uint32_t c = row0 + static_cast<uint32_t>(A.off[j]);  // out of range (also below 0): no entry
        gv[j] = A.val[static_cast<size_t>(j) * A.n_rows + row0];
        if (c < A.n_rows) {
          gr[j] = r_in[c];
          ga[j] = Ap_in[c];
          gp[j] = p_in[c];
          if constexpr (Jac) gd[j] = d[c];
        }
      }
    }
    own_r = r_in[row0];
    own_a = Ap_in[row0];
    own_p = p_in[row0];
    if constexpr (Jac) own_d = d[row0];
  }

  if (flags & kFlagDone) return;  // the same value for every lane: all return or none

  // ---- the scalar step of iteration k: the partial sums added in block_sum_partials_k's order
  double sums[3] = {0, 0, 0}, red3[3];
#pragma unroll
  for (uint32_t b = 0; b < kFusedPartials; ++b)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (tid + b * kBlock < pc[q]) sums[q] += pv[b][q];
  const uint32_t pmax = g_vec > g_prod ? g_vec : g_prod;
  for (uint32_t i = tid + kFusedPartials * kBlock; i < pmax; i += kBlock)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (i < pc[q]) sums[q] += ps[q][i];
  block_combine_k<3, double>(sums, red3);
  const double rr = red3[0], ApAp = red3[1], pAp = red3[2];
  const uint32_t iter = iter_prev + 1;
  const bool conv = fabs(rr / nrs) < tol * tol || fabs(rr) < abs_tol * abs_tol;
  const bool done = conv || iter >= max_iters;
  const double alpha = rr / pAp;
  const double beta = (alpha * alpha * ApAp - rr) / rr;
  if (blockIdx.x == 0 && tid == 0) {
    st->iter = iter;
    st->iter2[par ^ 1u] = iter;
    st->rr = rr;
    record_history(st, iter, rr);
    if (done) {
      if (conv) st->converged = 1;
      st->flags |= kFlagDone;
    } else {
      st->alpha = alpha;
      st->beta = beta;
    }
  }
  if (done) return;
  const T ta = static_cast<T>(alpha);
  const T tb = static_cast<T>(beta);

  // ---- update k + 1 (cg_step_kernel's update, chunk lw of g_vec, in for_packs' order)
  if (has_vec) {
    double acc[1] = {0};
    auto update = [&](uint32_t i, auto& vp, auto& vr, auto& vx, const auto& va, const auto& vd) {
      constexpr int V = sizeof(vp.v) / sizeof(T);
#pragma unroll
      for (int k = 0; k < V; ++k) {
        vx.v[k] += ta * vp.v[k];
        vr.v[k] -= ta * va.v[k];
        T z = vr.v[k];
        if constexpr (Jac) z = vr.v[k] / vd.v[k];
        vp.v[k] = z + tb * vp.v[k];
        acc[0] += static_cast<double>(vr.v[k] * z);
      }
      stp<V>(x, i, vx);
      stp<V>(p_out, i, vp);
      stp<V>(r_out, i, vr);
    };
#pragma unroll
    for (int k = 0; k < kFusedPacks; ++k) {
      const uint32_t i = c0 + W * tid + static_cast<uint32_t>(k) * W * kBlock;
      if (i < packed_end) update(i, fp[k], fr[k], fx[k], fa[k], fd[k]);
    }
    for (uint32_t i = c0 + W * tid + kFusedPacks * W * kBlock; i < packed_end; i += W * kBlock) {
      Pack<W, T> vp = ldp<W>(p_in, i), vr = ldp<W>(r_in, i), vx = ldp<W>(x, i);
      const Pack<W, T> va = ldp<W>(Ap_in, i);
      Pack<W, T> vd;
      if constexpr (Jac) vd = ldp<W>(d, i);
      update(i, vp, vr, vx, va, vd);
    }
    if (W > 1) {
      for (uint32_t i = packed_end + tid; i < c1; i += kBlock) {
        Pack<1, T> vp = ldp<1>(p_in, i), vr = ldp<1>(r_in, i), vx = ldp<1>(x, i);
        const Pack<1, T> va = ldp<1>(Ap_in, i);
        Pack<1, T> vd;
        if constexpr (Jac) vd = ldp<1>(d, i);
        update(i, vp, vr, vx, va, vd);
      }
    }
    block_store_partials<1, double>(acc, slot_ptr(part, par ? 0u : 3u), kMaxPartials, lw);
  }

  // ---- A p_{k+1} with EpiCg (dia_kernel's sum, chunk lw of g_prod)
  if (has_prod) {
    auto next = [&](T rj, T aj, T pj, T dj) -> T {
      rj -= ta * aj;
      T z = rj;
      if constexpr (Jac) z = rj / dj;
      return z + tb * pj;
    };
    auto p_next = [&](uint32_t j) -> T { return next(r_in[j], Ap_in[j], p_in[j], Jac ? d[j] : T(1)); };
    double acc[2] = {0, 0};
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t row = t * kBlock + tid;
      if (row < n) {
        T sum = 0;
        uint32_t d_from = 0;
        T p_row;
        if (t == t0) {  // the first group of diagonals from the registers loaded ahead
#pragma unroll
          for (uint32_t j = 0; j < kDiaGroup; ++j) {
            if (j < A.nd) {
              const uint32_t c = row + static_cast<uint32_t>(A.off[j]);
              const T xv = c < A.n_rows ? next(gr[j], ga[j], gp[j], gd[j]) : T(0);
              sum += (gv[j] > 0 || gv[j] < 0) ? xv * gv[j] : T(0);
            }
          }
          d_from = kDiaGroup;
          p_row = next(own_r, own_a, own_p, own_d);
        } else {
          p_row = p_next(row);
        }
        for (uint32_t d0 = d_from; d0 < A.nd; d0 += kDiaGroup) {
          T v[kDiaGroup], xv[kDiaGroup];
#pragma unroll
          for (uint32_t j = 0; j < kDiaGroup; ++j) {
            v[j] = T(0);
            xv[j] = T(0);
            if (d0 + j < A.nd) {
              const uint32_t dd = d0 + j;
              const uint32_t c = row + static_cast<uint32_t>(A.off[dd]);
              v[j] = A.val[static_cast<size_t>(dd) * A.n_rows + row];
              if (c < A.n_rows) xv[j] = p_next(c);
            }
          }
#pragma unroll
          for (uint32_t j = 0; j < kDiaGroup; ++j)
            if (d0 + j < A.nd) sum += (v[j] > 0 || v[j] < 0) ? xv[j] * v[j] : T(0);
        }
        Ap_out[row] = sum;
        if (Jac)
          acc[0] += static_cast<double>(sum * (sum / d[row]));
        else
          acc[0] += static_cast<double>(sum * sum);
        acc[1] += static_cast<double>(p_row * sum);
      }
    }
    block_store_partials<2, double>(acc, slot_ptr(part, par ? 1u : 4u), kMaxPartials, lw);
  }
}

// ------------------------------------------------------------------ BiCGStab kernels

// x = 0; r = p = r0* = M^-1 b; partials <b,b>, <r,r>  (bicgstab.hpp:105-131, :408-445)
template <typename T, bool Jac, int W>
__global__ __launch_bounds__(kBlock) void bicg_init_kernel(uint32_t n, const T* __restrict__ b, T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ p, T* __restrict__ r0,
                                                           const T* __restrict__ d, double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[2] = {0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vb = ldp<V>(b, i);
    Pack<V, T> v = vb, zero;
    if constexpr (Jac) {
      const Pack<V, T> vd = ldp<V>(d, i);
#pragma unroll
      for (int k = 0; k < V; ++k) v.v[k] = vb.v[k] / vd.v[k];
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
      zero.v[k] = 0;
      acc[0] += static_cast<double>(vb.v[k] * vb.v[k]);
      acc[1] += static_cast<double>(v.v[k] * v.v[k]);
    }
    stp<V>(x, i, zero);
    stp<V>(r, i, v);
    stp<V>(p, i, v);
    stp<V>(r0, i, v);
  });
  block_store_partials<2, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

__global__ __launch_bounds__(kBlock) void bicg_setup_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                            uint32_t g_vec)
{
  double sums[2];
  block_sum_partials_k<2, double>({slot_ptr(part, 0), slot_ptr(part, 1)}, {g_vec, g_vec}, sums);
  const double bb = sums[0], vv = sums[1];
  if (threadIdx.x != 0) return;
  const double norm_rhs = sqrt(bb);
  const double nv = sqrt(vv);
  st->nrs = norm_rhs;
  st->res = norm_rhs;
  st->rr = nv * nv;  // norm_2 then squared (bicgstab.hpp:131, :441-442)
  st->rr_from_partials = 0;
  st->iter = 0;
  if (norm_rhs <= st->abs_tol) {
    st->flags = kFlagDone;
    st->converged = 1;
    return;
  }
  if (st->max_iters == 0) st->flags = kFlagDone;
}

// alpha = <r,r0*> / <Ap,r0*>  (host_based/iterative_operations.hpp:536-544).  Also turns a pending
// finish into done, so the kernels of the iteration after the last one all return at once.
__global__ __launch_bounds__(kBlock) void bicg_alpha_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                            uint32_t g_vec, uint32_t g_prod, int detect_breakdown)
{
  const uint32_t flags = st->flags;
  if (flags & kFlagDone) return;
  if (flags & kFlagFinishing) {
    if (threadIdx.x == 0) st->flags = flags | kFlagDone;
    return;
  }
  const bool from_part = st->rr_from_partials != 0;
  double sums[2];
  block_sum_partials_k<2, double>({slot_ptr(part, 0), slot_ptr(part, 2)}, {from_part ? g_vec : 0u, g_prod}, sums);
  const double rsum = sums[0], ap_r0 = sums[1];
  if (threadIdx.x != 0) return;
  const double rr = from_part ? rsum : st->rr;
  st->rr = rr;
  st->ap_r0 = ap_r0;
  if (detect_breakdown && rr <= 0 && rr >= 0) {  // bicgstab.hpp:473: <r,r0*> == 0 asks for a restart
    st->breakdown = 1;
    st->flags = flags | kFlagDone;
    return;
  }
  st->alpha = rr / ap_r0;
}

// s = r - alpha Ap; partial <s,s>  (host_based/iterative_operations.hpp:552-560)
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicg_update_s_kernel(uint32_t n, T* __restrict__ s, const T* __restrict__ r,
                                                               const T* __restrict__ Ap, const KState* __restrict__ st,
                                                               double* __restrict__ part)
{
  if (st->flags & (kFlagDone | kFlagFinishing)) return;
  const T alpha = static_cast<T>(st->alpha);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vr = ldp<V>(r, i);
    const Pack<V, T> va = ldp<V>(Ap, i);
    Pack<V, T> vs;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vs.v[k] = vr.v[k] - alpha * va.v[k];
      acc[0] += static_cast<double>(vs.v[k] * vs.v[k]);
    }
    stp<V>(s, i, vs);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 3), kMaxPartials, lw);
}

// bicgstab.hpp:184-198: reduce, alpha/beta/omega, residual estimate, test
__global__ __launch_bounds__(kBlock) void bicg_scalars_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                              uint32_t g_vec, uint32_t g_prod, int detect_breakdown)
{
  if (st->flags & (kFlagDone | kFlagFinishing)) return;
  double sums[4];
  block_sum_partials_k<4, double>({slot_ptr(part, 3), slot_ptr(part, 4), slot_ptr(part, 5), slot_ptr(part, 6)},
                                  {g_vec, g_prod, g_prod, g_prod}, sums);
  const double s_dot_s = sums[0], As_dot_As = sums[1], As_dot_s = sums[2], As_dot_r0 = sums[3];
  if (threadIdx.x != 0) return;
  const double r_dot_r0 = st->rr;
  const double Ap_dot_r0 = st->ap_r0;
  const double alpha = r_dot_r0 / Ap_dot_r0;
  const double beta = -As_dot_r0 / Ap_dot_r0;
  const double omega = As_dot_s / As_dot_As;
  const double res = sqrt(s_dot_s - 2.0 * omega * As_dot_s + omega * omega * As_dot_As);
  st->alpha = alpha;
  st->beta = beta;
  st->omega = omega;
  st->res = res;
  const uint32_t iter = st->iter + 1;
  st->iter = iter;
  record_history(st, iter, res);
  st->rr_from_partials = 1;
  if (fabs(res / st->nrs) < st->tol || res < st->abs_tol) {
    st->final_mode = 1;
    st->converged = 1;
    st->flags |= kFlagFinishing;
  } else if (iter >= st->max_iters) {
    st->final_mode = 2;
    st->flags |= kFlagFinishing;
  } else {
    st->final_mode = 0;
    if (detect_breakdown && omega <= 0 && omega >= 0) {  // bicgstab.hpp:473: omega == 0 asks for a restart
      st->breakdown = 1;
      st->flags |= kFlagFinishing;
    }
  }
}

// x += alpha p + omega s; r = s - omega As; p = r + beta (p - omega Ap); partial <r,r0*>
// (host_based/iterative_operations.hpp:598-617).  final_mode 1 applies only the x update.
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicg_vec_update_kernel(uint32_t n, T* __restrict__ x, T* __restrict__ p,
                                                                 const T* __restrict__ s, T* __restrict__ r,
                                                                 const T* __restrict__ As, const T* __restrict__ Ap,
                                                                 const T* __restrict__ r0,
                                                                 const KState* __restrict__ st,
                                                                 double* __restrict__ part)
{
  if (st->flags & kFlagDone) return;
  const T alpha = static_cast<T>(st->alpha);
  const T beta = static_cast<T>(st->beta);
  const T omega = static_cast<T>(st->omega);
  const uint32_t mode = st->final_mode;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  if (mode == 1) {
    for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
      constexpr int V = decltype(w)::value;
      Pack<V, T> vx = ldp<V>(x, i);
      const Pack<V, T> vp = ldp<V>(p, i);
      const Pack<V, T> vs = ldp<V>(s, i);
#pragma unroll
      for (int k = 0; k < V; ++k) vx.v[k] += alpha * vp.v[k] + omega * vs.v[k];
      stp<V>(x, i, vx);
    });
    return;
  }
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vx = ldp<V>(x, i);
    Pack<V, T> vp = ldp<V>(p, i);
    const Pack<V, T> vs = ldp<V>(s, i);
    const Pack<V, T> vas = ldp<V>(As, i);
    const Pack<V, T> vap = ldp<V>(Ap, i);
    const Pack<V, T> vr0 = ldp<V>(r0, i);
    Pack<V, T> vr;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += alpha * vp.v[k] + omega * vs.v[k];
      vr.v[k] = vs.v[k] - omega * vas.v[k];
      vp.v[k] = vr.v[k] + beta * (vp.v[k] - omega * vap.v[k]);
      acc[0] += static_cast<double>(vr.v[k] * vr0.v[k]);
    }
    stp<V>(x, i, vx);
    stp<V>(r, i, vr);
    stp<V>(p, i, vp);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

// ---- compact BiCGStab for small systems (unpreconditioned): four kernels per iteration instead of
// six.  The alpha step runs at the head of the s update and the scalar step (alpha, beta, omega,
// residual estimate, test) at the head of the vector update: every workgroup reduces the few partial
// sums itself with block_sum_partials_k (the order of the scalar kernels, so every workgroup takes
// the same decision with the same scalars, and the solve is the six-kernel one bit for bit);
// workgroup 0 stores the state.  No workgroup reads a state field that another one writes in the
// same kernel, except the flags at the start, where either value leads to the same action; the
// iteration count is read from iter2[k & 1] and written to the other one.

// bicg_alpha_kernel + bicg_update_s_kernel
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicg_update_s_compact_kernel(uint32_t n, T* __restrict__ s,
                                                                       const T* __restrict__ r,
                                                                       const T* __restrict__ Ap,
                                                                       KState* __restrict__ st,
                                                                       double* __restrict__ part, uint32_t g_vec,
                                                                       uint32_t g_prod)
{
  const uint32_t flags = st->flags;
  if (flags & (kFlagDone | kFlagFinishing)) {
    // a pending finish becomes done (workgroup 0; the others return on either value)
    if (!(flags & kFlagDone) && blockIdx.x == 0 && threadIdx.x == 0) st->flags = flags | kFlagDone;
    return;
  }
  const bool from_part = st->rr_from_partials != 0;
  double sums[2];
  block_sum_partials_k<2, double>({slot_ptr(part, 0), slot_ptr(part, 2)}, {from_part ? g_vec : 0u, g_prod}, sums);
  const double rr = from_part ? sums[0] : st->rr;
  const double ap_r0 = sums[1];
  const double alpha_d = rr / ap_r0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (from_part) st->rr = rr;  // otherwise it is the value every workgroup read
    st->ap_r0 = ap_r0;
    st->alpha = alpha_d;
  }
  const T alpha = static_cast<T>(alpha_d);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, T> vr = ldp<V>(r, i);
    const Pack<V, T> va = ldp<V>(Ap, i);
    Pack<V, T> vs;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vs.v[k] = vr.v[k] - alpha * va.v[k];
      acc[0] += static_cast<double>(vs.v[k] * vs.v[k]);
    }
    stp<V>(s, i, vs);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 3), kMaxPartials, lw);
}

// bicg_scalars_kernel + bicg_vec_update_kernel.  A finish flag seen at the start can only be the one
// workgroup 0 sets in this kernel (the s update turns an earlier one into done), so only done stops.
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicg_vec_update_compact_kernel(
    uint32_t n, T* __restrict__ x, T* __restrict__ p, const T* __restrict__ s, T* __restrict__ r,
    const T* __restrict__ As, const T* __restrict__ Ap, const T* __restrict__ r0, KState* __restrict__ st,
    double* __restrict__ part, uint32_t par, uint32_t g_vec, uint32_t g_prod)
{
  if (st->flags & kFlagDone) return;
  double sums[4];
  block_sum_partials_k<4, double>({slot_ptr(part, 3), slot_ptr(part, 4), slot_ptr(part, 5), slot_ptr(part, 6)},
                                  {g_vec, g_prod, g_prod, g_prod}, sums);
  const double s_dot_s = sums[0], As_dot_As = sums[1], As_dot_s = sums[2], As_dot_r0 = sums[3];
  const double r_dot_r0 = st->rr;
  const double Ap_dot_r0 = st->ap_r0;
  const double alpha_d = r_dot_r0 / Ap_dot_r0;
  const double beta_d = -As_dot_r0 / Ap_dot_r0;
  const double omega_d = As_dot_s / As_dot_As;
  const double res = sqrt(s_dot_s - 2.0 * omega_d * As_dot_s + omega_d * omega_d * As_dot_As);
  const uint32_t iter = st->iter2[par] + 1;
  const bool conv = fabs(res / st->nrs) < st->tol || res < st->abs_tol;
  const uint32_t mode = conv ? 1u : (iter >= st->max_iters ? 2u : 0u);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha_d;
    st->beta = beta_d;
    st->omega = omega_d;
    st->res = res;
    st->iter = iter;
    st->iter2[par ^ 1u] = iter;
    record_history(st, iter, res);
    st->rr_from_partials = 1;
    st->final_mode = mode;
    if (conv) st->converged = 1;
    if (mode != 0) st->flags |= kFlagFinishing;
  }
  const T alpha = static_cast<T>(alpha_d);
  const T beta = static_cast<T>(beta_d);
  const T omega = static_cast<T>(omega_d);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  if (mode == 1) {
    for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
      constexpr int V = decltype(w)::value;
      Pack<V, T> vx = ldp<V>(x, i);
      const Pack<V, T> vp = ldp<V>(p, i);
      const Pack<V, T> vs = ldp<V>(s, i);
#pragma unroll
      for (int k = 0; k < V; ++k) vx.v[k] += alpha * vp.v[k] + omega * vs.v[k];
      stp<V>(x, i, vx);
    });
    return;
  }
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vx = ldp<V>(x, i);
    Pack<V, T> vp = ldp<V>(p, i);
    const Pack<V, T> vs = ldp<V>(s, i);
    const Pack<V, T> vas = ldp<V>(As, i);
    const Pack<V, T> vap = ldp<V>(Ap, i);
    const Pack<V, T> vr0 = ldp<V>(r0, i);
    Pack<V, T> vr;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += alpha * vp.v[k] + omega * vs.v[k];
      vr.v[k] = vs.v[k] - omega * vas.v[k];
      vp.v[k] = vr.v[k] + beta * (vp.v[k] - omega * vap.v[k]);
      acc[0] += static_cast<double>(vr.v[k] * vr0.v[k]);
    }
    stp<V>(x, i, vx);
    stp<V>(r, i, vr);
    stp<V>(p, i, vp);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

// ---- left-preconditioned BiCGStab with the reference's recurrences (bicgstab.hpp:447-480):
// explicit <r,r0*> and ||r|| after the x/r update, beta = <r,r0*>_new / <r,r0*> * alpha / omega.

// alpha = ip_rr0star / <M^-1 A p, r0*>  (bicgstab.hpp:448-450)
__global__ __launch_bounds__(kBlock) void bicgc_alpha_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_prod)
{
  if (st->flags & kFlagDone) return;
  const double t0r0 = block_sum_partials(slot_ptr(part, 2), g_prod);
  if (threadIdx.x != 0) return;
  st->alpha = st->rr / t0r0;
}

// omega = <t1,s> / (||t1|| ||t1||)  (bicgstab.hpp:454-457)
__global__ __launch_bounds__(kBlock) void bicgc_omega_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_prod)
{
  if (st->flags & kFlagDone) return;
  double sums[2];
  block_sum_partials_k<2, double>({slot_ptr(part, 4), slot_ptr(part, 5)}, {g_prod, g_prod}, sums);
  const double t1t1 = sums[0], t1s = sums[1];
  if (threadIdx.x != 0) return;
  const double nrm = sqrt(t1t1);
  st->omega = t1s / (nrm * nrm);
}

// x += alpha p + omega s; r = s - omega t1; partials <r,r>, <r,r0*>  (bicgstab.hpp:459-468)
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicgc_xr_kernel(uint32_t n, T* __restrict__ x, const T* __restrict__ p,
                                                          const T* __restrict__ s, T* __restrict__ r,
                                                          const T* __restrict__ t1, const T* __restrict__ r0,
                                                          const KState* __restrict__ st, double* __restrict__ part)
{
  if (st->flags & kFlagDone) return;
  const T alpha = static_cast<T>(st->alpha);
  const T omega = static_cast<T>(st->omega);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[2] = {0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vx = ldp<V>(x, i);
    const Pack<V, T> vp = ldp<V>(p, i);
    const Pack<V, T> vs = ldp<V>(s, i);
    const Pack<V, T> vt = ldp<V>(t1, i);
    const Pack<V, T> vr0 = ldp<V>(r0, i);
    Pack<V, T> vr;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += alpha * vp.v[k] + omega * vs.v[k];
      vr.v[k] = vs.v[k] - omega * vt.v[k];
      acc[0] += static_cast<double>(vr.v[k] * vr.v[k]);
      acc[1] += static_cast<double>(vr.v[k] * vr0.v[k]);
    }
    stp<V>(x, i, vx);
    stp<V>(r, i, vr);
  });
  block_store_partials<2, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

// ||r||, test, beta, restart request  (bicgstab.hpp:462-474)
__global__ __launch_bounds__(kBlock) void bicgc_check_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_vec)
{
  if (st->flags & kFlagDone) return;
  double sums[2];
  block_sum_partials_k<2, double>({slot_ptr(part, 0), slot_ptr(part, 1)}, {g_vec, g_vec}, sums);
  const double rr2 = sums[0], rr0 = sums[1];
  if (threadIdx.x != 0) return;
  const double res = sqrt(rr2);
  const uint32_t iter = st->iter + 1;
  st->iter = iter;
  st->res = res;
  record_history(st, iter, res);
  if (res / st->nrs < st->tol || res < st->abs_tol) {
    st->flags |= kFlagDone;
    st->converged = 1;
    return;
  }
  if (iter >= st->max_iters) {
    st->flags |= kFlagDone;
    return;
  }
  const double omega = st->omega;
  st->beta = rr0 / st->rr * st->alpha / omega;
  st->rr = rr0;
  if ((rr0 <= 0 && rr0 >= 0) || (omega <= 0 && omega >= 0)) {  // degenerate direction: restart (:473)
    st->breakdown = 1;
    st->flags |= kFlagDone;
  }
}

// p = r + beta (p - omega t0), in the reference's two steps (bicgstab.hpp:479-480)
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void bicgc_p_kernel(uint32_t n, T* __restrict__ p, const T* __restrict__ r,
                                                         const T* __restrict__ t0, const KState* __restrict__ st)
{
  if (st->flags & kFlagDone) return;
  const T beta = static_cast<T>(st->beta);
  const T omega = static_cast<T>(st->omega);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, T> vp = ldp<V>(p, i);
    const Pack<V, T> vr = ldp<V>(r, i);
    const Pack<V, T> vt = ldp<V>(t0, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const T pm = vp.v[k] - vt.v[k] * omega;
      vp.v[k] = vr.v[k] + pm * beta;
    }
    stp<V>(p, i, vp);
  });
}

// restart: <r,r> -> rr (norm_2 then squared, bicgstab.hpp:441-442)
__global__ __launch_bounds__(kBlock) void bicg_restart_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                              uint32_t g_prod)
{
  if (st->flags & (kFlagDone | kFlagFinishing)) return;
  const double vv = block_sum_partials(slot_ptr(part, 7), g_prod);
  if (threadIdx.x != 0) return;
  const double nv = sqrt(vv);
  st->rr = nv * nv;
  st->rr_from_partials = 0;
}

// Diagonal of A for Jacobi (linalg/detail row_info with SPARSE_ROW_DIAGONAL), failing on a zero entry.
template <typename T>
__global__ __launch_bounds__(kBlock) void count_zero_kernel(uint32_t n, const T* __restrict__ d,
                                                            uint32_t* __restrict__ count)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n && !(d[i] < 0 || d[i] > 0)) atomicAdd(count, 1u);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void sell_diag_kernel(SellDev<T> A, T* __restrict__ d)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= A.n_rows) return;
  const uint32_t s = row / A.C;
  const uint32_t lane = row - s * A.C;
  const size_t base = static_cast<size_t>(A.start[s]) + lane;
  T v = 0;
  for (uint32_t k = 0; k < A.cpb[s]; ++k) {
    const size_t off = base + static_cast<size_t>(k) * A.C;
    const bool on_diag = A.dcol ? A.dcol[off] == 0 : A.col[off] == row;
    if (on_diag && (A.val[off] < 0 || A.val[off] > 0)) { v = A.val[off]; break; }
  }
  d[row] = v;
}

// COO (row_info, host_based/sparse_matrix_operations.hpp:1156-1205): the row's last entry with
// column == row, whatever its value; 0 when there is none.
template <typename T>
__global__ __launch_bounds__(kBlock) void coo_diag_kernel(CsrDev<T> A, T* __restrict__ d)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= A.n_rows) return;
  T v = 0;
  for (uint32_t e = A.row_ptr[row]; e < A.row_ptr[row + 1]; ++e)
    if (A.col[2 * static_cast<size_t>(e)] == row) v = A.val[e];
  d[row] = v;
}

// ELL / HYB: the first nonzero entry with column == row (ELL part, then the CSR part).
template <typename T, bool Hyb>
__global__ __launch_bounds__(kBlock) void ell_diag_kernel(EllDev<T> A, T* __restrict__ d)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= A.n_rows) return;
  T v = 0;
  bool found = false;
  for (uint32_t k = 0; k < A.width && !found; ++k) {
    const size_t off = static_cast<size_t>(k) * A.ld + row;
    if (A.coords[off] == row && (A.val[off] < 0 || A.val[off] > 0)) {
      v = A.val[off];
      found = true;
    }
  }
  if constexpr (Hyb)
    for (uint32_t e = A.csr_rows[row]; e < A.csr_rows[row + 1] && !found; ++e)
      if (A.csr_cols[e] == row && (A.csr_val[e] < 0 || A.csr_val[e] > 0)) {
        v = A.csr_val[e];
        found = true;
      }
  d[row] = v;
}

inline int csr_row_info(const vclhip_csr* A, double* d, hipStream_t s) { return vclhip_csr_row_info_f64(A, d, 3, s); }
inline int csr_row_info(const vclhip_csr* A, float* d, hipStream_t s) { return vclhip_csr_row_info_f32(A, d, 3, s); }

// d = the offset-0 diagonal of a DIA matrix (0 where it has no entry).
template <typename T>
__global__ __launch_bounds__(kBlock) void dia_diag_kernel(DiaDev<T> A, T* __restrict__ d)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= A.n_rows) return;
  T v = 0;
  for (uint32_t q = 0; q < A.nd; ++q)
    if (A.off[q] == 0) v = A.val[static_cast<size_t>(q) * A.n_rows + row];
  d[row] = v;
}

// d = diag(A) for any format (0 where a row has no nonzero diagonal entry).
template <typename T>
static int matrix_diag(const DevOp<T>& op, T* d, hipStream_t s)
{
  const dim3 grid((op.n + kBlock - 1) / kBlock);
  switch (op.fmt) {
    case OpFormat::Csr:
      return csr_row_info(op.csr_desc, d, s);
    case OpFormat::Sell:
      hipLaunchKernelGGL(sell_diag_kernel<T>, grid, dim3(kBlock), 0, s, op.sm, d);
      break;
    case OpFormat::Coo:
      hipLaunchKernelGGL(coo_diag_kernel<T>, grid, dim3(kBlock), 0, s, op.csr, d);
      break;
    case OpFormat::Ell:
      hipLaunchKernelGGL((ell_diag_kernel<T, false>), grid, dim3(kBlock), 0, s, op.ell, d);
      break;
    case OpFormat::Hyb:
      hipLaunchKernelGGL((ell_diag_kernel<T, true>), grid, dim3(kBlock), 0, s, op.ell, d);
      break;
    case OpFormat::Dia:
      hipLaunchKernelGGL(dia_diag_kernel<T>, grid, dim3(kBlock), 0, s, op.dia, d);
      break;
  }
  return check_launch("diagonal");
}

template <typename T>
static int jacobi_diag(const DevOp<T>& op, T* d, uint32_t* zero_count, hipStream_t s)
{
  VCLHIP_TRY(matrix_diag(op, d, s));
  VCLHIP_HIP(hipMemsetAsync(zero_count, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(count_zero_kernel<T>, dim3((op.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, op.n, d, zero_count);
  VCLHIP_TRY(check_launch("count_zero_kernel"));
  uint32_t zeros = 0;
  VCLHIP_HIP(hipMemcpyAsync(&zeros, zero_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  if (zeros) {
    set_error("Jacobi preconditioner: %u zero diagonal entries", zeros);
    return VCLHIP_ERROR_ZERO_DIAGONAL;
  }
  return VCLHIP_SUCCESS;
}

// ------------------------------------------------------------------ iteration runner

struct PinnedState {
  KState* h = nullptr;
  ~PinnedState()
  {
    if (h) (void)hipHostFree(h);
  }
};

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~EventPair()
  {
    for (auto ev : e)
      if (ev) (void)hipEventDestroy(ev);
  }
};

struct NoTail {
  int operator()() const { return VCLHIP_SUCCESS; }
};

// Enqueues iterations in chunks and polls the state of the previous chunk.  launch_iter(i) enqueues
// iteration i (0-based); needs_hook(i) says whether iteration i needs extra work before it (a
// restart), which keeps it out of a captured graph; hook(i) enqueues that work.  tail() enqueues
// work that closes every chunk (the mixed-precision defect correction, whose kernels return at
// once unless an iteration of the chunk asked for it).  With no_alloc nothing is allocated: no
// pinned buffer, no events and no graph; the state of each chunk is copied to the host stack and
// waited for before the next chunk is enqueued.
template <class LaunchIter, class NeedsHook, class Hook, class Tail = NoTail>
static int run_iterations(hipStream_t s, KState* st_dev, uint32_t first_iter, uint32_t max_iters,
                          const vclhip_solver_params* P, const void* x_dev, LaunchIter&& launch_iter,
                          NeedsHook&& needs_hook, Hook&& hook, KState& final_state, bool bicg, Tail tail = Tail(),
                          bool no_alloc = false)
{
  PinnedState pin;
  EventPair ev;
  if (!no_alloc) {
    VCLHIP_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin.h), 2 * sizeof(KState), hipHostMallocDefault));
    VCLHIP_HIP(hipEventCreateWithFlags(&ev.e[0], hipEventDisableTiming));
    VCLHIP_HIP(hipEventCreateWithFlags(&ev.e[1], hipEventDisableTiming));
  }

  const bool monitor = P && P->monitor;
  uint32_t chunk = (P && P->chunk) ? P->chunk : 16;
  if (monitor) chunk = 1;
  const bool want_graph = P && P->use_graph && !monitor && chunk > 1 && !no_alloc;

  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  auto cleanup_graph = [&]() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  };

  const uint32_t end_mask = bicg ? (kFlagDone | kFlagFinishing) : kFlagDone;
  uint32_t launched = first_iter;
  int pending = -1;
  int slot = 0;
  bool finished = false;
  int status = VCLHIP_SUCCESS;
  while (!finished && launched < max_iters) {
    const uint32_t cnt = std::min(chunk, max_iters - launched);
    bool hook_inside = false;
    for (uint32_t k = 0; k < cnt; ++k) hook_inside = hook_inside || needs_hook(launched + k);
    if (want_graph && cnt == chunk && !hook_inside) {
      if (!gexec) {
        status = hip_status(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (status) break;
        int st2 = VCLHIP_SUCCESS;
        for (uint32_t k = 0; k < cnt && !st2; ++k) st2 = launch_iter(launched + k);
        if (!st2) st2 = tail();
        status = hip_status(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
        if (st2) status = st2;
        if (status) break;
        status = hip_status(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        if (status) break;
      }
      status = hip_status(hipGraphLaunch(gexec, s), "hipGraphLaunch");
      if (status) break;
    } else {
      for (uint32_t k = 0; k < cnt && !status; ++k) {
        if (needs_hook(launched + k)) status = hook(launched + k);
        if (!status) status = launch_iter(launched + k);
      }
      if (!status) status = tail();
      if (status) break;
    }
    launched += cnt;
    KState polled;  // no_alloc: the state after this chunk
    if (no_alloc) {
      status = hip_status(hipMemcpyAsync(&polled, st_dev, sizeof(KState), hipMemcpyDeviceToHost, s), "state D2H");
      if (!status) status = hip_status(hipStreamSynchronize(s), "solver sync");
    } else {
      status = hip_status(hipMemcpyAsync(&pin.h[slot], st_dev, sizeof(KState), hipMemcpyDeviceToHost, s), "state D2H");
      if (!status) status = hip_status(hipEventRecord(ev.e[slot], s), "hipEventRecord");
    }
    if (status) break;
    if (monitor) {
      if (!no_alloc) status = hip_status(hipEventSynchronize(ev.e[slot]), "hipEventSynchronize");
      if (status) break;
      const KState& h = no_alloc ? polled : pin.h[slot];
      if (h.iter > 0 && !h.breakdown) {
        const double rel = bicg ? std::fabs(h.res / h.nrs) : std::sqrt(std::fabs(h.rr / h.nrs));
        if (P->monitor(x_dev, rel, P->monitor_user)) {
          uint32_t done = kFlagDone;
          status = hip_status(hipMemcpyAsync(&st_dev->flags, &done, sizeof(uint32_t), hipMemcpyHostToDevice, s),
                              "stop flag");
          finished = true;
        }
      }
      if (h.flags & end_mask) finished = true;
    } else if (no_alloc) {
      if (polled.flags & end_mask) finished = true;
    } else if (pending >= 0) {
      status = hip_status(hipEventSynchronize(ev.e[pending]), "hipEventSynchronize");
      if (status) break;
      if (pin.h[pending].flags & end_mask) finished = true;
    }
    pending = slot;
    slot ^= 1;
    if (!finished && launched >= max_iters) {
      // Iterations that returned at once while an FP64 correction was pending did not count: when
      // the device has done fewer than max_iters, enqueue the rest.
      status = hip_status(hipStreamSynchronize(s), "solver sync");
      if (status) break;
      KState now;
      status = hip_status(hipMemcpy(&now, st_dev, sizeof(KState), hipMemcpyDeviceToHost), "state D2H");
      if (status) break;
      if ((now.flags & end_mask) || now.iter >= max_iters) break;
      launched = now.iter;
      pending = -1;
    }
  }
  cleanup_graph();
  int st_sync = hip_status(hipStreamSynchronize(s), "solver sync");
  if (!status) status = st_sync;
  if (!status)
    status = hip_status(hipMemcpy(&final_state, st_dev, sizeof(KState), hipMemcpyDeviceToHost), "final state");
  return status;
}

static void init_state(KState& h, const vclhip_solver_params* P, double* history_dev)
{
  std::memset(&h, 0, sizeof(h));
  h.tol = P->tol;
  h.abs_tol = P->abs_tol;
  h.max_iters = P->max_iters;
  h.history = history_dev;
  h.history_len = history_dev ? std::min(P->history_len, P->max_iters) : 0;
}

static int upload_state(KState* st_dev, const KState& h, hipStream_t s)
{
  VCLHIP_HIP(hipMemcpyAsync(st_dev, &h, sizeof(KState), hipMemcpyHostToDevice, s));
  VCLHIP_HIP(hipStreamSynchronize(s));  // h lives on the host stack
  return VCLHIP_SUCCESS;
}

static int copy_history(const vclhip_solver_params* P, const double* history_dev, uint32_t iters, hipStream_t s)
{
  if (!P->history_host || !history_dev) return VCLHIP_SUCCESS;
  const uint32_t cnt = std::min(std::min(P->history_len, P->max_iters), iters);
  if (cnt == 0) return VCLHIP_SUCCESS;
  VCLHIP_HIP(hipMemcpyAsync(P->history_host, history_dev, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  return VCLHIP_SUCCESS;
}

// ------------------------------------------------------------------ CG driver

// With an arena (arena_bytes of device memory) the solve allocates nothing: its vectors, partial
// sums and state are pieces of the arena, an optional DIA copy takes what is left or is not made,
// and the host waits for each chunk's state instead of using pinned memory, events or graphs.
template <typename T, class M>
static int cg_solve(const M* Ain, const T* b, T* x, int precond, const vclhip_solver_params* P,
                    vclhip_solver_result* R, void* stream, void* arena = nullptr, size_t arena_bytes = 0)
{
  if (!P || !R || !b || !x) return invalid("null argument");
  if (precond != VCLHIP_PRECOND_NONE && precond != VCLHIP_PRECOND_JACOBI) return invalid("precond");
  DevOp<T> op;
  VCLHIP_TRY(make_op(Ain, op));
  const uint32_t n = op.n;
  const bool jac = precond == VCLHIP_PRECOND_JACOBI;
  const uint32_t hist_len = (P->history_host && P->history_len && P->max_iters) ? std::min(P->history_len, P->max_iters) : 0;
  if (arena) {
    const size_t need = 3 * Workspace::arena_bytes(sizeof(T), n) + Workspace::arena_bytes(sizeof(double), kSlots * kMaxPartials) +
                        Workspace::arena_bytes(sizeof(KState), 1) + (hist_len ? Workspace::arena_bytes(sizeof(double), hist_len) : 0) +
                        (jac ? Workspace::arena_bytes(sizeof(T), n) + Workspace::arena_bytes(sizeof(uint32_t), 1) : 0);
    if (need > arena_bytes) {
      set_error("CG without allocations: n = %u needs %zu bytes of scratch memory, more than its %zu", n, need,
                arena_bytes);
      return VCLHIP_ERROR_TOO_LARGE;
    }
  }
  SolveStream ss(stream, P->use_graph != 0 && !arena);
  VCLHIP_TRY(ss.status());
  hipStream_t s = ss.get();
  Workspace ws = arena ? Workspace(s, arena, arena_bytes) : Workspace(s);
  T *r, *p, *Ap, *d = nullptr;
  double *part, *hist = nullptr;
  KState* st;
  VCLHIP_TRY(ws.alloc(&r, n));
  VCLHIP_TRY(ws.alloc(&p, n));
  VCLHIP_TRY(ws.alloc(&Ap, n));
  VCLHIP_TRY(ws.alloc(&part, kSlots * kMaxPartials));
  VCLHIP_TRY(ws.alloc(&st, 1));
  if (hist_len) VCLHIP_TRY(ws.alloc(&hist, hist_len));
  if (jac) {
    uint32_t* zc;
    VCLHIP_TRY(ws.alloc(&d, n));
    VCLHIP_TRY(ws.alloc(&zc, 1));
    VCLHIP_TRY(jacobi_diag(op, d, zc, s));
  }
  int used_dia = 0;
  VCLHIP_TRY(dia_operator(op, ws, s, P, &used_dia));
  KState h;
  init_state(h, P, hist);
  VCLHIP_TRY(upload_state(st, h, s));

  const uint32_t gv = solver_vec_grid(n);
  const bool v2 = aligned16({b, x});  // workspace vectors are 256-byte aligned
  // setup: x = 0, r = b, p = M^-1 b, Ap = A p, scalars
  if (jac)
    VCLHIP_LAUNCH_W(v2, (cg_init_kernel<T, true, kVecW<T> >), (cg_init_kernel<T, true, 1>), gv, s, n, b, x, r, p,
                    d, part);
  else
    VCLHIP_LAUNCH_W(v2, (cg_init_kernel<T, false, kVecW<T> >), (cg_init_kernel<T, false, 1>), gv, s, n, b, x, r, p,
                    d, part);
  VCLHIP_TRY(check_launch("cg_init_kernel"));
  if (jac)
    op.launch(s, p, EpiCg<T, true>{Ap, p, d}, slot_ptr(part, 1), nullptr, 0u);
  else
    op.launch(s, p, EpiCg<T, false>{Ap, p, d}, slot_ptr(part, 1), nullptr, 0u);
  VCLHIP_TRY(check_launch("cg prod"));
  if (jac)
    hipLaunchKernelGGL(cg_setup_kernel<true>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op.grid);
  else
    hipLaunchKernelGGL(cg_setup_kernel<false>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op.grid);
  VCLHIP_TRY(check_launch("cg_setup_kernel"));

  uint32_t* flags = &st->flags;
  auto launch_update = [&]() {
    if (jac)
      VCLHIP_LAUNCH_W(v2, (cg_update_kernel<T, true, kVecW<T> >), (cg_update_kernel<T, true, 1>), gv, s, n, x, p, r,
                      Ap, d, st, part);
    else
      VCLHIP_LAUNCH_W(v2, (cg_update_kernel<T, false, kVecW<T> >), (cg_update_kernel<T, false, 1>), gv, s, n, x, p,
                      r, Ap, d, st, part);
  };
  auto launch_prod = [&]() {
    if (jac)
      op.launch(s, p, EpiCg<T, true>{Ap, p, d}, slot_ptr(part, 1), flags, kFlagDone);
    else
      op.launch(s, p, EpiCg<T, false>{Ap, p, d}, slot_ptr(part, 1), flags, kFlagDone);
  };
  auto launch_iter = [&](uint32_t) -> int {
    launch_update();
    launch_prod();
    hipLaunchKernelGGL(cg_iter_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv, op.grid);
    return check_launch("cg iteration");
  };
  // Compact mode: two kernels per iteration (SpMV, then the scalar step fused with the next update)
  // where every update workgroup can afford to reduce all partials itself.  Not with a monitor
  // (the update of iteration k + 1 would run before the monitor sees iteration k) nor with graph
  // chunks of odd length (the captured parities must alternate across replays).
  const uint32_t chunk = P->chunk ? P->chunk : 16;
  const bool compact = !P->monitor && static_cast<uint64_t>(gv) * (gv + 2ull * op.grid) <= kCompactPartialReads &&
                       !(P->use_graph && (chunk & 1u));
  auto launch_iter_compact = [&](uint32_t k) -> int {
    launch_prod();
    if (jac)
      VCLHIP_LAUNCH_W(v2, (cg_step_kernel<T, true, kVecW<T> >), (cg_step_kernel<T, true, 1>), gv, s, n, x, p, r, Ap,
                      d, st, part, k & 1u, gv, op.grid);
    else
      VCLHIP_LAUNCH_W(v2, (cg_step_kernel<T, false, kVecW<T> >), (cg_step_kernel<T, false, 1>), gv, s, n, x, p, r,
                      Ap, d, st, part, k & 1u, gv, op.grid);
    return check_launch("cg iteration");
  };
  // Fused mode: one kernel per iteration (cg_fused_kernel) where compact mode runs on a DIA operator,
  // all workgroups together read at most kFusedPartialReads partial sums per iteration, and the grid
  // is resident at once (the kernel's registers leave room for three workgroups per CU; a second
  // round costs more than the launch it saves), unless params->variant is 1.  Its second copies of
  // r, p and A p are optional memory: without them the solve stays in compact mode (the same results
  // either way).
  T *r2 = nullptr, *p2 = nullptr, *Ap2 = nullptr;
  const uint32_t g_fused = std::max(gv, op.grid);
  bool fused = compact && op.fmt == OpFormat::Dia && !op.dia.code && P->variant != 1 &&
               static_cast<uint64_t>(g_fused) * (gv + 2ull * op.grid) <= kFusedPartialReads &&
               g_fused <= 3u * static_cast<uint32_t>(device_cu_count());
  if (fused) fused = ws.try_alloc(&r2, n) && ws.try_alloc(&p2, n) && ws.try_alloc(&Ap2, n);
  auto launch_iter_fused = [&](uint32_t k) -> int {
    const uint32_t par = k & 1u;
    T* const r_in = par ? r2 : r;
    T* const r_out = par ? r : r2;
    T* const p_in = par ? p2 : p;
    T* const p_out = par ? p : p2;
    T* const Ap_in = par ? Ap2 : Ap;
    T* const Ap_out = par ? Ap : Ap2;
    if (jac)
      VCLHIP_LAUNCH_W(v2, (cg_fused_kernel<T, true, kVecW<T> >), (cg_fused_kernel<T, true, 1>), g_fused, s, n, x,
                      r_in, r_out, p_in, p_out, Ap_in, Ap_out, d, op.dia, st, part, par, gv, op.grid);
    else
      VCLHIP_LAUNCH_W(v2, (cg_fused_kernel<T, false, kVecW<T> >), (cg_fused_kernel<T, false, 1>), g_fused, s, n, x,
                      r_in, r_out, p_in, p_out, Ap_in, Ap_out, d, op.dia, st, part, par, gv, op.grid);
    return check_launch("cg iteration");
  };
  KState fin;
  const bool no_alloc = arena != nullptr;
  if (compact) {
    launch_update();  // the update of iteration 0; iteration k then closes with that of k + 1
    VCLHIP_TRY(check_launch("cg update"));
    if (fused) {
      launch_prod();  // the product of iteration 0; iteration k's kernel then forms that of k + 1
      VCLHIP_TRY(check_launch("cg prod"));
      VCLHIP_TRY(run_iterations(
          s, st, 0u, P->max_iters, P, x, launch_iter_fused, [](uint32_t) { return false; },
          [](uint32_t) { return 0; }, fin, false, NoTail(), no_alloc));
    } else {
      VCLHIP_TRY(run_iterations(
          s, st, 0u, P->max_iters, P, x, launch_iter_compact, [](uint32_t) { return false; },
          [](uint32_t) { return 0; }, fin, false, NoTail(), no_alloc));
    }
  } else {
    VCLHIP_TRY(run_iterations(
        s, st, 0u, P->max_iters, P, x, launch_iter, [](uint32_t) { return false; }, [](uint32_t) { return 0; }, fin,
        false, NoTail(), no_alloc));
  }
  R->iters = fin.iter;
  R->error = fin.nrs != 0 ? std::sqrt(std::fabs(fin.rr) / fin.nrs) : 0.0;
  if (jac && fin.nrs != 0) R->error = std::sqrt(std::fabs(fin.rr / fin.nrs));
  R->restarts = 0;
  R->converged = static_cast<int>(fin.converged);
  R->operator_used = used_dia;
  R->kernels_per_iter = compact ? (fused ? 1u : 2u) : 3u;
  return copy_history(P, hist, fin.iter, s);
}

// ------------------------------------------------------------------ BiCGStab driver

template <typename T, class M>
static int bicgstab_solve(const M* Ain, const T* b, T* x, int precond, const vclhip_solver_params* P,
                          vclhip_solver_result* R, void* stream)
{
  if (!P || !R || !b || !x) return invalid("null argument");
  if (precond != VCLHIP_PRECOND_NONE && precond != VCLHIP_PRECOND_JACOBI) return invalid("precond");
  DevOp<T> op;
  VCLHIP_TRY(make_op(Ain, op));
  SolveStream ss(stream, P->use_graph != 0);
  VCLHIP_TRY(ss.status());
  hipStream_t s = ss.get();
  const uint32_t n = op.n;
  const bool jac = precond == VCLHIP_PRECOND_JACOBI;
  Workspace ws(s);
  int used_dia = 0;
  VCLHIP_TRY(dia_operator(op, ws, s, P, &used_dia));
  T *r, *p, *r0, *Ap, *sv, *As, *d = nullptr;
  double *part, *hist = nullptr;
  KState* st;
  VCLHIP_TRY(ws.alloc(&r, n));
  VCLHIP_TRY(ws.alloc(&p, n));
  VCLHIP_TRY(ws.alloc(&r0, n));
  VCLHIP_TRY(ws.alloc(&Ap, n));
  VCLHIP_TRY(ws.alloc(&sv, n));
  VCLHIP_TRY(ws.alloc(&As, n));
  VCLHIP_TRY(ws.alloc(&part, kSlots * kMaxPartials));
  VCLHIP_TRY(ws.alloc(&st, 1));
  if (P->history_host && P->history_len && P->max_iters) VCLHIP_TRY(ws.alloc(&hist, std::min(P->history_len, P->max_iters)));
  if (jac) {
    uint32_t* zc;
    VCLHIP_TRY(ws.alloc(&d, n));
    VCLHIP_TRY(ws.alloc(&zc, 1));
    VCLHIP_TRY(jacobi_diag(op, d, zc, s));
  }
  KState h;
  init_state(h, P, hist);
  VCLHIP_TRY(upload_state(st, h, s));

  const uint32_t gv = solver_vec_grid(n);
  const bool v2 = aligned16({b, x});
  if (jac)
    VCLHIP_LAUNCH_W(v2, (bicg_init_kernel<T, true, kVecW<T> >), (bicg_init_kernel<T, true, 1>), gv, s, n, b, x, r,
                    p, r0, d, part);
  else
    VCLHIP_LAUNCH_W(v2, (bicg_init_kernel<T, false, kVecW<T> >), (bicg_init_kernel<T, false, 1>), gv, s, n, b, x, r,
                    p, r0, d, part);
  VCLHIP_TRY(check_launch("bicg_init_kernel"));
  hipLaunchKernelGGL(bicg_setup_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv);
  VCLHIP_TRY(check_launch("bicg_setup_kernel"));

  uint32_t* flags = &st->flags;
  const uint32_t stop_all = kFlagDone | kFlagFinishing;
  const int brk = jac ? 1 : 0;
  const bool classic = jac && P->variant == 0;
  auto launch_classic = [&]() -> int {
    // t0 = M^-1 A p -> Ap; alpha; s = r - alpha t0; t1 = M^-1 A s -> As; omega; x, r; test; p
    op.launch(s, p, EpiBicg<T, true, false>{Ap, p, r0, d}, slot_ptr(part, 2), flags, kFlagDone);
    hipLaunchKernelGGL(bicgc_alpha_kernel, dim3(1), dim3(kBlock), 0, s, st, part, op.grid);
    VCLHIP_LAUNCH_W(v2, (bicg_update_s_kernel<T, kVecW<T> >), (bicg_update_s_kernel<T, 1>), gv, s, n, sv, r, Ap, st,
                    part);
    op.launch(s, sv, EpiBicg<T, true, true>{As, sv, r0, d}, slot_ptr(part, 4), flags, kFlagDone);
    hipLaunchKernelGGL(bicgc_omega_kernel, dim3(1), dim3(kBlock), 0, s, st, part, op.grid);
    VCLHIP_LAUNCH_W(v2, (bicgc_xr_kernel<T, kVecW<T> >), (bicgc_xr_kernel<T, 1>), gv, s, n, x, p, sv, r, As, r0, st,
                    part);
    hipLaunchKernelGGL(bicgc_check_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv);
    VCLHIP_LAUNCH_W(v2, (bicgc_p_kernel<T, kVecW<T> >), (bicgc_p_kernel<T, 1>), gv, s, n, p, r, Ap, st);
    return check_launch("bicgstab iteration");
  };
  // Compact mode (unpreconditioned, params->variant 0): four kernels per iteration where every
  // vector-update workgroup can afford to reduce all partials itself (as CG's compact mode); not with
  // hipGraph chunks of odd length (the captured parities must alternate across replays).
  const uint32_t chunk = P->chunk ? P->chunk : 16;
  const bool compact = !jac && P->variant == 0 &&
                       static_cast<uint64_t>(gv) * (gv + 3ull * op.grid) <= kCompactPartialReads &&
                       !(P->use_graph && (chunk & 1u) && !P->monitor);
  auto launch_iter_compact = [&](uint32_t k) -> int {
    op.launch(s, p, EpiBicg<T, false, false>{Ap, p, r0, d}, slot_ptr(part, 2), flags, stop_all);
    VCLHIP_LAUNCH_W(v2, (bicg_update_s_compact_kernel<T, kVecW<T> >), (bicg_update_s_compact_kernel<T, 1>), gv, s,
                    n, sv, r, Ap, st, part, gv, op.grid);
    op.launch(s, sv, EpiBicg<T, false, true>{As, sv, r0, d}, slot_ptr(part, 4), flags, stop_all);
    VCLHIP_LAUNCH_W(v2, (bicg_vec_update_compact_kernel<T, kVecW<T> >), (bicg_vec_update_compact_kernel<T, 1>), gv,
                    s, n, x, p, sv, r, As, Ap, r0, st, part, k & 1u, gv, op.grid);
    return check_launch("bicgstab iteration");
  };
  auto launch_iter = [&](uint32_t k) -> int {
    if (compact) return launch_iter_compact(k);
    if (classic) return launch_classic();
    if (jac)
      op.launch(s, p, EpiBicg<T, true, false>{Ap, p, r0, d}, slot_ptr(part, 2), flags, stop_all);
    else
      op.launch(s, p, EpiBicg<T, false, false>{Ap, p, r0, d}, slot_ptr(part, 2), flags, stop_all);
    hipLaunchKernelGGL(bicg_alpha_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv, op.grid, brk);
    VCLHIP_LAUNCH_W(v2, (bicg_update_s_kernel<T, kVecW<T> >), (bicg_update_s_kernel<T, 1>), gv, s, n, sv, r, Ap, st,
                    part);
    if (jac)
      op.launch(s, sv, EpiBicg<T, true, true>{As, sv, r0, d}, slot_ptr(part, 4), flags, stop_all);
    else
      op.launch(s, sv, EpiBicg<T, false, true>{As, sv, r0, d}, slot_ptr(part, 4), flags, stop_all);
    hipLaunchKernelGGL(bicg_scalars_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv, op.grid, brk);
    VCLHIP_LAUNCH_W(v2, (bicg_vec_update_kernel<T, kVecW<T> >), (bicg_vec_update_kernel<T, 1>), gv, s, n, x, p, sv,
                    r, As, Ap, r0, st, part);
    return check_launch("bicgstab iteration");
  };
  // Restarts of the preconditioned iteration (bicgstab.hpp:431-445, :473): iteration i restarts when
  // i - last_restart - 1 > max_iters_before_restart.
  const uint64_t mr = P->max_iters_before_restart;
  uint64_t last_restart = 0;
  auto needs_restart = [&](uint32_t i) { return jac && static_cast<uint64_t>(i) > last_restart + mr + 1; };
  auto restart = [&](uint32_t i) -> int {
    last_restart = i;
    if (jac)
      op.launch(s, x, EpiRestart<T, true>{b, d, r, p, r0}, slot_ptr(part, 7), flags, stop_all);
    else
      op.launch(s, x, EpiRestart<T, false>{b, d, r, p, r0}, slot_ptr(part, 7), flags, stop_all);
    hipLaunchKernelGGL(bicg_restart_kernel, dim3(1), dim3(kBlock), 0, s, st, part, op.grid);
    return check_launch("bicgstab restart");
  };

  KState fin;
  uint32_t restarts = 0;
  uint32_t first = 0;
  for (;;) {
    VCLHIP_TRY(run_iterations(s, st, first, P->max_iters, P, x, launch_iter, needs_restart, restart, fin, true));
    if (!(jac && fin.breakdown) || fin.iter >= P->max_iters) break;
    // breakdown: restart from the current x at iteration fin.iter and continue
    fin.breakdown = 0;
    fin.flags = 0;
    fin.final_mode = 0;
    VCLHIP_TRY(upload_state(st, fin, s));
    VCLHIP_TRY(restart(fin.iter));
    ++restarts;
    first = fin.iter;
  }
  R->iters = fin.iter;
  R->error = fin.nrs != 0 ? fin.res / fin.nrs : 0.0;
  R->restarts = restarts;
  R->converged = static_cast<int>(fin.converged);
  R->operator_used = used_dia;
  R->kernels_per_iter = compact ? 4u : (classic ? 8u : 6u);
  return copy_history(P, hist, fin.iter, s);
}

// ------------------------------------------------------------------ mixed-precision CG

// FP64 defect: r64 = b - A x64 and <r64,r64> (mixed_precision_cg.hpp:164-168).
struct EpiResid {
  using acc_t = double;
  static constexpr int K = 1;
  const double* __restrict__ b;
  double* __restrict__ r;
  __device__ __forceinline__ void operator()(uint32_t row, double s, acc_t* acc) const
  {
    const double v = b[row] - s;
    r[row] = v;
    acc[0] += v * v;
  }
};

// Partial-sum slots of the mixed solver (each kernel's inner products are adjacent slots):
// 0 <r,z>, 1 <r,r>, 2 <Ap, D^-1 Ap>, 3 <p,Ap>; 4-6 <r,z>, <r,r>, <b,b> of the start; 7 <r64,r64>.
constexpr uint32_t kSlotRZ = 0, kSlotRR = 1, kSlotApAp = 2, kSlotPAp = 3, kSlotInit = 4, kSlotR64 = 7;

// x64 = 0; x32 = 0; r32 = (float) b; p32 = D^-1 r32; partials <r,z>, <r,r>, <b,b>.
template <bool Jac, int W>
__global__ __launch_bounds__(kBlock) void mixed_init_kernel(uint32_t n, const double* __restrict__ b,
                                                            double* __restrict__ x, float* __restrict__ x32,
                                                            float* __restrict__ r32, float* __restrict__ p32,
                                                            const float* __restrict__ d, double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[3] = {0, 0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, double> vb = ldp<V>(b, i);
    Pack<V, double> zero64;
    Pack<V, float> r, z, zero32;
    Pack<V, float> vd;
    if constexpr (Jac) vd = ldp<V>(d, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      zero64.v[k] = 0;
      zero32.v[k] = 0;
      r.v[k] = static_cast<float>(vb.v[k]);
      z.v[k] = Jac ? r.v[k] / vd.v[k] : r.v[k];
      acc[0] += static_cast<double>(r.v[k] * z.v[k]);
      acc[1] += static_cast<double>(r.v[k] * r.v[k]);
      acc[2] += vb.v[k] * vb.v[k];
    }
    stp<V>(x, i, zero64);
    stp<V>(x32, i, zero32);
    stp<V>(r32, i, r);
    stp<V>(p32, i, z);
  });
  block_store_partials<3, double>(acc, slot_ptr(part, kSlotInit), kMaxPartials, lw);
}

// Scalars after the start (mode 1 also has Ap = A p of the first direction).
template <bool Jac>
__global__ __launch_bounds__(kBlock) void mixed_setup_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_vec, uint32_t g_prod, int mode)
{
  double sums[3];
  block_sum_partials_k<3, double>({slot_ptr(part, kSlotInit), slot_ptr(part, kSlotInit + 1),
                                   slot_ptr(part, kSlotInit + 2)},
                                  {g_vec, Jac ? g_vec : 0u, g_vec}, sums);
  const double rz = sums[0];
  const double rr = Jac ? sums[1] : rz;
  const double bb = sums[2];
  double ApAp = 0, pAp = 0;
  if (mode == 1) {
    ApAp = block_sum_partials(slot_ptr(part, kSlotApAp), g_prod);
    pAp = block_sum_partials(slot_ptr(part, kSlotPAp), g_prod);
  }
  if (threadIdx.x != 0) return;
  st->nrs = bb;
  st->iter = 0;
  st->updates = 0;
  st->res = 0;
  st->flags = kFlagNoPause | kFlagTailIdle;
  if (bb <= 0) {  // x = 0 solves b = 0 (mixed_precision_cg.hpp:112-113)
    st->flags |= kFlagDone;
    st->converged = 1;
    return;
  }
  if (mode == 0) {
    st->inner = static_cast<float>(bb);
    st->initial = static_cast<float>(bb);
  } else {
    st->rr = rz;
    st->rr2 = rr;
    st->rr_ref = rr;
    const double alpha = rz / pAp;
    st->alpha = alpha;
    st->beta = (alpha * alpha * ApAp - rz) / rz;
  }
  if (st->max_iters == 0) st->flags |= kFlagDone;
}

// ---- mode 0: the reference's schedule (FP32 classic CG, restart on inner_tol)

// alpha = inner / <t,p>  in FP32 (mixed_precision_cg.hpp:148).
__global__ __launch_bounds__(kBlock) void mixed0_alpha_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                              uint32_t g_prod)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  const double pAp = block_sum_partials(slot_ptr(part, kSlotPAp), g_prod);
  if (threadIdx.x != 0) return;
  st->alpha = static_cast<double>(st->inner / static_cast<float>(pAp));
  st->flags |= kFlagTailIdle;  // a restart that preceded this iteration is complete
}

// x32 += alpha p; r32 -= alpha t; partial <r,r>  (mixed_precision_cg.hpp:149-152)
template <int W>
__global__ __launch_bounds__(kBlock) void mixed0_xr_kernel(uint32_t n, float* __restrict__ x, float* __restrict__ r,
                                                           const float* __restrict__ p, const float* __restrict__ t,
                                                           const KState* __restrict__ st, double* __restrict__ part)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  const float alpha = static_cast<float>(st->alpha);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, float> vx = ldp<V>(x, i);
    Pack<V, float> vr = ldp<V>(r, i);
    const Pack<V, float> vp = ldp<V>(p, i);
    const Pack<V, float> vt = ldp<V>(t, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] = vx.v[k] + vp.v[k] * alpha;
      vr.v[k] = vr.v[k] + vt.v[k] * (-alpha);
      acc[0] += static_cast<double>(vr.v[k] * vr.v[k]);
    }
    stp<V>(x, i, vx);
    stp<V>(r, i, vr);
  });
  block_store_partials<1, double>(acc, slot_ptr(part, kSlotRZ), kMaxPartials, lw);
}

// beta = new/inner in FP32; restart test (mixed_precision_cg.hpp:154-161)
__global__ __launch_bounds__(kBlock) void mixed0_beta_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_vec)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  const double s = block_sum_partials(slot_ptr(part, kSlotRZ), g_vec);
  if (threadIdx.x != 0) return;
  const uint32_t iter = st->iter + 1;
  st->iter = iter;
  const float nw = static_cast<float>(s);
  record_history(st, iter, static_cast<double>(nw));
  st->beta = static_cast<double>(nw / st->inner);
  st->inner = nw;
  if (nw < st->inner_tol * st->initial || iter >= st->max_iters)
    st->flags = (st->flags | kFlagPause) & ~kFlagNoPause;
}

// p32 = r32 + beta p32
template <int W>
__global__ __launch_bounds__(kBlock) void mixed0_p_kernel(uint32_t n, float* __restrict__ p, const float* __restrict__ r,
                                                          const KState* __restrict__ st)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  const float beta = static_cast<float>(st->beta);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, float> vp = ldp<V>(p, i);
    const Pack<V, float> vr = ldp<V>(r, i);
#pragma unroll
    for (int k = 0; k < V; ++k) vp.v[k] = vr.v[k] + vp.v[k] * beta;
    stp<V>(p, i, vp);
  });
}

// ---- mode 1: FP32 pipelined CG with reliable updates

// x32 += alpha p; r32 -= alpha Ap; z = D^-1 r; p = z + beta p; partials <r,z> (and <r,r> with Jacobi)
template <bool Jac, int W>
__global__ __launch_bounds__(kBlock) void mixed1_update_kernel(uint32_t n, float* __restrict__ x,
                                                               float* __restrict__ p, float* __restrict__ r,
                                                               const float* __restrict__ Ap,
                                                               const float* __restrict__ d,
                                                               const KState* __restrict__ st,
                                                               double* __restrict__ part)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  const float alpha = static_cast<float>(st->alpha);
  const float beta = static_cast<float>(st->beta);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[2] = {0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, float> vp = ldp<V>(p, i);
    Pack<V, float> vr = ldp<V>(r, i);
    Pack<V, float> vx = ldp<V>(x, i);
    const Pack<V, float> va = ldp<V>(Ap, i);
    Pack<V, float> vd;
    if constexpr (Jac) vd = ldp<V>(d, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] += alpha * vp.v[k];
      vr.v[k] -= alpha * va.v[k];
      float z = vr.v[k];
      if constexpr (Jac) z = vr.v[k] / vd.v[k];
      vp.v[k] = z + beta * vp.v[k];
      acc[0] += static_cast<double>(vr.v[k] * z);
      if constexpr (Jac) acc[1] += static_cast<double>(vr.v[k] * vr.v[k]);
    }
    stp<V>(x, i, vx);
    stp<V>(p, i, vp);
    stp<V>(r, i, vr);
  });
  if constexpr (Jac) {
    block_store_partials<2, double>(acc, slot_ptr(part, kSlotRZ), kMaxPartials, lw);
  } else {
    const double a0[1] = {acc[0]};
    block_store_partials<1, double>(a0, slot_ptr(part, kSlotRZ), kMaxPartials, lw);
  }
}

// Scalars of one iteration; asks for an FP64 correction when the FP32 <r,r> has fallen by
// inner_tol since the last one, when it says the tolerance is met, or at the iteration limit.
template <bool Jac>
__global__ __launch_bounds__(kBlock) void mixed1_iter_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_vec, uint32_t g_prod)
{
  if (st->flags & (kFlagDone | kFlagPause)) return;
  double sums[4];
  block_sum_partials_k<4, double>({slot_ptr(part, kSlotRZ), slot_ptr(part, kSlotRR), slot_ptr(part, kSlotApAp),
                                   slot_ptr(part, kSlotPAp)},
                                  {g_vec, Jac ? g_vec : 0u, g_prod, g_prod}, sums);
  const double rz = sums[0];
  const double rr = Jac ? sums[1] : rz;
  const double ApAp = sums[2], pAp = sums[3];
  if (threadIdx.x != 0) return;
  const uint32_t iter = st->iter + 1;
  st->iter = iter;
  st->rr = rz;
  st->rr2 = rr;
  record_history(st, iter, rr);
  if (rr < static_cast<double>(st->inner_tol) * st->rr_ref || rr / st->nrs < st->tol * st->tol ||
      iter >= st->max_iters) {
    st->flags = (st->flags | kFlagPause) & ~kFlagNoPause;
    return;
  }
  const double alpha = rz / pAp;
  st->alpha = alpha;
  st->beta = (alpha * alpha * ApAp - rz) / rz;
}

// ---- FP64 correction (both modes)

// x64 += x32; x32 = 0  (mixed_precision_cg.hpp:163-164)
template <int W>
__global__ __launch_bounds__(kBlock) void mixed_fold_kernel(uint32_t n, double* __restrict__ x,
                                                            float* __restrict__ x32, const KState* __restrict__ st)
{
  if (st->flags & (kFlagDone | kFlagNoPause)) return;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    Pack<V, double> vx = ldp<V>(x, i);
    const Pack<V, float> vs = ldp<V>(x32, i);
    Pack<V, float> zero;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      vx.v[k] = vx.v[k] + static_cast<double>(vs.v[k]);
      zero.v[k] = 0;
    }
    stp<V>(x, i, vx);
    stp<V>(x32, i, zero);
  });
}

// Tests ||b - A x||^2 / ||b||^2 < tol^2 in FP64 (mixed_precision_cg.hpp:167-169).
__global__ __launch_bounds__(kBlock) void mixed_check_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                             uint32_t g_prod, int mode)
{
  if (st->flags & (kFlagDone | kFlagNoPause)) return;
  const double rr64 = block_sum_partials(slot_ptr(part, kSlotR64), g_prod);
  if (threadIdx.x != 0) return;
  st->updates += 1;
  st->res = rr64;
  if (rr64 / st->nrs < st->tol * st->tol) {
    st->flags |= kFlagDone;
    st->converged = 1;
    return;
  }
  if (st->iter >= st->max_iters) {
    st->flags |= kFlagDone;
    return;
  }
  if (mode == 0) {
    st->initial = static_cast<float>(rr64);
    st->inner = static_cast<float>(rr64);
  }
  st->flags = (st->flags & ~(kFlagPause | kFlagTailIdle)) | kFlagNoPause;
}

// mode 0: p32 = r32 = (float) r64 (mixed_precision_cg.hpp:171-174)
template <int W>
__global__ __launch_bounds__(kBlock) void mixed0_restart_kernel(uint32_t n, const double* __restrict__ r64,
                                                                float* __restrict__ r32, float* __restrict__ p32,
                                                                const KState* __restrict__ st)
{
  if (st->flags & (kFlagDone | kFlagTailIdle)) return;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, double> v = ldp<V>(r64, i);
    Pack<V, float> f;
#pragma unroll
    for (int k = 0; k < V; ++k) f.v[k] = static_cast<float>(v.v[k]);
    stp<V>(r32, i, f);
    stp<V>(p32, i, f);
  });
}

// mode 1: r32 = (float) r64; p32 += D^-1 (r_new - r_old) so that p = z + beta p_prev still holds;
// partials <r,z>, <r,r>.
template <bool Jac, int W>
__global__ __launch_bounds__(kBlock) void mixed1_replace_kernel(uint32_t n, const double* __restrict__ r64,
                                                                float* __restrict__ r32, float* __restrict__ p32,
                                                                const float* __restrict__ d,
                                                                const KState* __restrict__ st,
                                                                double* __restrict__ part)
{
  if (st->flags & (kFlagDone | kFlagTailIdle)) return;
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range_w(n, W, lw, gridDim.x, c0, c1);
  double acc[2] = {0, 0};
  for_packs<W>(c0, c1, [&](uint32_t i, auto w) {
    constexpr int V = decltype(w)::value;
    const Pack<V, double> v = ldp<V>(r64, i);
    Pack<V, float> vr = ldp<V>(r32, i);
    Pack<V, float> vp = ldp<V>(p32, i);
    Pack<V, float> vd;
    if constexpr (Jac) vd = ldp<V>(d, i);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const float rn = static_cast<float>(v.v[k]);
      const float zo = Jac ? vr.v[k] / vd.v[k] : vr.v[k];
      const float zn = Jac ? rn / vd.v[k] : rn;
      vp.v[k] = vp.v[k] + (zn - zo);
      vr.v[k] = rn;
      acc[0] += static_cast<double>(rn * zn);
      acc[1] += static_cast<double>(rn * rn);
    }
    stp<V>(r32, i, vr);
    stp<V>(p32, i, vp);
  });
  block_store_partials<2, double>(acc, slot_ptr(part, kSlotRZ), kMaxPartials, lw);
}

// mode 1: scalars of the refreshed state; ends the correction.
__global__ __launch_bounds__(kBlock) void mixed1_resume_kernel(KState* __restrict__ st, const double* __restrict__ part,
                                                               uint32_t g_vec, uint32_t g_prod)
{
  if (st->flags & (kFlagDone | kFlagTailIdle)) return;
  double sums[4];
  block_sum_partials_k<4, double>({slot_ptr(part, kSlotRZ), slot_ptr(part, kSlotRR), slot_ptr(part, kSlotApAp),
                                   slot_ptr(part, kSlotPAp)},
                                  {g_vec, g_vec, g_prod, g_prod}, sums);
  const double rz = sums[0], rr = sums[1], ApAp = sums[2], pAp = sums[3];
  if (threadIdx.x != 0) return;
  st->rr = rz;
  st->rr2 = rr;
  st->rr_ref = rr;
  const double alpha = rz / pAp;
  st->alpha = alpha;
  st->beta = (alpha * alpha * ApAp - rz) / rz;
  st->flags |= kFlagTailIdle;
}

__global__ __launch_bounds__(kBlock) void convert_values_kernel(size_t n, const double* __restrict__ v,
                                                                float* __restrict__ f)
{
  for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * kBlock)
    f[i] = static_cast<float>(v[i]);
}

// ---- FP32 SELL-64 copy of a CSR matrix, built on the device (the inner operator of the mixed CG)

constexpr uint32_t kInnerC = kWave;  // slice height: one wavefront per slice

// width[s] = longest row of slice s (one wavefront per slice); *max_offset = max |col - row|.
__global__ __launch_bounds__(kBlock) void sell_width_kernel(const uint32_t* __restrict__ rp,
                                                            const uint32_t* __restrict__ col, uint32_t n,
                                                            uint32_t slices, uint32_t* __restrict__ width,
                                                            uint32_t* __restrict__ max_offset)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  uint32_t len = 0, dev = 0;
  if (row < n) {
    const uint32_t a = rp[row];
    len = rp[row + 1] - a;
    for (uint32_t k = 0; k < len; ++k) {
      const uint32_t c = col[a + k];
      dev = max(dev, c > row ? c - row : row - c);
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    len = max(len, static_cast<uint32_t>(__shfl_xor(static_cast<int>(len), o, kWave)));
    dev = max(dev, static_cast<uint32_t>(__shfl_xor(static_cast<int>(dev), o, kWave)));
  }
  const uint32_t s = row / kInnerC;
  if ((threadIdx.x & (kWave - 1)) == 0 && s < slices) {
    width[s] = len;
    atomicMax(max_offset, dev);
  }
}

// start[s] = C * (width[0] + ... + width[s-1]); total = C * sum of all widths.  One workgroup.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void sell_scan_kernel(const uint32_t* __restrict__ width, uint32_t slices,
                                                                 uint32_t* __restrict__ start,
                                                                 unsigned long long* __restrict__ total)
{
  __shared__ unsigned long long part[kScanThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (slices + kScanThreads - 1) / kScanThreads;
  const uint32_t lo = min(t * per, slices), hi = min(lo + per, slices);
  unsigned long long sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += static_cast<unsigned long long>(width[i]) * kInnerC;
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
    const unsigned long long v = t >= static_cast<uint32_t>(o) ? part[t - o] : 0ull;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned long long run = part[t] - sum;
  for (uint32_t i = lo; i < hi; ++i) {
    start[i] = static_cast<uint32_t>(run);
    run += static_cast<unsigned long long>(width[i]) * kInnerC;
  }
  if (t == kScanThreads - 1) *total = part[t];
}

// Fills the slices: entry k of row r at start[s] + k*C + lane, FP32 values; padding is value 0,
// column 0 (sliced_ell_matrix.hpp:180-200).
// ci or dci (16-bit column - row) receives the columns; the other pointer is null.
__global__ __launch_bounds__(kBlock) void sell_fill_kernel(const uint32_t* __restrict__ rp,
                                                           const uint32_t* __restrict__ col,
                                                           const double* __restrict__ val, uint32_t n,
                                                           uint32_t slices, const uint32_t* __restrict__ width,
                                                           const uint32_t* __restrict__ start,
                                                           uint32_t* __restrict__ ci, int16_t* __restrict__ dci,
                                                           float* __restrict__ el)
{
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t s = row / kInnerC;
  if (s >= slices) return;
  const uint32_t lane = row - s * kInnerC;
  const uint32_t w = width[s];
  const uint32_t a = row < n ? rp[row] : 0u;
  const uint32_t len = row < n ? rp[row + 1] - a : 0u;
  const size_t base = static_cast<size_t>(start[s]) + lane;
  for (uint32_t k = 0; k < w; ++k) {
    const size_t off = base + static_cast<size_t>(k) * kInnerC;
    const uint32_t c = k < len ? col[a + k] : (dci ? row : 0u);
    if (dci)
      dci[off] = static_cast<int16_t>(static_cast<int32_t>(c) - static_cast<int32_t>(row));
    else
      ci[off] = c;
    el[off] = k < len ? static_cast<float>(val[a + k]) : 0.0f;
  }
}

// Builds the SELL-64 FP32 copy of A into workspace memory when it stores at most `max_fill` times
// nnz entries (or always when forced).  The columns are stored as 16-bit offsets from the row when
// every |col - row| fits and `allow16` holds (*dcol receives them), else as 32-bit columns.
// *built tells whether a SELL copy was made.
static int build_inner_sell(const vclhip_csr* A, Workspace& ws, hipStream_t s, double max_fill, bool force,
                            bool allow16, vclhip_sell* out, const int16_t** dcol, size_t* stored_out, bool* built)
{
  *built = false;
  *dcol = nullptr;
  *stored_out = 0;
  const uint32_t n = A->n_rows;
  const uint32_t slices = (n + kInnerC - 1) / kInnerC;
  uint32_t *width, *start, *maxoff;
  unsigned long long* total;
  VCLHIP_TRY(ws.alloc(&width, slices));
  VCLHIP_TRY(ws.alloc(&start, slices));
  VCLHIP_TRY(ws.alloc(&total, 1));
  VCLHIP_TRY(ws.alloc(&maxoff, 1));
  VCLHIP_HIP(hipMemsetAsync(maxoff, 0, sizeof(uint32_t), s));
  const uint32_t g = (slices * kInnerC + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(sell_width_kernel, dim3(g), dim3(kBlock), 0, s, A->row_ptr, A->col_idx, n, slices, width,
                     maxoff);
  VCLHIP_TRY(check_launch("sell_width_kernel"));
  hipLaunchKernelGGL(sell_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, width, slices, start, total);
  VCLHIP_TRY(check_launch("sell_scan_kernel"));
  unsigned long long stored = 0;
  uint32_t max_offset = 0;
  VCLHIP_HIP(hipMemcpyAsync(&stored, total, sizeof(stored), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipMemcpyAsync(&max_offset, maxoff, sizeof(max_offset), hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  if (stored >= (1ull << 32)) return VCLHIP_SUCCESS;  // block_start is 32-bit
  if (!force && static_cast<double>(stored) > max_fill * static_cast<double>(A->nnz)) return VCLHIP_SUCCESS;
  const bool use16 = allow16 && max_offset <= 32767u;
  uint32_t* ci = nullptr;
  int16_t* dci = nullptr;
  float* el;
  if (use16)
    VCLHIP_TRY(ws.alloc(&dci, stored ? stored : 1));
  else
    VCLHIP_TRY(ws.alloc(&ci, stored ? stored : 1));
  VCLHIP_TRY(ws.alloc(&el, stored ? stored : 1));
  hipLaunchKernelGGL(sell_fill_kernel, dim3(g), dim3(kBlock), 0, s, A->row_ptr, A->col_idx,
                     static_cast<const double*>(A->values), n, slices, width, start, ci, dci, el);
  VCLHIP_TRY(check_launch("sell_fill_kernel"));
  out->n_rows = n;
  out->n_cols = A->n_cols;
  out->C = kInnerC;
  out->num_slices = slices;
  out->columns_per_block = width;
  out->column_indices = ci;
  out->block_start = start;
  out->elements = el;
  *dcol = dci;
  *stored_out = static_cast<size_t>(stored);
  *built = true;
  return VCLHIP_SUCCESS;
}

inline const double* values_of(const vclhip_csr* A) { return static_cast<const double*>(A->values); }
inline const double* values_of(const vclhip_sell* A) { return static_cast<const double*>(A->elements); }
// Number of stored values: nnz for CSR; block_start[last] + columns_per_block[last] * C for SELL
// (read back from the device).
inline int value_count(const vclhip_csr* A, size_t* count)
{
  *count = A->nnz;
  return VCLHIP_SUCCESS;
}
inline int value_count(const vclhip_sell* A, size_t* count)
{
  *count = 0;
  if (!A || !A->block_start || !A->columns_per_block || A->num_slices == 0) return VCLHIP_SUCCESS;
  uint32_t start = 0, width = 0;
  VCLHIP_HIP(hipMemcpy(&start, A->block_start + (A->num_slices - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
  VCLHIP_HIP(hipMemcpy(&width, A->columns_per_block + (A->num_slices - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
  *count = static_cast<size_t>(start) + static_cast<size_t>(width) * A->C;
  return VCLHIP_SUCCESS;
}
inline vclhip_csr with_values(const vclhip_csr* A, const float* v)
{
  vclhip_csr c = *A;
  c.values = v;
  return c;
}
inline vclhip_sell with_values(const vclhip_sell* A, const float* v)
{
  vclhip_sell c = *A;
  c.elements = v;
  return c;
}

// mixed_precision_cg.hpp:95-187 on the device.  values32: FP32 copy of the matrix values, or null
// (converted into a workspace here, as the reference converts at every solve).
template <class M>
static int mixed_cg_solve(const M* A, const float* values32, const double* b, double* x, int precond,
                          const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  if (!P || !R || !b || !x) return invalid("null argument");
  if (precond != VCLHIP_PRECOND_NONE && precond != VCLHIP_PRECOND_JACOBI) return invalid("precond");
  const int mode = P->mixed_mode;
  if (mode != 0 && mode != 1) return invalid("mixed_mode must be 0 or 1");
  if (mode == 0 && precond != VCLHIP_PRECOND_NONE)
    return invalid("mixed_mode 0 (the reference's restart schedule) has no preconditioner");
  DevOp<double> op64;
  VCLHIP_TRY(make_op(A, op64));
  SolveStream ss(stream, P->use_graph != 0);
  VCLHIP_TRY(ss.status());
  hipStream_t s = ss.get();
  const uint32_t n = op64.n;
  const bool jac = precond == VCLHIP_PRECOND_JACOBI;
  Workspace ws(s);
  // FP32 inner operator: the caller's FP32 values, or a copy converted here -- as SELL-64 when A is
  // CSR and the slices pad little (one wavefront per slice, no row pointers), else in A's format.
  DevOp<float> op32;
  vclhip_sell inner_sell{};
  const int16_t* inner_dcol = nullptr;
  size_t inner_stored = 0;
  bool inner_is_sell = false;
  DiaDev<float> inner_dia{};
  bool inner_is_dia = false;
  if (P->inner_format < VCLHIP_INNER_AUTO || P->inner_format > VCLHIP_INNER_DIA) return invalid("inner_format");
  if constexpr (std::is_same<M, vclhip_csr>::value) {
    if (!values32 && (P->inner_format == VCLHIP_INNER_AUTO || P->inner_format == VCLHIP_INNER_DIA)) {
      // the inner copy is summed like the SELL copy (zeros skipped), so stored zeros do not refuse it
      VCLHIP_TRY((build_dia<double, float>(A, ws, s, 1.25, dia_memory_cap(), false, WorkspaceDiaAlloc{ws},
                                           &inner_dia, &inner_is_dia)));
      if (P->inner_format == VCLHIP_INNER_DIA && !inner_is_dia) {
        set_error("inner_format DIA: more than %u diagonals, or their padding exceeds 25%% of nnz", kDiaMax);
        return VCLHIP_ERROR_NOT_SUPPORTED;
      }
    }
    if (!values32 && !inner_is_dia && P->inner_format != VCLHIP_INNER_CSR) {
      const bool forced = P->inner_format != VCLHIP_INNER_AUTO;
      const bool allow16 = P->inner_format == VCLHIP_INNER_AUTO || P->inner_format == VCLHIP_INNER_SELL64_D16;
      VCLHIP_TRY(build_inner_sell(A, ws, s, 1.25, forced, allow16, &inner_sell, &inner_dcol, &inner_stored,
                                  &inner_is_sell));
      if (P->inner_format == VCLHIP_INNER_SELL64_D16 && !inner_dcol) {
        set_error("inner_format SELL64_D16: a column is more than 32767 from its row");
        return VCLHIP_ERROR_NOT_SUPPORTED;
      }
    }
  }
  M A32{};
  if (inner_is_dia) {
    op32.fmt = OpFormat::Dia;
    op32.sell = false;
    op32.dia = inner_dia;
    op32.dia_nt = dia_stream_loads((static_cast<size_t>(inner_dia.nd) + 4) * n * sizeof(float));
    op32.n = n;
    op32.grid = chunk_grid((n + kBlock - 1) / kBlock, sell_gmax());
  } else if (inner_is_sell) {
    VCLHIP_TRY(make_op(&inner_sell, op32));
    op32.sm.dcol = inner_dcol;
    op32.nt = stream_loads(inner_stored * (sizeof(float) + (inner_dcol ? 2 : 4)));
  } else {
    float* v32 = const_cast<float*>(values32);
    if (!v32) {
      size_t value_count32 = 0;
      VCLHIP_TRY(value_count(A, &value_count32));
      VCLHIP_TRY(ws.alloc(&v32, value_count32 ? value_count32 : 1));
      if (value_count32) {
        const uint32_t g = static_cast<uint32_t>(std::min<size_t>((value_count32 + kBlock - 1) / kBlock, 8192));
        hipLaunchKernelGGL(convert_values_kernel, dim3(g), dim3(kBlock), 0, s, value_count32, values_of(A), v32);
        VCLHIP_TRY(check_launch("convert_values_kernel"));
      }
    }
    A32 = with_values(A, v32);
    VCLHIP_TRY(make_op(&A32, op32));
  }

  float *x32, *r32, *p32, *Ap32, *d32 = nullptr;
  double *r64, *part, *hist = nullptr;
  KState* st;
  VCLHIP_TRY(ws.alloc(&x32, n));
  VCLHIP_TRY(ws.alloc(&r32, n));
  VCLHIP_TRY(ws.alloc(&p32, n));
  VCLHIP_TRY(ws.alloc(&Ap32, n));
  VCLHIP_TRY(ws.alloc(&r64, n));
  VCLHIP_TRY(ws.alloc(&part, kSlots * kMaxPartials));
  VCLHIP_TRY(ws.alloc(&st, 1));
  if (P->history_host && P->history_len && P->max_iters) VCLHIP_TRY(ws.alloc(&hist, std::min(P->history_len, P->max_iters)));
  if (jac) {
    uint32_t* zc;
    VCLHIP_TRY(ws.alloc(&d32, n));
    VCLHIP_TRY(ws.alloc(&zc, 1));
    VCLHIP_TRY(jacobi_diag(op32, d32, zc, s));
  }
  KState h;
  init_state(h, P, hist);
  h.inner_tol = P->inner_tol;
  VCLHIP_TRY(upload_state(st, h, s));

  const uint32_t gv = solver_vec_grid(n);
  const bool v4 = aligned16({b, x});
  if (jac)
    VCLHIP_LAUNCH_W(v4, (mixed_init_kernel<true, 4>), (mixed_init_kernel<true, 1>), gv, s, n, b, x, x32, r32, p32, d32,
                    part);
  else
    VCLHIP_LAUNCH_W(v4, (mixed_init_kernel<false, 4>), (mixed_init_kernel<false, 1>), gv, s, n, b, x, x32, r32, p32,
                    d32, part);
  VCLHIP_TRY(check_launch("mixed_init_kernel"));
  if (mode == 1) {
    if (jac)
      op32.launch(s, p32, EpiCg<float, true>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), nullptr, 0u);
    else
      op32.launch(s, p32, EpiCg<float, false>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), nullptr, 0u);
    VCLHIP_TRY(check_launch("mixed prod"));
  }
  if (jac)
    hipLaunchKernelGGL(mixed_setup_kernel<true>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op32.grid, mode);
  else
    hipLaunchKernelGGL(mixed_setup_kernel<false>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op32.grid, mode);
  VCLHIP_TRY(check_launch("mixed_setup_kernel"));

  uint32_t* flags = &st->flags;
  const uint32_t run_mask = kFlagDone | kFlagPause;
  auto launch_iter = [&](uint32_t) -> int {
    if (mode == 0) {
      op32.launch(s, p32, EpiCg<float, false>{Ap32, p32, nullptr}, slot_ptr(part, kSlotApAp), flags, run_mask);
      hipLaunchKernelGGL(mixed0_alpha_kernel, dim3(1), dim3(kBlock), 0, s, st, part, op32.grid);
      hipLaunchKernelGGL(mixed0_xr_kernel<4>, dim3(gv), dim3(kBlock), 0, s, n, x32, r32, p32, Ap32, st, part);
      hipLaunchKernelGGL(mixed0_beta_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv);
      hipLaunchKernelGGL(mixed0_p_kernel<4>, dim3(gv), dim3(kBlock), 0, s, n, p32, r32, st);
      return check_launch("mixed CG iteration");
    }
    if (jac) {
      hipLaunchKernelGGL((mixed1_update_kernel<true, 4>), dim3(gv), dim3(kBlock), 0, s, n, x32, p32, r32, Ap32, d32, st,
                         part);
      op32.launch(s, p32, EpiCg<float, true>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), flags, run_mask);
      hipLaunchKernelGGL(mixed1_iter_kernel<true>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op32.grid);
    } else {
      hipLaunchKernelGGL((mixed1_update_kernel<false, 4>), dim3(gv), dim3(kBlock), 0, s, n, x32, p32, r32, Ap32, d32,
                         st, part);
      op32.launch(s, p32, EpiCg<float, false>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), flags, run_mask);
      hipLaunchKernelGGL(mixed1_iter_kernel<false>, dim3(1), dim3(kBlock), 0, s, st, part, gv, op32.grid);
    }
    return check_launch("mixed CG iteration");
  };
  auto tail = [&]() -> int {
    if (v4)
      hipLaunchKernelGGL(mixed_fold_kernel<4>, dim3(gv), dim3(kBlock), 0, s, n, x, x32, st);
    else
      hipLaunchKernelGGL(mixed_fold_kernel<1>, dim3(gv), dim3(kBlock), 0, s, n, x, x32, st);
    op64.launch(s, x, EpiResid{b, r64}, slot_ptr(part, kSlotR64), flags, kFlagDone | kFlagNoPause);
    hipLaunchKernelGGL(mixed_check_kernel, dim3(1), dim3(kBlock), 0, s, st, part, op64.grid, mode);
    if (mode == 0) {
      hipLaunchKernelGGL(mixed0_restart_kernel<4>, dim3(gv), dim3(kBlock), 0, s, n, r64, r32, p32, st);
      return check_launch("mixed CG restart");
    }
    if (jac) {
      hipLaunchKernelGGL((mixed1_replace_kernel<true, 4>), dim3(gv), dim3(kBlock), 0, s, n, r64, r32, p32, d32, st, part);
      op32.launch(s, p32, EpiCg<float, true>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), flags,
                  kFlagDone | kFlagTailIdle);
    } else {
      hipLaunchKernelGGL((mixed1_replace_kernel<false, 4>), dim3(gv), dim3(kBlock), 0, s, n, r64, r32, p32, d32, st,
                         part);
      op32.launch(s, p32, EpiCg<float, false>{Ap32, p32, d32}, slot_ptr(part, kSlotApAp), flags,
                  kFlagDone | kFlagTailIdle);
    }
    hipLaunchKernelGGL(mixed1_resume_kernel, dim3(1), dim3(kBlock), 0, s, st, part, gv, op32.grid);
    return check_launch("mixed CG correction");
  };
  KState fin;
  VCLHIP_TRY(run_iterations(
      s, st, 0u, P->max_iters, P, x, launch_iter, [](uint32_t) { return false; }, [](uint32_t) { return 0; }, fin,
      false, tail));
  R->iters = fin.iter;
  R->error = fin.nrs > 0 ? std::sqrt(fin.res / fin.nrs) : 0.0;
  R->restarts = fin.updates;
  R->converged = static_cast<int>(fin.converged);
  R->kernels_per_iter = 0;
  R->operator_used = inner_is_dia    ? (inner_dia.code ? VCLHIP_INNER_DIA_CODED : VCLHIP_INNER_DIA)
                     : inner_is_sell ? (inner_dcol ? VCLHIP_INNER_SELL64_D16 : VCLHIP_INNER_SELL64)
                                     : VCLHIP_INNER_CSR;
  return copy_history(P, hist, fin.iter, s);
}

// ------------------------------------------------------------------ pipelined ops with the reference's buffer

// Sums slot q over `count` partials into buf[dest[q]] and zeroes buf[dest[q]+1 .. dest[q]+chunk).
__global__ __launch_bounds__(kBlock) void write_chunks_kernel(const double* __restrict__ part, uint32_t count, int K,
                                                              uint32_t d0, uint32_t d1, uint32_t d2,
                                                              double* __restrict__ buf, uint32_t chunk)
{
  const uint32_t dest[3] = {d0, d1, d2};
  for (int q = 0; q < K; ++q) {
    const double v = block_sum_partials(slot_ptr(part, q), count);
    for (uint32_t i = threadIdx.x; i < chunk; i += kBlock) buf[dest[q] + i] = (i == 0) ? v : 0.0;
    __syncthreads();
  }
}

// alpha = sum(buf[0..chunk)) / sum(buf[3chunk..4chunk)), sequential as host_based :539-544
__global__ void bicg_buffer_alpha_kernel(const double* __restrict__ buf, uint32_t chunk, double* __restrict__ alpha)
{
  if (threadIdx.x != 0) return;
  double a = 0, c = 0;
  for (uint32_t i = 0; i < chunk; ++i) {
    a += buf[i];
    c += buf[i + 3 * chunk];
  }
  *alpha = a / c;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bicg_update_s_ptr_kernel(uint32_t n, T* __restrict__ s,
                                                                   const T* __restrict__ r, const T* __restrict__ Ap,
                                                                   const double* __restrict__ alpha_dev,
                                                                   double* __restrict__ part)
{
  const T alpha = static_cast<T>(*alpha_dev);
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    const T vs = r[i] - alpha * Ap[i];
    acc[0] += static_cast<double>(vs * vs);
    s[i] = vs;
  }
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void cg_update_scalar_kernel(uint32_t n, T* __restrict__ x, T alpha,
                                                                  T* __restrict__ p, T* __restrict__ r,
                                                                  const T* __restrict__ Ap, T beta,
                                                                  double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    T vp = p[i];
    T vr = r[i];
    x[i] += alpha * vp;
    vr -= alpha * Ap[i];
    vp = vr + beta * vp;
    acc[0] += static_cast<double>(vr * vr);
    p[i] = vp;
    r[i] = vr;
  }
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bicg_vec_update_scalar_kernel(uint32_t n, T* __restrict__ x, T alpha,
                                                                        T* __restrict__ p, T omega,
                                                                        const T* __restrict__ s, T* __restrict__ r,
                                                                        const T* __restrict__ As, T beta,
                                                                        const T* __restrict__ Ap,
                                                                        const T* __restrict__ r0,
                                                                        double* __restrict__ part)
{
  const uint32_t lw = xcd_logical_block();
  uint32_t c0, c1;
  chunk_range(n, lw, gridDim.x, c0, c1);
  double acc[1] = {0};
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
    T vx = x[i];
    T vp = p[i];
    const T vs = s[i];
    vx += alpha * vp + omega * vs;
    const T vr = vs - omega * As[i];
    vp = vr + beta * (vp - omega * Ap[i]);
    acc[0] += static_cast<double>(vr * r0[i]);
    x[i] = vx;
    r[i] = vr;
    p[i] = vp;
  }
  block_store_partials<1, double>(acc, slot_ptr(part, 0), kMaxPartials, lw);
}

template <class M>
static int pipelined_prod(const M* A, const double* p, double* Ap, const double* r0, double* buf, uint32_t chunk,
                          uint32_t offset, void* stream)
{
  if (!p || !Ap || !buf || chunk == 0) return invalid("null argument");
  DevOp<double> op;
  VCLHIP_TRY(make_op(A, op));
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double* part;
  VCLHIP_TRY(ws.alloc(&part, kSlots * kMaxPartials));
  if (r0) {
    op.launch(s, p, EpiBicg<double, false, true>{Ap, p, r0, nullptr}, slot_ptr(part, 0), nullptr, 0u);
    VCLHIP_TRY(check_launch("pipelined bicgstab prod"));
    hipLaunchKernelGGL(write_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, op.grid, 3, chunk, 2 * chunk, offset,
                       buf, chunk);
  } else {
    op.launch(s, p, EpiCg<double, false>{Ap, p, nullptr}, slot_ptr(part, 0), nullptr, 0u);
    VCLHIP_TRY(check_launch("pipelined cg prod"));
    hipLaunchKernelGGL(write_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, op.grid, 2, chunk, 2 * chunk, 0u, buf,
                       chunk);
  }
  return check_launch("write_chunks_kernel");
}

}  // namespace vclhip

using namespace vclhip;

namespace vclhip {

// The copy's arrays in device memory owned by a vclhip_dia (freed by vclhip_dia_destroy).
struct OwnedDiaAlloc {
  vclhip_dia* d;
  bool operator()(size_t off_bytes, size_t val_bytes, DiaArrays* a) const
  {
    void* o = nullptr;
    void* v = nullptr;
    if (hipMalloc(&o, off_bytes) != hipSuccess || hipMalloc(&v, val_bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (o) (void)hipFree(o);
      return false;
    }
    d->off = static_cast<int32_t*>(o);
    d->val = v;
    a->off = d->off;
    a->val = v;
    return true;
  }
};

template <typename T>
static int dia_create(const vclhip_csr* A, vclhip_dia** out, void* stream)
{
  if (!out) return invalid("dia is NULL");
  *out = nullptr;
  if (!A || !A->row_ptr || !A->col_idx || !A->values) return invalid("CSR matrix arrays");
  hipStream_t s = static_cast<hipStream_t>(stream);
  vclhip_dia* d = new vclhip_dia();
  d->single = sizeof(T) == 4 ? 1 : 0;
  d->matrix = *A;
  d->built = false;
  d->val = nullptr;
  d->off = nullptr;
  int status = VCLHIP_SUCCESS;
  {
    Workspace ws(s);
    DiaDev<T> dia{};
    bool built = false;
    status = build_dia<T, T>(A, ws, s, 1.25, dia_memory_cap(), true, OwnedDiaAlloc{d}, &dia, &built);
    if (!status) {
      d->built = built;
      if constexpr (sizeof(T) == 4)
        d->d32 = dia;
      else
        d->d64 = dia;
    }
  }
  if (!status) status = hip_status(hipStreamSynchronize(s), "dia create");  // the workspace is released
  if (status) {
    (void)vclhip_dia_destroy(d);
    return status;
  }
  *out = d;
  return VCLHIP_SUCCESS;
}

}  // namespace vclhip

// ------------------------------------------------------------------ CG without allocations

// Scratch memory of vclhip_cg_solve_csr_f64_noalloc: a file-scope device array that every device the
// library runs on holds until the process ends.  Host code never passes `scratch` itself (its host
// name is a stand-in, not a device address): scratch_address() takes each device's own address once,
// with that device current, and the solver receives that address.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_addr[kScratchDevices] = {};

static int scratch_address(void** out)
{
  int dev = 0;
  VCLHIP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return invalid("device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    VCLHIP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return hip_status(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_addr[dev] = ptr;
  }
  *out = g_scratch_addr[dev];
  return VCLHIP_SUCCESS;
}

extern "C" {

int vclhip_dia_create_f64(const vclhip_csr* A, vclhip_dia** dia, void* stream)
{
  return vclhip::dia_create<double>(A, dia, stream);
}

int vclhip_dia_create_f32(const vclhip_csr* A, vclhip_dia** dia, void* stream)
{
  return vclhip::dia_create<float>(A, dia, stream);
}

int vclhip_dia_info(const vclhip_dia* dia, int* built, uint32_t* nd)
{
  if (!dia || !built || !nd) return vclhip::invalid("null argument");
  *built = dia->built ? 1 : 0;
  *nd = !dia->built ? 0u : (dia->single ? dia->d32.nd : dia->d64.nd);
  return VCLHIP_SUCCESS;
}

int vclhip_dia_destroy(vclhip_dia* dia)
{
  if (!dia) return VCLHIP_SUCCESS;
  int status = VCLHIP_SUCCESS;
  if (dia->val) status = vclhip::hip_status(hipFree(dia->val), "hipFree");
  if (dia->off) {
    const int st2 = vclhip::hip_status(hipFree(dia->off), "hipFree");
    if (!status) status = st2;
  }
  delete dia;
  return status;
}

void vclhip_solver_params_default(vclhip_solver_params* P)
{
  if (!P) return;
  std::memset(P, 0, sizeof(*P));
  P->tol = 1e-8;                     // cg_tag / bicgstab_tag / mixed_precision_cg_tag default
  P->abs_tol = 0;
  P->max_iters = 300;                // cg_tag default (cg.hpp:56)
  P->max_iters_before_restart = 200; // bicgstab_tag default (bicgstab.hpp:56)
  P->inner_tol = 1e-2f;              // mixed_precision_cg_tag default (mixed_precision_cg.hpp:56)
  P->chunk = 16;
  P->use_graph = 0;
  P->mixed_mode = 0;
  P->krylov_dim = 20;                // gmres_tag default (gmres.hpp:56)
}

int vclhip_cg_solve_csr_f64(const vclhip_csr* A, const double* b, double* x, int precond,
                            const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_sell_f64(const vclhip_sell* A, const double* b, double* x, int precond,
                             const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_csr_f64_noalloc(const vclhip_csr* A, const double* b, double* x, int precond,
                                    const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  if (P && P->use_graph) return invalid("use_graph must be 0: instantiating a graph allocates");
  void* arena = nullptr;
  VCLHIP_TRY(scratch_address(&arena));
  return cg_solve(A, b, x, precond, P, R, stream, arena, kScratchBytes);
}

int vclhip_bicgstab_solve_csr_f64(const vclhip_csr* A, const double* b, double* x, int precond,
                                  const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_bicgstab_solve_sell_f64(const vclhip_sell* A, const double* b, double* x, int precond,
                                   const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_csr_f32(const vclhip_csr* A, const float* b, float* x, int precond,
                            const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_sell_f32(const vclhip_sell* A, const float* b, float* x, int precond,
                             const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_bicgstab_solve_csr_f32(const vclhip_csr* A, const float* b, float* x, int precond,
                                  const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_bicgstab_solve_sell_f32(const vclhip_sell* A, const float* b, float* x, int precond,
                                   const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_f64(const vclhip_matrix* A, const double* b, double* x, int precond,
                        const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_cg_solve_f32(const vclhip_matrix* A, const float* b, float* x, int precond,
                        const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return cg_solve(A, b, x, precond, P, R, stream);
}

int vclhip_bicgstab_solve_f64(const vclhip_matrix* A, const double* b, double* x, int precond,
                              const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_bicgstab_solve_f32(const vclhip_matrix* A, const float* b, float* x, int precond,
                              const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return bicgstab_solve(A, b, x, precond, P, R, stream);
}

int vclhip_pipelined_cg_prod_f64(const vclhip_matrix* A, const double* p, double* Ap, double* buf, uint32_t chunk,
                                 void* stream)
{
  return pipelined_prod(A, p, Ap, nullptr, buf, chunk, 0u, stream);
}

int vclhip_pipelined_bicgstab_prod_f64(const vclhip_matrix* A, const double* p, double* Ap, const double* r0,
                                       double* buf, uint32_t chunk, uint32_t offset, void* stream)
{
  if (!r0) return invalid("null argument");
  return pipelined_prod(A, p, Ap, r0, buf, chunk, offset, stream);
}

int vclhip_matrix_diag_f64(const vclhip_matrix* A, double* d, void* stream)
{
  if (!d) return invalid("null argument");
  DevOp<double> op;
  VCLHIP_TRY(make_op(A, op));
  return matrix_diag(op, d, static_cast<hipStream_t>(stream));
}

int vclhip_matrix_diag_f32(const vclhip_matrix* A, float* d, void* stream)
{
  if (!d) return invalid("null argument");
  DevOp<float> op;
  VCLHIP_TRY(make_op(A, op));
  return matrix_diag(op, d, static_cast<hipStream_t>(stream));
}

int vclhip_mixed_cg_solve_csr(const vclhip_csr* A, const float* values32, const double* b, double* x, int precond,
                              const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return mixed_cg_solve(A, values32, b, x, precond, P, R, stream);
}

int vclhip_mixed_cg_solve_sell(const vclhip_sell* A, const float* elements32, const double* b, double* x,
                               int precond, const vclhip_solver_params* P, vclhip_solver_result* R, void* stream)
{
  return mixed_cg_solve(A, elements32, b, x, precond, P, R, stream);
}

int vclhip_pipelined_cg_vector_update_f64(uint32_t n, double* x, double alpha, double* p, double* r, const double* Ap,
                                          double beta, double* buf, uint32_t chunk, void* stream)
{
  if (!x || !p || !r || !Ap || !buf || chunk == 0) return invalid("null argument");
  if (n == 0) return VCLHIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double* part;
  VCLHIP_TRY(ws.alloc(&part, kMaxPartials));
  const uint32_t gv = vec_grid_n(n);
  hipLaunchKernelGGL(cg_update_scalar_kernel<double>, dim3(gv), dim3(kBlock), 0, s, n, x, alpha, p, r, Ap, beta, part);
  VCLHIP_TRY(check_launch("cg_update_scalar_kernel"));
  hipLaunchKernelGGL(write_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, gv, 1, 0u, 0u, 0u, buf, chunk);
  return check_launch("write_chunks_kernel");
}

int vclhip_pipelined_cg_prod_csr_f64(const vclhip_csr* A, const double* p, double* Ap, double* buf, uint32_t chunk,
                                     void* stream)
{
  return pipelined_prod(A, p, Ap, nullptr, buf, chunk, 0u, stream);
}

int vclhip_pipelined_cg_prod_sell_f64(const vclhip_sell* A, const double* p, double* Ap, double* buf, uint32_t chunk,
                                      void* stream)
{
  return pipelined_prod(A, p, Ap, nullptr, buf, chunk, 0u, stream);
}

// pipelined_gmres_prod is pipelined_cg_prod (host_based/iterative_operations.hpp:931-937)
int vclhip_pipelined_gmres_prod_csr_f64(const vclhip_csr* A, const double* p, double* Ap, double* buf, uint32_t chunk,
                                        void* stream)
{
  return pipelined_prod(A, p, Ap, nullptr, buf, chunk, 0u, stream);
}

int vclhip_pipelined_gmres_prod_sell_f64(const vclhip_sell* A, const double* p, double* Ap, double* buf,
                                         uint32_t chunk, void* stream)
{
  return pipelined_prod(A, p, Ap, nullptr, buf, chunk, 0u, stream);
}

int vclhip_pipelined_bicgstab_prod_csr_f64(const vclhip_csr* A, const double* p, double* Ap, const double* r0,
                                           double* buf, uint32_t chunk, uint32_t offset, void* stream)
{
  if (!r0) return invalid("r0star is NULL");
  return pipelined_prod(A, p, Ap, r0, buf, chunk, offset, stream);
}

int vclhip_pipelined_bicgstab_prod_sell_f64(const vclhip_sell* A, const double* p, double* Ap, const double* r0,
                                            double* buf, uint32_t chunk, uint32_t offset, void* stream)
{
  if (!r0) return invalid("r0star is NULL");
  return pipelined_prod(A, p, Ap, r0, buf, chunk, offset, stream);
}

int vclhip_pipelined_bicgstab_update_s_f64(uint32_t n, double* s_vec, const double* r, const double* Ap, double* buf,
                                           uint32_t chunk, uint32_t offset, void* stream)
{
  if (!s_vec || !r || !Ap || !buf || chunk == 0) return invalid("null argument");
  if (n == 0) return VCLHIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double *part, *alpha;
  VCLHIP_TRY(ws.alloc(&part, kMaxPartials));
  VCLHIP_TRY(ws.alloc(&alpha, 1));
  hipLaunchKernelGGL(bicg_buffer_alpha_kernel, dim3(1), dim3(64), 0, s, buf, chunk, alpha);
  VCLHIP_TRY(check_launch("bicg_buffer_alpha_kernel"));
  const uint32_t gv = vec_grid_n(n);
  hipLaunchKernelGGL(bicg_update_s_ptr_kernel<double>, dim3(gv), dim3(kBlock), 0, s, n, s_vec, r, Ap, alpha, part);
  VCLHIP_TRY(check_launch("bicg_update_s_ptr_kernel"));
  hipLaunchKernelGGL(write_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, gv, 1, offset, 0u, 0u, buf, chunk);
  return check_launch("write_chunks_kernel");
}

int vclhip_pipelined_bicgstab_vector_update_f64(uint32_t n, double* x, double alpha, double* p, double omega,
                                                const double* s_vec, double* r, const double* As, double beta,
                                                const double* Ap, const double* r0, double* buf, uint32_t chunk,
                                                void* stream)
{
  if (!x || !p || !s_vec || !r || !As || !Ap || !r0 || !buf || chunk == 0) return invalid("null argument");
  if (n == 0) return VCLHIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Workspace ws(s);
  double* part;
  VCLHIP_TRY(ws.alloc(&part, kMaxPartials));
  const uint32_t gv = vec_grid_n(n);
  hipLaunchKernelGGL(bicg_vec_update_scalar_kernel<double>, dim3(gv), dim3(kBlock), 0, s, n, x, alpha, p, omega, s_vec,
                     r, As, beta, Ap, r0, part);
  VCLHIP_TRY(check_launch("bicg_vec_update_scalar_kernel"));
  hipLaunchKernelGGL(write_chunks_kernel, dim3(1), dim3(kBlock), 0, s, part, gv, 1, 0u, 0u, 0u, buf, chunk);
  return check_launch("write_chunks_kernel");
}

}  // extern "C"
