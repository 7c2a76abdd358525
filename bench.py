#!/usr/bin/env python3
"""Benchmark of the sparse Krylov hot path on one MI355X.

Step = one FP64 CSR SpMV y = A x on the 3-D 7-point Laplacian 128^3 (n = 2,097,152,
nnz = 14,581,760) with all inputs resident in HBM.  Before each timed SpMV a 1 GiB read
evicts the 256 MB Infinity Cache (leaving only clean, unrelated lines), so `value` is an HBM
figure; the back-to-back (cache-assisted) rate is reported as `warm_gbs`.  The JSON line also
carries the roofline of the SpMV kernel, a streaming reference (z = a x + b y over 1.5 GB), the
CPU baseline (the reference's own OpenMP host backend, compiled from its sources into
oracle/_ref), and time-to-solution of the FP64 Krylov solvers.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --kernels-only     # only the SpMV and streaming kernels (for rocprofv3 --pmc)
"""
import argparse
import ctypes as C
import json
import os
import pathlib
import subprocess
import sys
import time

# The CPU baseline pins its OpenMP threads (close binding, one thread per core); libgomp reads
# these when the reference build is first loaded.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "viennacl-dev_amd"))

METRIC = "CG time-to-1e-10 rel. residual (FP64 vs mixed, 1 GPU); SpMV GB/s vs HBM peak"
HBM_PEAK_GBS = 8000.0
EVICT_DOUBLES = 128 << 20   # 1 GiB
STREAM_N = 64 << 20         # 3 x 512 MB for the streaming reference


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def csr_bytes(n, nnz, val_bytes=8, vec_bytes=8):
    """Algorithmic bytes of y = A x: values + column indices + row pointers + x once + y written."""
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * vec_bytes * n


def sell_bytes(stored, slices, n):
    """Stored values + column indices + per-slice (columns, start) + x once + y written."""
    return 12 * stored + 8 * slices + 16 * n


class Events:
    def __init__(self, vh, count):
        self.vh = vh
        self.ev = []
        for _ in range(count):
            h = C.c_void_p()
            vh.call("vclhip_event_create", C.byref(h))
            self.ev.append(h)

    def record(self, i, stream):
        self.vh.call("vclhip_event_record", self.ev[i], stream)

    def ms(self, i, j):
        t = C.c_float()
        self.vh.call("vclhip_event_elapsed_ms", self.ev[i], self.ev[j], C.byref(t))
        return t.value

    def close(self):
        for h in self.ev:
            self.vh.call("vclhip_event_destroy", h)


def evict_caches(vh, evict, stream):
    """Reads 1 GiB (4x the Infinity Cache) so that the caches hold only clean, unrelated lines."""
    r = C.c_double()
    vh.call("vclhip_norm2_f64", evict.size, evict.ptr, 0, 1, C.byref(r), stream.handle)


def time_launches(vh, launch, stream, steps, warmup, evict=None):
    """Kernel durations of `launch` (one SpMV or streaming kernel), each measured by events that the
    kernel itself records at its start and end (vclhip_time_next_launch).  Returns the per-launch
    cold durations (each after a 1 GiB read that evicts the caches; None without `evict`), the mean
    warm (back-to-back) duration, and the wall time of the cold loop."""
    for _ in range(warmup):
        launch()
    stream.synchronize()
    ev = Events(vh, 2 * steps)
    t0 = time.perf_counter()
    cold = None
    if evict is not None:
        for k in range(steps):
            evict_caches(vh, evict, stream)
            vh.call("vclhip_time_next_launch", ev.ev[2 * k], ev.ev[2 * k + 1])
            launch()
        stream.synchronize()
        cold = np.array([ev.ms(2 * k, 2 * k + 1) for k in range(steps)])
    wall = time.perf_counter() - t0
    for k in range(steps):
        vh.call("vclhip_time_next_launch", ev.ev[2 * k], ev.ev[2 * k + 1])
        launch()
    stream.synchronize()
    warm = float(np.mean([ev.ms(2 * k, 2 * k + 1) for k in range(steps)]))
    ev.close()
    return cold, warm, wall


def stream_reference(vh, stream, steps, evict=None, spmv_bytes=None, spmv_blocks=None):
    """The streaming rates of this device, kernel-timed like the SpMV:
    * axpby: z = a x + b y over 64M doubles (1.5 GB moved per launch);
    * read_1gib: a read-only kernel (vclhip_stream_read_f64) over 1 GiB, 16-byte loads, 8 in flight
      per lane, one contiguous chunk per workgroup: the read ceiling;
    * read_spmv_shaped: the same kernel over as many bytes as the CSR SpMV reads, one contiguous
      chunk per row block of the SpMV (about 24 KB each), cold (after the 1 GiB eviction read) and
      warm: what a perfect stream of the SpMV's bytes in its launch shape would take."""
    x = vh.DeviceArray.zeros(STREAM_N, np.float64)
    y = vh.DeviceArray.zeros(STREAM_N, np.float64)
    z = vh.DeviceArray.zeros(STREAM_N, np.float64)

    def launch():
        vh.call("vclhip_axpby_f64", STREAM_N, 1.5, x.ptr, -0.5, y.ptr, z.ptr, stream.handle)

    _, warm, _ = time_launches(vh, launch, stream, steps, 2)
    out = {"axpby": {"kernel": "axpby_kernel<double>", "bytes_per_launch": 24 * STREAM_N, "ms": round(warm, 6),
                     "gbs": round(24 * STREAM_N / (warm * 1e-3) / 1e9, 2)}}
    y.free()
    z.free()
    # read-only: 1 GiB (x and y's storage, 128M doubles; larger than the Infinity Cache)
    big = vh.DeviceArray.zeros(2 * STREAM_N, np.float64)
    sums = vh.DeviceArray.zeros(1 << 16, np.float64)
    grid = 8 * 256 * 8  # eight workgroups per CU
    _, warm, _ = time_launches(vh, lambda: vh.call("vclhip_stream_read_f64", C.c_uint64(2 * STREAM_N), big.ptr,
                                                   sums.ptr, grid, stream.handle), stream, steps, 2)
    out["read_1gib"] = {"kernel": "stream_read_kernel", "bytes_per_launch": 16 * STREAM_N, "workgroups": grid,
                        "ms": round(warm, 6), "gbs": round(16 * STREAM_N / (warm * 1e-3) / 1e9, 2)}
    if spmv_bytes and spmv_blocks and evict is not None:
        m = spmv_bytes // 8
        g = (spmv_blocks + 7) // 8 * 8
        cold, warm, _ = time_launches(vh, lambda: vh.call("vclhip_stream_read_f64", C.c_uint64(m), big.ptr, sums.ptr,
                                                          g, stream.handle), stream, steps, 2, evict)
        out["read_spmv_shaped"] = {"kernel": "stream_read_kernel", "bytes_per_launch": 8 * m, "workgroups": g,
                                   "cold_ms": round(float(cold.mean()), 6), "warm_ms": round(warm, 6),
                                   "cold_gbs": round(8 * m / (cold.mean() * 1e-3) / 1e9, 2),
                                   "warm_gbs": round(8 * m / (warm * 1e-3) / 1e9, 2)}
    big.free()
    sums.free()
    x.free()
    return out


def true_relres(vh, A, b, x, n):
    """||b - A x|| / ||b|| in FP64 on the device."""
    r = vh.DeviceArray.zeros(n, np.float64)
    vh.call("vclhip_memcpy_d2d", r.ptr, b.ptr, C.c_size_t(8 * n), None)
    A.spmv(x, r, -1.0, 1.0)
    nr, nb = C.c_double(), C.c_double()
    vh.call("vclhip_norm2_f64", n, r.ptr, 0, 1, C.byref(nr), None)
    vh.call("vclhip_norm2_f64", n, b.ptr, 0, 1, C.byref(nb), None)
    return nr.value / nb.value


def timed_solve(vh, kind, A, b, n, precond=0, reps=2, **kw):
    """Best-of-`reps` wall time of one solve (device-synchronized), with its FP64 true residual."""
    x = vh.DeviceArray.zeros(n, np.float64)
    best = None
    res = None
    for _ in range(reps):
        vh.call("vclhip_device_synchronize")
        t0 = time.perf_counter()
        if kind == "gmres":
            res = vh.gmres(A, b, x, **kw)
        else:
            res = (vh.cg if kind == "cg" else vh.bicgstab)(A, b, x, precond=precond, **kw)
        vh.call("vclhip_device_synchronize")
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    rel = true_relres(vh, A, b, x, n)
    out = {"seconds": round(best, 6), "iters": int(res.iters), "error": float(res.error),
           "true_relres": float(rel), "converged": bool(res.converged),
           "operator": {0: "as_given", 1: "dia", 2: "dia_coded"}.get(int(res.operator_used), int(res.operator_used))}
    if kind == "gmres":
        out["cycles"] = int(res.restarts)
    else:
        out["kernels_per_iter"] = int(res.kernels_per_iter)
        out["us_per_iter"] = round(best * 1e6 / max(int(res.iters), 1), 2)
    return out


def gmres_cycle_bytes(n, spmv_bytes, K, vec_bytes=8):
    """Algorithmic bytes of one GMRES(K) cycle as the HIP solver runs it: K SpMVs (the first also
    reads r), per step k >= 1 one pass over k + 1 basis vectors (<v_j, v_k>) and one reading k + 2
    and writing one (v_k -= sum h_j v_j, ||v_k||^2, <r, v_k>), the update x += c0 r + sum c_j v_{j-1}
    (x read and written, r and K - 1 basis vectors read) and the restart r = b - A x (SpMV plus b)."""
    passes = 1 + sum(2 * k + 4 for k in range(1, K)) + (K + 2) + 1
    return K * spmv_bytes + spmv_bytes + passes * vec_bytes * n


def gmres_bench(vh, grid, reps=2):
    """Nonsymmetric 7-point convection-diffusion (upwind, beta = (0.5, 0.3, 0.2)) on a grid^3 mesh,
    GMRES(K) without preconditioner to 1e-8 (gmres_tag), CSR and SELL-64, FP64 and FP32."""
    rp, col, val = vh.convection_diffusion(grid, grid, grid)
    n, nnz = rp.size - 1, int(rp[-1])
    A = vh.CsrMatrix(rp, col, val)
    S = vh.SellMatrix(rp, col, val, 64)
    b = vh.DeviceArray.from_numpy(vh.random_uniform(0, n))
    out = {"workload": f"convection-diffusion {grid}^3 (n={n}, nnz={nnz}), GMRES(K) to 1e-8, no preconditioner",
           "n": n, "nnz": nnz}
    spmv = csr_bytes(n, nnz)
    for name, M, K, kw in (("gmres20_fp64_csr", A, 20, {}),
                           ("gmres20_fp64_csr_dia_uncoded", A, 20, {"operator_format": vh.OPERATOR_DIA_UNCODED}),
                           ("gmres20_fp64_csr_graph", A, 20, {"use_graph": True}),
                           ("gmres20_fp64_sell64", S, 20, {}), ("gmres40_fp64_csr", A, 40, {}),
                           ("gmres10_fp64_csr", A, 10, {})):
        rec = timed_solve(vh, "gmres", M, b, n, reps=reps, tol=1e-8, max_iters=6000, krylov_dim=K, **kw)
        moved = rec["cycles"] * gmres_cycle_bytes(n, spmv, K)
        rec["ms_per_step"] = round(rec["seconds"] * 1e3 / max(rec["iters"], 1), 4)
        rec["effective_gbs"] = round(moved / rec["seconds"] / 1e9, 1)
        out[name] = rec
        log("gmres", name, rec)
    # FP32 to 1e-5 (FP32 matrix and basis, FP64-accumulated inner products)
    A32 = vh.CsrMatrix(rp, col, val, dtype=np.float32)
    b32 = vh.DeviceArray.from_numpy(b.to_numpy().astype(np.float32))
    x32 = vh.DeviceArray.zeros(n, np.float32)
    best = None
    for _ in range(reps):
        vh.call("vclhip_device_synchronize")
        t0 = time.perf_counter()
        res = vh.gmres(A32, b32, x32, tol=1e-5, max_iters=6000, krylov_dim=20)
        vh.call("vclhip_device_synchronize")
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    x64 = vh.DeviceArray.from_numpy(x32.to_numpy().astype(np.float64))
    out["gmres20_fp32_csr_1e-5"] = {"seconds": round(best, 6), "iters": int(res.iters), "cycles": int(res.restarts),
                                    "error": float(res.error), "converged": bool(res.converged),
                                    "true_relres_fp64": float(true_relres(vh, A, b, x64, n))}
    log("gmres", "gmres20_fp32_csr_1e-5", out["gmres20_fp32_csr_1e-5"])
    for a in (b, b32, x32, x64):
        a.free()
    for M in (A, S, A32):
        M.free()
    return out


def powerlaw_bench(vh, stream, steps):
    """Random nonsymmetric matrix, n = 2,097,152, row lengths P(k) ~ k^-2 on 1..512 (seed 1),
    diagonal 1.1 x row abs-sum: cold CSR SpMV rate and BiCGStab (plain, Jacobi) to 1e-10."""
    n = 2097152
    t = time.perf_counter()
    rp, col, val = vh.powerlaw(n, 512, 2.0, 1.1, 1)
    nnz = int(rp[-1])
    log(f"generated power-law matrix: n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    A = vh.CsrMatrix(rp, col, val)
    x = vh.DeviceArray.from_numpy(vh.random_uniform(1, n))
    y = vh.DeviceArray.zeros(n, np.float64)
    evict = vh.DeviceArray(EVICT_DOUBLES, np.float64)
    vh.call("vclhip_fill_zero", evict.ptr, C.c_size_t(evict.nbytes), None)
    cold, warm, _ = time_launches(vh, lambda: A.spmv(x, y, stream=stream.handle), stream, steps, 3, evict)
    # the x gathers alone (the matrix's own column indices, the SpMV's launch shape), cold as well
    g = (A.num_blocks + 7) // 8 * 8
    sums = vh.DeviceArray.zeros(g, np.float64)
    gcold, gwarm, _ = time_launches(vh, lambda: vh.call("vclhip_gather_read_f64", nnz, A.d_col.ptr, x.ptr, n, sums.ptr,
                                                        g, stream.handle), stream, steps, 3, evict)
    sums.free()
    evict.free()
    by = csr_bytes(n, nnz)
    out = {"workload": f"power-law nonsymmetric matrix n={n} nnz={nnz}, CSR, tol 1e-10", "n": n, "nnz": nnz,
           "spmv_cold_us": round(float(cold.mean()) * 1e3, 2), "spmv_cold_gbs": round(float(by / (cold.mean() * 1e-3) / 1e9), 1),
           "spmv_warm_gbs": round(by / (warm * 1e-3) / 1e9, 1), "algorithmic_bytes_per_spmv": by,
           "gather_floor": {"kernel": "gather_read_kernel", "what": "x[col[e]] summed over every entry, the SpMV's "
                            "column indices and launch shape, no values, row pointers or y",
                            "cold_us": round(float(gcold.mean()) * 1e3, 2), "warm_us": round(gwarm * 1e3, 2),
                            "spmv_over_gathers": round(float(cold.mean() / gcold.mean()), 3)}}
    b = vh.DeviceArray.from_numpy(vh.random_uniform(0, n))
    # without a preconditioner BiCGStab stagnates on this badly scaled matrix; 400 iterations show it
    out["bicgstab_fp64_csr_400"] = timed_solve(vh, "bicgstab", A, b, n, reps=1, tol=1e-10, max_iters=400)
    out["jacobi_bicgstab_fp64_csr"] = timed_solve(vh, "bicgstab", A, b, n, precond=1, tol=1e-10, max_iters=5000)
    log("powerlaw", out)
    for a in (x, y, b):
        a.free()
    A.free()
    return out


def formats_bench(vh, stream, steps):
    """Cold and warm SpMV of coordinate_matrix, ell_matrix and hyb_matrix beside CSR: the 128^3
    Laplacian and the n = 2M power-law matrix (where HYB moves the few long rows out of the ELL
    part).  Algorithmic bytes: the stored arrays the kernel must read (ELL padding values included,
    their columns are not read), x once and y written."""
    out = {}
    evict = vh.DeviceArray(EVICT_DOUBLES, np.float64)
    vh.call("vclhip_fill_zero", evict.ptr, C.c_size_t(evict.nbytes), None)
    for wname, gen in (("lap3d_128", lambda: vh.laplace_3d(128)),
                       ("powerlaw_2M", lambda: vh.powerlaw(2097152, 512, 2.0, 1.1, 1))):
        rp, col, val = gen()
        n, nnz = rp.size - 1, int(rp[-1])
        x = vh.DeviceArray.from_numpy(vh.random_uniform(1, n))
        y = vh.DeviceArray.zeros(n, np.float64)
        rec = {"n": n, "nnz": nnz}
        mats = {"csr": (vh.CsrMatrix(rp, col, val), csr_bytes(n, nnz)),
                "coo": (vh.CooMatrix(rp, col, val), nnz * 16 + 4 * (n + 1) + 16 * n)}
        width = int(np.diff(rp.astype(np.int64)).max())
        if n * width <= 6 * nnz:
            E = vh.EllMatrix(rp, col, val)
            mats["ell"] = (E, nnz * 12 + (E.ld * E.width - nnz) * 8 + 16 * n)
        else:  # the power-law matrix's ELL would store 512 columns for every row
            rec["ell_skipped"] = f"ELL width {width} would store {n * width / nnz:.0f}x the entries"
        # HYB with the plan of its CSR part (long CSR parts summed by a wavefront), and HYB summed one
        # lane per row in the reference's order
        H = vh.HybMatrix(rp, col, val)
        csr_real = int(H.host[2][-1])
        ell_real = nnz - csr_real
        hyb_bytes = ell_real * 12 + (H.ld * H.width - ell_real) * 8 + csr_real * 12 + 4 * (n + 1) + 16 * n
        mats["hyb"] = (H, hyb_bytes)
        mats["hyb_lane_per_row"] = (vh.HybMatrix(rp, col, val, plan=False), hyb_bytes)
        for name, (M, by) in mats.items():
            cold, warm, _ = time_launches(vh, lambda: M.spmv(x, y, stream=stream.handle), stream, steps, 3, evict)
            cold_ms = float(cold.mean())
            rec[name] = {"cold_us": round(cold_ms * 1e3, 2), "warm_us": round(warm * 1e3, 2),
                         "algorithmic_bytes": int(by), "cold_gbs": round(by / (cold_ms * 1e-3) / 1e9, 1),
                         "warm_gbs": round(by / (warm * 1e-3) / 1e9, 1)}
            M.free()
        rec["hyb_ell_width"], rec["hyb_csr_nnz"] = H.width, H.csr_nnz
        out[wname] = rec
        log("formats", wname, rec)
        x.free()
        y.free()
    evict.free()
    return out


def mixed_vs_fp64(vh, grid, reps=2):
    """FP64 CG vs mixed-precision CG to ||b - A x|| / ||b|| <= 1e-10 on the 3-D variable-coefficient
    diffusion operator (k = exp(g), g ~ N(0,1) per cell, harmonic-mean faces; seed 1 for k, 0 for b)."""
    t = time.perf_counter()
    n = grid ** 3
    k = np.exp(vh.random_normal(1, n))
    rp, col, val = vh.diffusion(grid, grid, grid, k)
    nnz = int(rp[-1])
    log(f"generated diffusion {grid}^3: n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    A = vh.CsrMatrix(rp, col, val)
    b = vh.DeviceArray.from_numpy(vh.random_uniform(0, n))
    out = {"workload": f"3-D variable-coefficient diffusion {grid}^3 (n={n}, nnz={nnz}), CSR, tol 1e-10, "
                       f"b ~ U[-1,1]", "n": n, "nnz": nnz}

    def run(name, fn, x):
        best, res = None, None
        for _ in range(reps):
            vh.call("vclhip_device_synchronize")
            t0 = time.perf_counter()
            res = fn(x)
            vh.call("vclhip_device_synchronize")
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        rec = {"seconds": round(best, 6), "iters": int(res.iters), "error": float(res.error),
               "true_relres": float(true_relres(vh, A, b, x, n)), "converged": bool(res.converged)}
        if "mixed" in name:
            rec["fp64_corrections"] = int(res.restarts)
        out[name] = rec
        log(name, rec)
        return rec

    x64 = vh.DeviceArray.zeros(n, np.float64)
    f64 = run("cg_fp64", lambda x: vh.cg(A, b, x, tol=1e-10, max_iters=20000), x64)
    x64h = x64.to_numpy()
    xg = vh.DeviceArray.zeros(n, np.float64)
    g64 = run("cg_fp64_csr_as_given", lambda x: vh.cg(A, b, x, tol=1e-10, max_iters=20000, operator_format=1), xg)
    g64["same_x_as_cg_fp64"] = bool(np.array_equal(xg.to_numpy(), x64h))
    xg.free()
    xm = vh.DeviceArray.zeros(n, np.float64)
    for name, kw in (("mixed_reliable", dict(mixed_mode=1)),
                     ("mixed_reliable_sell16_inner", dict(mixed_mode=1, inner_format=3)),
                     ("mixed_reliable_csr_inner", dict(mixed_mode=1, inner_format=1)),
                     ("mixed_reference_schedule", dict(mixed_mode=0))):
        rec = run(name, lambda x: vh.mixed_cg(A, b, x, tol=1e-10, max_iters=20000, **kw), xm)
        xh = xm.to_numpy()
        rec["rel_error_vs_fp64_solution"] = float(np.linalg.norm(xh - x64h) / np.linalg.norm(x64h))
        rec["time_ratio_vs_fp64"] = round(rec["seconds"] / f64["seconds"], 4)
        rec["time_ratio_vs_fp64_csr_as_given"] = round(rec["seconds"] / g64["seconds"], 4)
    S = vh.SellMatrix(rp, col, val, 64)
    s64 = run("cg_fp64_sell64", lambda x: vh.cg(S, b, x, tol=1e-10, max_iters=20000), xm)
    rec = run("mixed_reliable_sell64", lambda x: vh.mixed_cg(S, b, x, tol=1e-10, max_iters=20000, mixed_mode=1), xm)
    rec["rel_error_vs_fp64_solution"] = float(np.linalg.norm(xm.to_numpy() - x64h) / np.linalg.norm(x64h))
    rec["time_ratio_vs_fp64"] = round(rec["seconds"] / f64["seconds"], 4)
    rec["time_ratio_vs_fp64_sell64"] = round(rec["seconds"] / s64["seconds"], 4)
    xj = vh.DeviceArray.zeros(n, np.float64)
    j64 = run("jacobi_cg_fp64", lambda x: vh.cg(A, b, x, precond=1, tol=1e-10, max_iters=20000), xj)
    xjh = xj.to_numpy()
    rec = run("jacobi_mixed_reliable", lambda x: vh.mixed_cg(A, b, x, precond=1, tol=1e-10, max_iters=20000,
                                                              mixed_mode=1), xm)
    rec["rel_error_vs_fp64_solution"] = float(np.linalg.norm(xm.to_numpy() - xjh) / np.linalg.norm(xjh))
    rec["time_ratio_vs_fp64"] = round(rec["seconds"] / j64["seconds"], 4)
    for a in (x64, xm, xj, b):
        a.free()
    A.free()
    return out


def batched_cg_bench(vh, num_systems, reps=3):
    """4096 graph-Laplacian systems (n_i ~ U{64..4096}, average degree 8, weights U[0.1,1], shift 0.01),
    per-system tol 1e-8, autotuned lanes per row, one persistent launch replayed as a hipGraph."""
    t = time.perf_counter()
    sys_ptr, rp, col, val = vh.graph_laplacian_batch(num_systems, 64, 4096, 8.0, 0.1, 1.0, 0.01, 1)
    n, nnz = rp.size - 1, int(rp[-1])
    log(f"generated batch of {num_systems} systems: rows={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    b = vh.random_uniform(0, n)
    out = {"workload": f"{num_systems} graph-Laplacian systems, n_i ~ U{{64..4096}} (total rows {n}, nnz {nnz}), "
                       f"CSR, per-system tol 1e-8, b ~ U[-1,1]", "rows": n, "nnz": nnz}
    plan = vh.BatchPlan(sys_ptr, rp)
    stream = vh.Stream()
    # FP32 at the configuration's 1e-8 stops on its recursive residual, which drifts from the true
    # one by about the FP32 resolution times the condition number; the FP32 run at 1e-5 is one whose
    # tolerance the true residual can meet.  "converged" counts systems whose FP64 true residual is
    # at most the tolerance; "converged_recursive" the device's own flags.
    for name, dtype, tol in (("fp64", np.float64, 1e-8), ("fp32", np.float32, 1e-8), ("fp32_tol1e-5", np.float32, 1e-5)):
        A = vh.CsrMatrix(rp, col, val, dtype=dtype)
        t0 = time.perf_counter()
        plan.autotune(A, single=(dtype == np.float32), stream=stream.handle)
        tune_s = time.perf_counter() - t0
        db = vh.DeviceArray.from_numpy(b.astype(dtype))
        dx = vh.DeviceArray.zeros(n, dtype)
        it = vh.DeviceArray.zeros(num_systems, np.uint32)
        er = vh.DeviceArray.zeros(num_systems, np.float64)
        cv = vh.DeviceArray.zeros(num_systems, np.uint32)
        ev = Events(vh, 2)
        best = None
        for r in range(reps + 1):
            ev.record(0, stream.handle)
            plan.solve_into(A, db, dx, it, er, cv, stream.handle, tol=tol, max_iters=20000, use_graph=True)
            ev.record(1, stream.handle)
            stream.synchronize()
            ms = ev.ms(0, 1)
            if r > 0:
                best = ms if best is None else min(best, ms)
        ev.close()
        iters, conv = it.to_numpy(), cv.to_numpy()
        # per-system FP64 true residual with the FP64 matrix
        A64 = A if dtype == np.float64 else vh.CsrMatrix(rp, col, val)
        x64 = vh.DeviceArray.from_numpy(dx.to_numpy().astype(np.float64))
        bb = b.astype(dtype).astype(np.float64)
        y = vh.DeviceArray.zeros(n, np.float64)
        A64.spmv(x64, y)
        r = bb - y.to_numpy()
        seg = np.add.reduceat(r * r, sys_ptr[:-1].astype(np.int64))
        segb = np.add.reduceat(bb * bb, sys_ptr[:-1].astype(np.int64))
        rel = np.sqrt(seg / segb)
        vb = 4 if dtype == np.float32 else 8
        # bytes per system iteration: values + 16-bit column offsets + row pointers (the vectors stay
        # on chip), plus the 6 bytes per entry of the column conversion once per solve; the CSR
        # equivalent counts 32-bit columns instead
        rows_i = np.diff(sys_ptr).astype(np.float64)
        nnz_i = np.diff(rp[sys_ptr].astype(np.int64)).astype(np.float64)
        moved = float(np.sum(iters * ((vb + 2) * nnz_i + 4 * (rows_i + 1)))) + 6.0 * nnz
        csr_equiv = float(np.sum(iters * ((vb + 4) * nnz_i + 4 * (rows_i + 1))))
        out[name] = {"tol": tol, "ms": round(best, 4), "systems_per_s": round(num_systems / (best * 1e-3), 1),
                     "iters_total": int(iters.sum()), "iters_max": int(iters.max()), "iters_mean": float(iters.mean()),
                     "converged": int((rel <= tol).sum()), "converged_recursive": int((conv == 1).sum()),
                     "max_true_relres": float(rel.max()),
                     "median_true_relres": float(np.median(rel)), "matrix_gbs": round(moved / (best * 1e-3) / 1e9, 1),
                     "csr_equivalent_gbs": round(csr_equiv / (best * 1e-3) / 1e9, 1),
                     "autotune_s": round(tune_s, 3), "lanes_log2_by_size_class": plan.lanes(dtype == np.float32).tolist()}
        log("batched", name, out[name])
        for a in (db, dx, it, er, cv, x64, y):
            a.free()
        A.free()
        if A64 is not A:
            A64.free()
    plan.close()
    return out


def adjoint_bench(vh, grid, n_fd=64):
    """Differentiable solve on the 2-D variable-coefficient diffusion operator (theta = exp(N(0,0.5))),
    L = 0.5 ||x - x_t||^2: forward CG + adjoint CG + fused gradient kernel, chained to dL/dtheta by
    torch; central finite differences (h = 1e-6) on n_fd random entries."""
    import torch
    from vclhip import torch_solve as ts
    nx = ny = grid
    n = nx * ny
    theta = np.exp(0.5 * vh.random_normal(1, n))
    b = vh.random_uniform(0, n)
    xt = vh.random_uniform(2, n)
    rp, col, pos, masks = ts.diffusion_pattern_2d(nx, ny)
    pat = ts.CsrPattern(rp, col)
    dev = torch.device("cuda")
    bt = torch.tensor(b, dtype=torch.float64, device=dev)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev)

    def loss(th, tol):
        vals = ts.diffusion_values_2d(th, nx, ny, pos, masks)
        x = ts.sparse_solve(vals, bt, pat, tol=tol, max_iters=50000)
        return 0.5 * torch.sum((x - xtt) ** 2)

    best = None
    g = None
    for r in range(3):
        th = torch.tensor(theta.reshape(ny, nx), dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = loss(th, 1e-10)
        L.backward()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if r > 0:
            best = dt if best is None else min(best, dt)
        g = th.grad.detach().cpu().numpy().ravel()
    rng = np.random.default_rng(0)
    idx = rng.choice(n, n_fd, replace=False)
    h = 1e-6
    th_ref = torch.tensor(theta.reshape(ny, nx), dtype=torch.float64, device=dev, requires_grad=True)
    loss(th_ref, 1e-13).backward()
    gref = th_ref.grad.detach().cpu().numpy().ravel()
    errs = []
    with torch.no_grad():
        for i in idx:
            tp, tm = theta.copy(), theta.copy()
            tp[i] += h
            tm[i] -= h
            fd = (loss(torch.tensor(tp.reshape(ny, nx), device=dev), 1e-13).item()
                  - loss(torch.tensor(tm.reshape(ny, nx), device=dev), 1e-13).item()) / (2 * h)
            errs.append(abs(fd - gref[i]))
    # errors on a floor of 1e-2 |g|_inf: a finite difference of L (about 2.5e7 here) at h = 1e-6 carries
    # an absolute error of about 1e-2 from the rounding of L and of the perturbed matrix, whatever g_i is
    g_inf = float(np.abs(gref).max())
    scaled = [e / max(abs(gref[i]), 1e-2 * g_inf) for e, i in zip(errs, idx)]
    out = {"workload": f"2-D variable-coefficient diffusion {grid}^2 (n={n}, nnz={int(rp[-1])}), FP64 CG forward "
                       f"and adjoint to 1e-10, dL/dA on the CSR pattern chained to dL/dtheta",
           "forward_backward_s": round(best, 6), "fd_entries": n_fd, "fd_h": h,
           "fd_max_abs_err": float(max(errs)), "fd_max_err_on_floor": float(max(scaled)),
           "fd_median_err_on_floor": float(np.median(scaled)), "fd_floor": "max(|g_i|, 1e-2 |g|_inf)",
           "grad_inf": g_inf, "grad_norm": float(np.linalg.norm(g))}
    log("adjoint", out)
    return out


def host_solve(po, R, rp, col, val, b, tol, max_iters, kind, reps=1):
    """The reference's own solve() on its host backend (ref_bench_solve: kind 0 CG, 2 mixed-precision
    CG), best of `reps`, timed around solve() only."""
    rp, col, val, b = (np.ascontiguousarray(a) for a in (rp, col, val, b))
    x = np.zeros(rp.size - 1)
    it, err = C.c_uint32(), C.c_double()
    best = None
    for _ in range(reps):
        t = R.ref_bench_solve(rp.size - 1, po.p(rp), po.p(col), po.p(val), po.p(b), po.p(x), tol, max_iters, kind, 0,
                              C.byref(it), C.byref(err))
        best = t if best is None else min(best, t)
    r = b - po.csr_spmv(rp, col, val, x)
    return {"seconds": round(best, 4), "iters": int(it.value), "error": float(err.value),
            "ms_per_iter": round(best * 1e3 / max(it.value, 1), 4),
            "true_relres": float(np.linalg.norm(r) / np.linalg.norm(b))}


def cpu_solver_baseline(vh, po, R, mixed_grid):
    """The CG metric on the host: the reference's own CG (cg.hpp:128-187) on the 128^3 Laplacian
    (C2) and the FP64 and mixed-precision CG (mixed_precision_cg.hpp:95-187) on the 160^3 diffusion
    operator (C3), same inputs as the device runs, to 1e-10."""
    out = {}
    rp, col, val = vh.laplace_3d(128)
    out["c2_cg_lap3d_128"] = host_solve(po, R, rp, col, val, vh.random_uniform(0, rp.size - 1), 1e-10, 5000, 0)
    log("CPU baseline c2", out["c2_cg_lap3d_128"])
    if mixed_grid:
        g = mixed_grid
        k = np.exp(vh.random_normal(1, g ** 3))
        rp, col, val = vh.diffusion(g, g, g, k)
        b = vh.random_uniform(0, rp.size - 1)
        out[f"c3_cg_fp64_diffusion_{g}"] = host_solve(po, R, rp, col, val, b, 1e-10, 20000, 0)
        log("CPU baseline c3 fp64", out[f"c3_cg_fp64_diffusion_{g}"])
        out[f"c3_mixed_cg_diffusion_{g}"] = host_solve(po, R, rp, col, val, b, 1e-10, 20000, 2)
        log("CPU baseline c3 mixed", out[f"c3_mixed_cg_diffusion_{g}"])
    return out


def cpu_baseline(rp, col, val, x, n, nnz, seconds, c1=None):
    """The reference host backend (OpenMP) timed on this machine's cores over a bounded sample; with
    c1 = (rp, col, val, b), also the reference's own CG solve of the 256x256 case."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    if not po.ref_available():
        return None
    R = po.ref()
    y = np.zeros(n)
    one = R.ref_bench_csr_spmv(n, po.p(rp), po.p(col), po.p(val), po.p(x), po.p(y), 1)
    reps = max(3, int(seconds / max(one, 1e-6)))
    t = R.ref_bench_csr_spmv(n, po.p(rp), po.p(col), po.p(val), po.p(x), po.p(y), reps)
    gbs = csr_bytes(n, nnz) * reps / t / 1e9
    out = {"value": round(gbs, 2), "unit": "GB/s", "cores": int(R.ref_max_threads()), "kind": "reference",
           "sample": f"{reps} FP64 CSR SpMVs y=Ax of the same 128^3 Laplacian (viennacl::linalg::prod, "
                     f"host_based OpenMP backend), {t:.1f} s",
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
           "host_cpus": os.cpu_count()}
    if c1 is not None:
        crp, ccol, cval, cb = (np.ascontiguousarray(a) for a in c1)
        cx = np.zeros(crp.size - 1)
        it, err = C.c_uint32(), C.c_double()
        best = None
        for _ in range(3):
            ts = R.ref_bench_solve(crp.size - 1, po.p(crp), po.p(ccol), po.p(cval), po.p(cb), po.p(cx), 1e-8, 2000,
                                   0, 0, C.byref(it), C.byref(err))
            best = ts if best is None else min(best, ts)
        out["c1_cg_fdm_256"] = {"seconds": round(best, 6), "iters": int(it.value), "error": float(err.value),
                                "sample": "solve(A, b, cg_tag(1e-8, 2000)), generate_fdm_laplace(256, 256), "
                                          "best of 3"}
    return out, y


def latency_reference(vh):
    """What a kernel boundary and a device-wide synchronisation cost on this device, the two ways of
    ordering one solver iteration after the previous one: back-to-back launches of an empty kernel
    (256 workgroups), and steps of vclhip_grid_sync_probe (co-resident workgroups that publish their
    stores to all XCDs and synchronise through flags, as a single-launch solver would per
    iteration), each timed by events around the whole sequence."""
    stream = vh.Stream()
    ev = Events(vh, 2)
    out = {}
    launches = 2000
    for _ in range(2):  # the second pass is the measurement
        ev.record(0, stream.handle)
        for _ in range(launches):
            vh.call("vclhip_empty_launch", 256, stream.handle)
        ev.record(1, stream.handle)
        stream.synchronize()
    out["empty_kernel_us_per_launch"] = round(ev.ms(0, 1) * 1e3 / launches, 3)
    flags = vh.DeviceArray.zeros(1024, np.uint32)
    steps = 2000
    sync = {}
    for g in (64, 256, 512):
        best = None
        for _ in range(2):
            ev.record(0, stream.handle)
            vh.call("vclhip_grid_sync_probe", g, steps, flags.ptr, stream.handle)
            ev.record(1, stream.handle)
            stream.synchronize()
            gave_up = int(flags.to_numpy()[g + 1])
            if gave_up:
                raise RuntimeError(f"grid_sync_probe: a workgroup gave up waiting (workgroups {g})")
            t = ev.ms(0, 1)
            best = t if best is None else min(best, t)
        sync[str(g)] = round(best * 1e3 / steps, 3)
    out["grid_sync_us_per_step_by_workgroups"] = sync
    flags.free()
    ev.close()
    stream.close()
    log("latency reference", out)
    return out


def c1_solve_bench(vh):
    """The reference's CPU-runnable case on the device: generate_fdm_laplace(256, 256) (n = 65,536,
    nnz = 326,656), b ~ U[-1,1] (seed 0), cg_tag(1e-8, 2000).  5 MB per SpMV stays in the caches,
    so the device solve is bound by launch latency, not bandwidth; hipGraph chunks take most of the
    launches off the host."""
    rp, col, val = vh.fdm_laplace_2d(256, 256)
    n = rp.size - 1
    bh = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val)
    b = vh.DeviceArray.from_numpy(bh)
    rec = {"workload": f"generate_fdm_laplace(256, 256) (n={n}, nnz={int(rp[-1])}), CG to 1e-8", "n": n,
           "nnz": int(rp[-1])}
    rec["cg_fp64_csr"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-8, max_iters=2000)
    rec["cg_fp64_csr_graph"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-8, max_iters=2000, use_graph=True)
    rec["cg_fp64_csr_graph_chunk64"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-8, max_iters=2000,
                                                   use_graph=True, chunk=64)
    rec["cg_fp64_csr_two_kernels"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-8, max_iters=2000, variant=1)
    rec["latency_reference"] = latency_reference(vh)
    log("c1 cg", rec)
    b.free()
    A.free()
    return rec, (rp, col, val, bh)


def profile_workloads(vh, names, steps, warmup):
    """Short runs of single kernels for rocprofv3 counter passes (one workload per dispatch group,
    each printed as it starts)."""
    out = {}
    stream = vh.Stream()
    for name in names:
        log("profile workload", name)
        if name == "spmv":
            rp, col, val = vh.laplace_3d(128)
            n = rp.size - 1
            A, S = vh.CsrMatrix(rp, col, val), vh.SellMatrix(rp, col, val, 64)
            x, y = vh.DeviceArray.from_numpy(vh.random_uniform(1, n)), vh.DeviceArray.zeros(n, np.float64)
            _, warm, _ = time_launches(vh, lambda: A.spmv(x, y, stream=stream.handle), stream, steps, warmup)
            _, swarm, _ = time_launches(vh, lambda: S.spmv(x, y, stream=stream.handle), stream, steps, warmup)
            out[name] = {"csr_ms": warm, "sell_ms": swarm, "stream": stream_reference(vh, stream, steps)}
            for a in (x, y, A, S):
                a.free()
        elif name == "powerlaw":
            n = 2097152
            rp, col, val = vh.powerlaw(n, 512, 2.0, 1.1, 1)
            A = vh.CsrMatrix(rp, col, val)
            x, y = vh.DeviceArray.from_numpy(vh.random_uniform(1, n)), vh.DeviceArray.zeros(n, np.float64)
            _, warm, _ = time_launches(vh, lambda: A.spmv(x, y, stream=stream.handle), stream, steps, warmup)
            # then the gathers alone (gather_read_kernel over the matrix's column indices)
            g = (A.num_blocks + 7) // 8 * 8
            sums = vh.DeviceArray.zeros(g, np.float64)
            _, gwarm, _ = time_launches(vh, lambda: vh.call("vclhip_gather_read_f64", int(rp[-1]), A.d_col.ptr, x.ptr,
                                                            n, sums.ptr, g, stream.handle), stream, steps, warmup)
            out[name] = {"csr_ms": warm, "gather_ms": gwarm, "nnz": int(rp[-1]),
                         "algorithmic_bytes": csr_bytes(n, int(rp[-1]))}
            for a in (x, y, A, sums):
                a.free()
        elif name == "dia_cg":
            g = 160
            rp, col, val = vh.diffusion(g, g, g, np.exp(vh.random_normal(1, g ** 3)))
            A = vh.CsrMatrix(rp, col, val)
            b, x = vh.DeviceArray.from_numpy(vh.random_uniform(0, g ** 3)), vh.DeviceArray.zeros(g ** 3, np.float64)
            res = vh.cg(A, b, x, tol=1e-30, max_iters=50, stream=stream.handle)
            stream.synchronize()
            out[name] = {"iters": int(res.iters), "operator_used": int(res.operator_used)}
            for a in (b, x, A):
                a.free()
        elif name in ("c1_cg", "c1_bicgstab", "c2_cg", "c2_bicgstab", "c3_cg_fp64", "c3_mixed"):
            # one solve of the configuration as the bench line times it
            if name.startswith("c1"):
                rp, col, val = vh.fdm_laplace_2d(256, 256)
                kw = dict(tol=1e-8, max_iters=2000)
            elif name.startswith("c2"):
                rp, col, val = vh.laplace_3d(128)
                kw = dict(tol=1e-10, max_iters=5000)
            else:
                g = 160
                rp, col, val = vh.diffusion(g, g, g, np.exp(vh.random_normal(1, g ** 3)))
                kw = dict(tol=1e-10, max_iters=20000)
            n = rp.size - 1
            A = vh.CsrMatrix(rp, col, val)
            b, x = vh.DeviceArray.from_numpy(vh.random_uniform(0, n)), vh.DeviceArray.zeros(n, np.float64)
            if name.endswith("bicgstab"):
                res = vh.bicgstab(A, b, x, stream=stream.handle, **kw)
            elif name == "c3_mixed":
                res = vh.mixed_cg(A, b, x, stream=stream.handle, mixed_mode=1, **kw)
            else:
                res = vh.cg(A, b, x, stream=stream.handle, **kw)
            stream.synchronize()
            out[name] = {"n": n, "iters": int(res.iters), "error": float(res.error), "converged": bool(res.converged)}
            for a in (b, x, A):
                a.free()
        elif name == "batched":
            sys_ptr, rp, col, val = vh.graph_laplacian_batch(4096, 64, 4096, 8.0, 0.1, 1.0, 0.01, 1)
            n = rp.size - 1
            plan = vh.BatchPlan(sys_ptr, rp)
            A = vh.CsrMatrix(rp, col, val)
            plan.autotune(A, single=False, stream=stream.handle)
            db, dx = vh.DeviceArray.from_numpy(vh.random_uniform(0, n)), vh.DeviceArray.zeros(n, np.float64)
            it, cv = vh.DeviceArray.zeros(4096, np.uint32), vh.DeviceArray.zeros(4096, np.uint32)
            er = vh.DeviceArray.zeros(4096, np.float64)
            plan.solve_into(A, db, dx, it, er, cv, stream.handle, tol=1e-8, max_iters=20000, use_graph=False)
            stream.synchronize()
            out[name] = {"iters_total": int(it.to_numpy().sum())}
            for a in (db, dx, it, cv, er, A):
                a.free()
            plan.close()
        else:
            raise SystemExit(f"unknown profile workload {name}")
    return out


def device_banner(vh):
    name = C.create_string_buffer(64)
    mem, cu, l2, llc = C.c_size_t(), C.c_int(), C.c_int(), C.c_int()
    vh.call("vclhip_device_info", 0, name, 64, C.byref(mem), C.byref(cu), C.byref(l2), C.byref(llc))
    log(f"device {name.value.decode()} CUs={cu.value} HBM={mem.value / 2**30:.0f} GiB L2={l2.value >> 20} MiB")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solvers", action="store_true", help="skip the 128^3 CG/BiCGStab solves")
    ap.add_argument("--mixed-grid", type=int, default=160,
                    help="grid of the FP64-vs-mixed CG comparison (0 skips it)")
    ap.add_argument("--gmres-grid", type=int, default=128, help="grid of the GMRES runs (0 skips them)")
    ap.add_argument("--no-formats", action="store_true", help="skip the COO / ELL / HYB SpMV runs")
    ap.add_argument("--batch", type=int, default=4096, help="systems of the batched CG run (0 skips it)")
    ap.add_argument("--adjoint-grid", type=int, default=512, help="grid of the differentiable solve (0 skips it)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiable-solve measurement (a child process: torch is imported "
                         "before libvclhip so that the process holds one HIP runtime)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only warm CSR/SELL SpMV and the streaming kernel (for counter collection)")
    ap.add_argument("--profile", default=None,
                    help="comma-separated short workloads for rocprofv3 runs, nothing else: spmv (CSR and "
                         "SELL-64 SpMV of the 128^3 Laplacian and the streaming kernels), powerlaw (CSR SpMV of the "
                         "power-law matrix, then its gathers alone), dia_cg (50 FP64 CG iterations on the 160^3 diffusion operator's DIA "
                         "copy), batched (one FP64 batched CG solve of the 4096 systems), and one solve of a "
                         "configuration: c1_cg, c1_bicgstab, c2_cg, c2_bicgstab, c3_cg_fp64, c3_mixed")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sharded (row-partitioned) path is not built; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.adjoint_only:
        import torch  # noqa: F401  (before libvclhip: one HIP runtime in the process)
        import vclhip as vh
        vh.call("vclhip_set_device", 0)
        print(json.dumps(adjoint_bench(vh, args.adjoint_grid)), flush=True)
        return

    import vclhip as vh
    assert vh.device_count() > 0, "no HIP device"
    vh.call("vclhip_set_device", 0)
    device_banner(vh)
    if args.profile:
        print(json.dumps(profile_workloads(vh, args.profile.split(","), args.steps, args.warmup)), flush=True)
        return

    N = args.grid
    t = time.perf_counter()
    rp, col, val = vh.laplace_3d(N)
    n, nnz = rp.size - 1, int(rp[-1])
    log(f"generated 3-D 7-point Laplacian {N}^3: n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    stream = vh.Stream()
    A = vh.CsrMatrix(rp, col, val)
    S = vh.SellMatrix(rp, col, val, 64)
    xh = vh.random_uniform(1, n)
    x = vh.DeviceArray.from_numpy(xh)
    y = vh.DeviceArray.zeros(n, np.float64)
    bytes_csr = csr_bytes(n, nnz)
    bytes_sell = sell_bytes(S.stored, (n + 63) // 64, n)

    if args.kernels_only:
        _, warm, _ = time_launches(vh, lambda: A.spmv(x, y, stream=stream.handle), stream, args.steps, args.warmup)
        _, swarm, _ = time_launches(vh, lambda: S.spmv(x, y, stream=stream.handle), stream, args.steps, args.warmup)
        ref = stream_reference(vh, stream, args.steps)
        print(json.dumps({"csr_spmv_ms": warm, "csr_bytes": bytes_csr, "sell_spmv_ms": swarm,
                          "sell_bytes": bytes_sell, "stream": ref}), flush=True)
        return

    evict = vh.DeviceArray(EVICT_DOUBLES, np.float64)
    vh.call("vclhip_fill_zero", evict.ptr, C.c_size_t(evict.nbytes), None)
    cold, warm, wall = time_launches(vh, lambda: A.spmv(x, y, stream=stream.handle), stream, args.steps,
                                     args.warmup, evict)
    ms = float(cold.mean())
    gbs = bytes_csr / (ms * 1e-3) / 1e9
    warm_gbs = bytes_csr / (warm * 1e-3) / 1e9
    log(f"CSR SpMV cold {ms * 1e3:.1f} us ({gbs:.0f} GB/s), warm {warm * 1e3:.1f} us ({warm_gbs:.0f} GB/s)")
    s_cold, s_warm, _ = time_launches(vh, lambda: S.spmv(x, y, stream=stream.handle), stream, args.steps,
                                      args.warmup, evict)
    sell_gbs = bytes_sell / (s_cold.mean() * 1e-3) / 1e9
    log(f"SELL-64 SpMV cold {s_cold.mean() * 1e3:.1f} us ({sell_gbs:.0f} GB/s), warm {s_warm * 1e3:.1f} us")
    # the FP32 product of the same matrix (values 4 bytes, x and y 4 bytes each)
    A32 = vh.CsrMatrix(rp, col, val, dtype=np.float32)
    x32 = vh.DeviceArray.from_numpy(xh.astype(np.float32))
    y32 = vh.DeviceArray.zeros(n, np.float32)
    f_cold, f_warm, _ = time_launches(vh, lambda: A32.spmv(x32, y32, stream=stream.handle), stream, args.steps,
                                      args.warmup, evict)
    bytes_csr32 = 8 * nnz + 4 * (n + 1) + 8 * n
    csr32 = {"cold_gbs": round(bytes_csr32 / (f_cold.mean() * 1e-3) / 1e9, 2),
             "cold_ms": round(float(f_cold.mean()), 6), "warm_ms": round(f_warm, 6),
             "algorithmic_bytes": bytes_csr32}
    log("CSR SpMV FP32", csr32)
    for a in (A32, x32, y32):
        a.free()
    y_gpu = y.to_numpy()
    ref = stream_reference(vh, stream, 10, evict, bytes_csr, A.num_blocks)
    evict.free()
    log("streaming reference", ref)

    traffic = None
    tfile = ROOT / "profiles" / "csr_spmv_lap3d_128_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 6), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FP64 CSR SpMV y=Ax, 3-D 7-point Laplacian {N}^3 (n={n}, nnz={nnz}), "
                               f"Infinity Cache evicted by a 1 GiB read before each timed SpMV",
                   "n": n, "nnz": nnz, "format": "CSR (uint32 indices)", "algorithmic_bytes_per_spmv": bytes_csr},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "warm_gbs": round(warm_gbs, 2),
        "cold_us": {"mean": round(ms * 1e3, 2), "median": round(float(np.median(cold)) * 1e3, 2),
                    "min": round(float(cold.min()) * 1e3, 2), "max": round(float(cold.max()) * 1e3, 2)},
        "timed_region_wall_s": round(wall, 4),
        "sell64": {"cold_gbs": round(sell_gbs, 2), "cold_ms": round(float(s_cold.mean()), 6),
                   "warm_ms": round(s_warm, 6), "algorithmic_bytes": bytes_sell},
        "csr_fp32": csr32,
        "stream_reference": ref,
    }

    if not args.no_solvers:
        b = vh.DeviceArray.from_numpy(vh.random_uniform(0, n))
        sol = {}
        # a CSR stencil is solved on a DIA copy by default (the same results), here read through its
        # value tables (one value per diagonal); "_dia_uncoded" reads the copy's values as stored,
        # "_as_given" keeps CSR
        sol["cg_fp64_csr"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-10, max_iters=5000)
        sol["cg_fp64_csr_dia_uncoded"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-10, max_iters=5000,
                                                     operator_format=vh.OPERATOR_DIA_UNCODED)
        sol["cg_fp64_csr_as_given"] = timed_solve(vh, "cg", A, b, n, reps=3, tol=1e-10, max_iters=5000,
                                                  operator_format=1)
        sol["cg_fp64_csr_graph"] = timed_solve(vh, "cg", A, b, n, tol=1e-10, max_iters=5000, use_graph=True)
        sol["cg_fp64_sell64"] = timed_solve(vh, "cg", S, b, n, tol=1e-10, max_iters=5000)
        sol["jacobi_cg_fp64_csr"] = timed_solve(vh, "cg", A, b, n, precond=1, tol=1e-10, max_iters=5000)
        sol["bicgstab_fp64_csr"] = timed_solve(vh, "bicgstab", A, b, n, tol=1e-10, max_iters=5000)
        sol["bicgstab_fp64_csr_dia_uncoded"] = timed_solve(vh, "bicgstab", A, b, n, tol=1e-10, max_iters=5000,
                                                           operator_format=vh.OPERATOR_DIA_UNCODED)
        sol["bicgstab_fp64_csr_as_given"] = timed_solve(vh, "bicgstab", A, b, n, tol=1e-10, max_iters=5000,
                                                        operator_format=1)
        sol["jacobi_bicgstab_fp64_csr"] = timed_solve(vh, "bicgstab", A, b, n, precond=1, tol=1e-10,
                                                      max_iters=5000)
        sol["jacobi_cg_fp64_sell64"] = timed_solve(vh, "cg", S, b, n, precond=1, tol=1e-10, max_iters=5000)
        sol["bicgstab_fp64_sell64"] = timed_solve(vh, "bicgstab", S, b, n, tol=1e-10, max_iters=5000)
        sol["jacobi_bicgstab_fp64_sell64"] = timed_solve(vh, "bicgstab", S, b, n, precond=1, tol=1e-10,
                                                         max_iters=5000)
        for k, v in sol.items():
            log(k, v)
        out["solvers_lap3d_128"] = sol
        out["c1_cg_fdm_256"], c1_inputs = c1_solve_bench(vh)
        out["powerlaw_2M"] = powerlaw_bench(vh, stream, args.steps)
    if args.mixed_grid:
        out["mixed_vs_fp64"] = mixed_vs_fp64(vh, args.mixed_grid)
    if not args.no_formats:
        out["formats_spmv"] = formats_bench(vh, stream, args.steps)
    if args.gmres_grid:
        out["gmres_convdiff"] = gmres_bench(vh, args.gmres_grid)
    if args.batch:
        out["batched_cg"] = batched_cg_bench(vh, args.batch)
    if args.adjoint_grid:
        log("differentiable solve (child process) ...")
        p = subprocess.run([sys.executable, str(pathlib.Path(__file__).resolve()), "--adjoint-only",
                            "--adjoint-grid", str(args.adjoint_grid)], capture_output=True, text=True, timeout=900)
        sys.stderr.write(p.stderr[-2000:])
        if p.returncode != 0:
            raise RuntimeError(f"adjoint measurement failed with status {p.returncode}")
        out["adjoint"] = json.loads(p.stdout.strip().splitlines()[-1])

    if not args.no_cpu:
        log("CPU baseline (reference host backend) ...")
        cb = cpu_baseline(rp, col, val, xh, n, nnz, args.cpu_seconds, None if args.no_solvers else c1_inputs)
        if cb is not None:
            base, y_cpu = cb
            if not args.no_solvers:
                import pyoracle as po
                base.update(cpu_solver_baseline(vh, po, po.ref(), args.mixed_grid))
            out["cpu_baseline"] = base
            out["cpu_vs_gpu_relinf"] = float(np.max(np.abs(y_cpu - y_gpu)) / np.max(np.abs(y_cpu)))
            log("CPU baseline", base)
        else:
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
