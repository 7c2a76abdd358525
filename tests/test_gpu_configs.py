"""The benchmark workloads at their full sizes, through the C ABI: C1 CG on
generate_fdm_laplace(256, 256); C2 the 128^3 Laplacian (CG, Jacobi-PCG, BiCGStab, Jacobi-BiCGStab,
CSR and SELL-64) and the 2 M power-law matrix; C3 FP64 and mixed CG on the 160^3 diffusion
operator; C4 the batched CG of 4096 systems; C5 the 512^2 differentiable solve.

C1 is checked against the reference itself (oracle/_ref, the reference's host backend built from
its sources; the restated oracle where that build is absent).  At the larger sizes the checks are
the ones that do not depend on size: the FP64 true residual ||b - A x|| / ||b||, iteration counts
against the oracle's (same recurrences), the mixed solve's distance to the FP64 solution, and
every system of the batch.
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def iters_close(got, want):
    return abs(int(got) - int(want)) <= max(2, int(np.ceil(0.02 * want)))


def device_relres(vh, A, b, x):
    """||b - A x|| / ||b|| in FP64 with the device SpMV (for the large configurations)."""
    import ctypes as C
    n = b.size
    r = vh.DeviceArray.from_numpy(b)
    A.spmv(vh.DeviceArray.from_numpy(x), r, -1.0, 1.0)
    nr = C.c_double()
    vh.call("vclhip_norm2_f64", n, r.ptr, 0, 1, C.byref(nr), None)
    return nr.value / np.linalg.norm(b)


def test_c1_cg_fdm_256_vs_reference(dev, vh, oracle):
    """generate_fdm_laplace(256, 256), b ~ U[-1,1] seed 0, cg_tag(1e-8, 2000)."""
    rp, col, val = vh.fdm_laplace_2d(256, 256)
    b = vh.random_uniform(0, rp.size - 1)
    A = vh.CsrMatrix(rp, col, val)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(b.size, np.float64)
    res = vh.cg(A, db, dx, tol=1e-8, max_iters=2000)
    if oracle.ref_available():
        oracle.set_threads(1)
        xr, it_r, err_r = oracle.ref_cg(rp, col, val, b, 1e-8, 2000)
    else:
        xr, it_r, err_r, _ = oracle.cg_pipelined(oracle.CsrHost(rp, col, val), b, 1e-8, 2000)
    assert res.converged and iters_close(res.iters, it_r), (res.iters, it_r)
    assert res.error < 1e-8 and abs(res.error - err_r) < 0.05 * err_r
    x = dx.to_numpy()
    r = b - oracle.csr_spmv(rp, col, val, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-6


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_c2_cg_lap3d_128(dev, vh, oracle, fmt):
    """3-D Laplacian 128^3, tol 1e-10: iterations against the oracle's pipelined CG (8 threads)."""
    rp, col, val = vh.laplace_3d(128)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val) if fmt == "csr" else vh.SellMatrix(rp, col, val, 64)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, np.float64)
    res = vh.cg(A, db, dx, tol=1e-10, max_iters=5000)
    oracle.set_threads(8)
    try:
        _, it_o, err_o, _ = oracle.cg_pipelined(oracle.CsrHost(rp, col, val), b, 1e-10, 5000)
    finally:
        oracle.set_threads(1)
    assert res.converged and iters_close(res.iters, it_o), (res.iters, it_o)
    Acsr = vh.CsrMatrix(rp, col, val)
    assert device_relres(vh, Acsr, b, dx.to_numpy()) < 1e-10


_C2_ORACLE = {}


def c2_oracle(vh, oracle, solver):
    """The oracle's solve of the 128^3 Laplacian (8 threads, its OpenMP reduction order) for one of
    the solvers of config C2, computed once per session: (iterations, error, x)."""
    if solver not in _C2_ORACLE:
        rp, col, val = vh.laplace_3d(128)
        b = vh.random_uniform(0, rp.size - 1)
        Ah = oracle.CsrHost(rp, col, val)
        d = np.full(rp.size - 1, 6.0)  # the Laplacian's diagonal: the Jacobi preconditioner
        oracle.set_threads(8)
        try:
            if solver == "jacobi_cg":  # cg.hpp:257-322 with jacobi_precond
                x, it, err, _ = oracle.cg_classic(Ah, b, 1e-10, 5000, diag=d)
            elif solver == "bicgstab":  # bicgstab.hpp:97-215
                x, it, err, _ = oracle.bicgstab_pipelined(Ah, b, 1e-10, 5000)
            else:  # bicgstab.hpp:398-489 with jacobi_precond
                x, it, err, _ = oracle.bicgstab_classic(Ah, b, 1e-10, 5000, diag=d)
        finally:
            oracle.set_threads(1)
        _C2_ORACLE[solver] = (it, err, x)
    return _C2_ORACLE[solver]


def c2_bicgstab_envelope(solver):
    """The range of BiCGStab iteration counts on the 128^3 Laplacian over the reference's thread
    counts (1-16) and the oracle's reduction orders."""
    g = np.load(GOLDEN / "c2_bicgstab_lap3d_128_iters.npz")
    counts = np.concatenate([g["ref_" + solver], g["oracle_" + solver]])
    return int(counts.min()), int(counts.max())


@pytest.mark.parametrize("fmt", ["csr", "sell"])
@pytest.mark.parametrize("solver", ["jacobi_cg", "bicgstab", "jacobi_bicgstab"])
def test_c2_solvers_lap3d_128(dev, vh, oracle, solver, fmt):
    """The other solvers of config C2 on the 128^3 Laplacian (n = 2,097,152), tol 1e-10, in CSR
    (solved on its DIA copy, the default) and SELL-64: the iteration count within 2% of the
    oracle's (8 threads) for Jacobi-PCG, the reported error below the tolerance, x close to the
    oracle's, and the FP64 true residual.  BiCGStab's count at this size is set by the rounding of
    its inner products: the reference's own solve() took 365-394 iterations (306-345 with Jacobi)
    over 1-16 OpenMP threads, and not the same count twice at 4 threads (its reductions combine
    the threads' sums in arrival order); the oracle takes 348-398 (313-348) over its reduction
    orders.  These counts are the fixture c2_bicgstab_lap3d_128_iters (tests/golden/make_golden.py,
    from the reference build), and the BiCGStab counts are held to within 2% of their range.
    The criteria differ by solver: Jacobi-PCG stops on the M^-1-norm of the
    residual (cg.hpp:306-309), which with the constant diagonal 6 of this operator equals the
    2-norm ratio, so ||b - A x|| / ||b|| < tol; Jacobi-BiCGStab stops on ||M^-1 r|| / ||b|| =
    ||r|| / (6 ||b||) (bicgstab.hpp:460-461), so its 2-norm ratio is below 6 tol; plain BiCGStab
    stops on its residual estimate (bicgstab.hpp:195), so below tol up to the estimate's drift."""
    rp, col, val = vh.laplace_3d(128)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val) if fmt == "csr" else vh.SellMatrix(rp, col, val, 64)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, np.float64)
    jac = 1 if solver.startswith("jacobi") else 0
    if solver == "jacobi_cg":
        res = vh.cg(A, db, dx, precond=jac, tol=1e-10, max_iters=5000)
    else:
        res = vh.bicgstab(A, db, dx, precond=jac, tol=1e-10, max_iters=5000)
    it_o, err_o, x_o = c2_oracle(vh, oracle, solver)
    print(f"{solver} {fmt}: {res.iters} iterations (oracle {it_o}), error {res.error:.3e} (oracle {err_o:.3e})")
    assert res.converged and res.error < 1e-10
    if solver == "jacobi_cg":
        assert iters_close(res.iters, it_o), (res.iters, it_o)
    else:
        lo, hi = c2_bicgstab_envelope(solver)
        assert int(np.floor(0.98 * lo)) <= res.iters <= int(np.ceil(1.02 * hi)), (res.iters, lo, hi)
    if fmt == "csr":  # the 7-point stencil's diagonals hold one value each: the coded DIA copy
        assert res.operator_used == vh.USED_DIA_CODED
    x = dx.to_numpy()
    bound = {"jacobi_cg": 1.0, "bicgstab": 1.05, "jacobi_bicgstab": 6.0}[solver] * 1e-10
    rel = device_relres(vh, vh.CsrMatrix(rp, col, val), b, x)
    print(f"  true residual {rel:.3e} (bound {bound:.2e})")
    assert rel < bound
    assert np.linalg.norm(x - x_o) / np.linalg.norm(x_o) < 1e-6


def test_c2_powerlaw_bicgstab_full_size(dev, vh, oracle):
    """Power-law nonsymmetric matrix, n = 2,097,152 (row lengths P(k) ~ k^-2 on 1..512), Jacobi
    BiCGStab to 1e-10: FP64 true residual at the tolerance (the preconditioned criterion is
    ||M^-1 r|| / ||b||, so the plain 2-norm is allowed a factor of 10).  Without a preconditioner
    BiCGStab stagnates near 1e-2 on this badly scaled matrix, as the restated reference does on a
    smaller one of the same family.  The SpMV itself matches the oracle on this matrix."""
    n = 2097152
    rp, col, val = vh.powerlaw(n, 512, 2.0, 1.1, 1)
    b = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, np.float64)
    res = vh.bicgstab(A, db, dx, precond=1, tol=1e-10, max_iters=2000)
    assert res.converged
    assert device_relres(vh, A, b, dx.to_numpy()) < 1e-9
    y = vh.DeviceArray.zeros(n, np.float64)
    A.spmv(db, y)
    ref = oracle.csr_spmv(rp, col, val, b)
    assert np.max(np.abs(y.to_numpy() - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_c3_mixed_vs_fp64_160(dev, vh):
    """160^3 variable-coefficient diffusion (k = exp(N(0,1)), harmonic faces), tol 1e-10: the mixed
    solve reaches the FP64-verified tolerance, close to the FP64 CG solution, with at most 12% more
    iterations."""
    g = 160
    k = np.exp(vh.random_normal(1, g ** 3))
    rp, col, val = vh.diffusion(g, g, g, k)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val)
    db = vh.DeviceArray.from_numpy(b)
    x64 = vh.DeviceArray.zeros(n, np.float64)
    r64 = vh.cg(A, db, x64, tol=1e-10, max_iters=20000)
    xm = vh.DeviceArray.zeros(n, np.float64)
    rm = vh.mixed_cg(A, db, xm, tol=1e-10, max_iters=20000, mixed_mode=1)
    assert r64.converged and rm.converged
    assert device_relres(vh, A, b, xm.to_numpy()) < 1e-10
    assert rm.iters <= 1.12 * r64.iters
    a, c = xm.to_numpy(), x64.to_numpy()
    assert np.linalg.norm(a - c) / np.linalg.norm(c) < 1e-6


@pytest.mark.parametrize("single", [False, True], ids=["fp64", "fp32"])
def test_c4_batched_4096(dev, vh, single):
    """4096 graph-Laplacian systems, n_i ~ U{64..4096}, in FP64 at tol 1e-8 and in FP32 at tol 1e-5
    (an FP32 solve cannot reach 1e-8: its recursive residual drifts from the true one by about the
    FP32 resolution times the condition number).  Every system converges on its recursive residual,
    and its FP64 true residual ||b_i - A_i x_i|| / ||b_i|| (x_i and b_i as the solve holds them,
    A_i in FP64) is checked: FP64 at the tolerance, FP32 within TRUE_FP32 of it."""
    sys_ptr, rp, col, val = vh.graph_laplacian_batch(4096, 64, 4096, 8.0, 0.1, 1.0, 0.01, 1)
    n = rp.size - 1
    dtype = np.float32 if single else np.float64
    tol = 1e-5 if single else 1e-8
    b = vh.random_uniform(0, n).astype(dtype)
    A = vh.CsrMatrix(rp, col, val, dtype=dtype)
    plan = vh.BatchPlan(sys_ptr, rp)
    plan.autotune(A, single=single)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, dtype)
    it, err, conv = plan.solve(A, db, dx, tol=tol, max_iters=5000)
    assert (conv == 1).all() and (err < tol).all()
    A64 = vh.CsrMatrix(rp, col, val)
    y = vh.DeviceArray.zeros(n, np.float64)
    A64.spmv(vh.DeviceArray.from_numpy(dx.to_numpy().astype(np.float64)), y)
    b64 = b.astype(np.float64)
    r = b64 - y.to_numpy()
    starts = sys_ptr[:-1].astype(np.int64)
    rel = np.sqrt(np.add.reduceat(r * r, starts) / np.add.reduceat(b64 * b64, starts))
    print(f"{'fp32' if single else 'fp64'}: true residual max {rel.max():.3e} median {np.median(rel):.3e}")
    assert rel.max() < (TRUE_FP32 if single else 1.05e-8), rel.max()


# FP64 true residual that the FP32 batched solve at tol 1e-5 stays below on every system
TRUE_FP32 = 5e-5
