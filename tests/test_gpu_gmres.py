"""GPU parity of pipelined GMRES (gmres.hpp:182-370) and its operations
(iterative_operations.hpp:247-418).

The HIP solver computes the Gram-Schmidt inner products over all n entries, as the reference's
device kernels do; the oracle runs the same way (gs_skip_tail=False).  Its reference-range variant
is pinned bit-exact to the reference build by tests/test_oracle_golden.py, and the full-range
variant to the Krylov minimizer by tests/test_oracle_math.py.  Bars: iteration counts within 2% of
the oracle, the FP64 true residual ||b - A x|| / ||b|| below the tolerance when the solve reports
convergence, x close to the oracle's, and the operations bit-exact where the reference's
arithmetic is sequential (chunk sums, element-wise updates) and within 1e-13 where only the
summation order of an inner product changes.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def iters_close(got, want):
    return abs(int(got) - int(want)) <= max(2, int(np.ceil(0.02 * want)))


def true_relres(rp, col, val, x, b):
    import pyoracle
    r = b - pyoracle.csr_spmv(rp, col, val, np.asarray(x, np.float64))
    return np.linalg.norm(r) / np.linalg.norm(b)


def run_gpu(vh, rp, col, val, b, fmt="csr", dtype=np.float64, **kw):
    A = vh.CsrMatrix(rp, col, val, dtype=dtype) if fmt == "csr" else vh.SellMatrix(rp, col, val, 64, dtype=dtype)
    db = vh.DeviceArray.from_numpy(np.asarray(b, dtype))
    dx = vh.DeviceArray.zeros(b.size, dtype)
    res = vh.gmres(A, db, dx, **kw)
    return dx.to_numpy(), res


def check_against_oracle(oracle, vh, rp, col, val, b, tol, mi, kd, fmt="csr"):
    A = oracle.CsrHost(rp, col, val)
    xo, io, eo, _ = oracle.gmres(A, b, tol, mi, kd)
    x, res = run_gpu(vh, rp, col, val, b, fmt, tol=tol, max_iters=mi, krylov_dim=kd)
    print(f"[gmres] n={rp.size - 1} fmt={fmt} tol={tol} max_iters={mi} K={kd}: gpu {res.iters} it err {res.error:.3e}"
          f" | oracle {io} it err {eo:.3e}", flush=True)
    if max(res.error, eo) < 1e-12:
        # Both runs drove the last cycle down to round-off, where the reference's round-off check
        # (|xi_i| >= rho, gmres.hpp:316-320) truncates the cycle at a step set by rounding noise:
        # the same number of cycles is the bar there.
        assert -(-int(res.iters) // kd) == -(-int(io) // kd), (res.iters, io)
    else:
        assert iters_close(res.iters, io), (res.iters, io)
    rel = true_relres(rp, col, val, x, b)
    orel = true_relres(rp, col, val, xo, b)
    if res.converged:
        assert rel < tol * (1 + 1e-9), rel
    # the same stopping state: the estimate and the solution agree with the oracle's
    assert abs(res.error - eo) <= 1e-3 * eo + 1e-12
    assert np.linalg.norm(x - xo) <= 1e-5 * np.linalg.norm(xo) + 10 * max(rel, orel) * np.linalg.norm(xo)
    return res, io


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_gmres_fixture_cases(dev, vh, oracle, fmt):
    g = load("solve_gmres_convdiff_20x15x3")
    for tol, mi, kd in g["cases"]:
        check_against_oracle(oracle, vh, g["row_ptr"], g["col"], g["val"], g["b"], float(tol), int(mi), int(kd), fmt)


@pytest.mark.parametrize("kd", [1, 16, 17, 33, 40])
def test_gmres_krylov_dims(dev, vh, oracle, kd):
    """krylov_dim across the 16-vector Gram-Schmidt tiles (k = 16, 17, 32, 33, 39)."""
    rp, col, val = vh.convection_diffusion(12, 10, 5, beta=(0.8, 0.3, 0.1))
    b = vh.random_uniform(9, rp.size - 1)
    check_against_oracle(oracle, vh, rp, col, val, b, 1e-9, 800, kd)


def test_gmres_laplacian_fixture(dev, vh, oracle):
    g = load("solve_gmres_lap2d_40x40")
    for tol, mi, kd in g["cases"][:3]:
        check_against_oracle(oracle, vh, g["row_ptr"], g["col"], g["val"], g["b"], float(tol), int(mi), int(kd))


def test_gmres_fp32(dev, vh, oracle):
    """FP32 basis and vectors, inner products accumulated in FP64."""
    rp, col, val = vh.convection_diffusion(16, 12, 6)
    b = vh.random_uniform(2, rp.size - 1)
    x, res = run_gpu(vh, rp, col, val, b, dtype=np.float32, tol=1e-5, max_iters=600, krylov_dim=20)
    xo, io, eo, _ = oracle.gmres(oracle.CsrHost(rp, col, val), b, 1e-5, 600, 20)
    print(f"[gmres fp32] {res.iters} it err {res.error:.3e} | oracle fp64 {io} it", flush=True)
    assert res.converged and res.iters <= 1.25 * io + 20
    assert true_relres(rp, col, val, x, b) < 2e-5


def test_gmres_edge_cases(dev, vh, oracle):
    rp, col, val = vh.convection_diffusion(7, 6, 2)
    n = rp.size - 1
    # b = 0: x = 0 at once (gmres.hpp:224-225)
    x, res = run_gpu(vh, rp, col, val, np.zeros(n), tol=1e-8, max_iters=100, krylov_dim=10)
    assert res.iters == 0 and res.converged and not x.any()
    b = vh.random_uniform(4, n)
    # max_iters below krylov_dim and max_iters = 0: one full cycle (max_restarts = 0)
    for mi in (5, 0):
        x, res = run_gpu(vh, rp, col, val, b, tol=1e-12, max_iters=mi, krylov_dim=10)
        xo, io, eo, _ = oracle.gmres(oracle.CsrHost(rp, col, val), b, 1e-12, mi, 10)
        assert res.iters == io == 10 and res.restarts == 1
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    # krylov_dim > n: the basis becomes dependent and the premature-convergence check cuts it
    rp2, col2, val2 = vh.convection_diffusion(3, 2, 2)
    b2 = vh.random_uniform(5, rp2.size - 1)
    check_against_oracle(oracle, vh, rp2, col2, val2, b2, 1e-10, 60, 20)
    # n = 1: r/rho_0 and v_0 are +-1, so xi_0 = +-rho and the estimator keeps no step
    # (gmres.hpp:316-320); the reference leaves x = 0, and so does the HIP solver
    one = (np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([4.0]))
    x, res = run_gpu(vh, *one, np.array([2.0]), tol=1e-10, max_iters=10, krylov_dim=5)
    xo, io, eo, _ = oracle.gmres(oracle.CsrHost(*one), np.array([2.0]), 1e-10, 10, 5)
    assert (res.iters, res.error, x[0]) == (io, eo, xo[0])
    # invalid krylov_dim
    with pytest.raises(vh.VclHipError):
        run_gpu(vh, rp, col, val, b, krylov_dim=0)
    with pytest.raises(vh.VclHipError):
        run_gpu(vh, rp, col, val, b, krylov_dim=2000)


def test_gmres_history_monitor_graph(dev, vh, oracle):
    rp, col, val = vh.convection_diffusion(14, 11, 3)
    b = vh.random_uniform(7, rp.size - 1)
    xo, io, eo, ho = oracle.gmres(oracle.CsrHost(rp, col, val), b, 1e-9, 400, 15, history=True)
    hist = np.zeros(400)
    x0, res0 = run_gpu(vh, rp, col, val, b, tol=1e-9, max_iters=400, krylov_dim=15, history=hist)
    assert iters_close(res0.iters, io)
    m = min(io, res0.iters, 3 * 15)
    assert np.allclose(hist[:m], ho[:m], rtol=1e-6, atol=0)
    # graph replay of the cycle: the same result bit for bit
    x1, res1 = run_gpu(vh, rp, col, val, b, tol=1e-9, max_iters=400, krylov_dim=15, use_graph=True)
    assert np.array_equal(x0, x1) and (res0.iters, res0.error) == (res1.iters, res1.error)
    # monitor: once per cycle with the cycle's estimate (gmres.hpp:366-367); nonzero stops
    seen = []

    @vh.MONITOR
    def mon(xp, err, user):
        seen.append(err)
        return 0

    x2, res2 = run_gpu(vh, rp, col, val, b, tol=1e-9, max_iters=400, krylov_dim=15, monitor=mon)
    assert np.array_equal(x0, x2) and len(seen) == res2.restarts and seen[-1] == res2.error
    stop_after = []

    @vh.MONITOR
    def stop(xp, err, user):
        stop_after.append(err)
        return 1

    x3, res3 = run_gpu(vh, rp, col, val, b, tol=1e-9, max_iters=400, krylov_dim=15, monitor=stop)
    assert len(stop_after) == 1 and res3.iters == 15 and res3.restarts == 1


def test_gmres_operations(dev, vh, oracle):
    """The five pipelined_gmres_* operations against the oracle (all entries in stage 1)."""
    g = load("gmres_ops_1001")
    n, ld, K, k, chunk = (int(g[q]) for q in ("n", "ld", "K", "k", "chunk"))
    basis = np.ascontiguousarray(g["basis"])
    residual = np.ascontiguousarray(g["residual"])
    # stage 1
    want = np.zeros(K * chunk)
    oracle.gmres_stage1(basis, n, ld, k, want, chunk)
    dB = vh.DeviceArray.from_numpy(basis)
    out = vh.DeviceArray.from_numpy(np.full(K * chunk, 7.0))
    vh.call("vclhip_pipelined_gmres_gram_schmidt_stage1_f64", dB.ptr, n, ld, k, out.ptr, chunk, None)
    got = out.to_numpy()
    for j in range(k):
        assert abs(got[j * chunk] - want[j * chunk]) <= 1e-13 * n
        assert not got[j * chunk + 1:(j + 1) * chunk].any()
    assert np.all(got[k * chunk:] == 7.0)
    # stage 2
    vi = np.ascontiguousarray(g["vi_in_vk"])
    B2, R2, buf2 = basis.copy(), np.zeros(K * K), np.zeros(3 * chunk)
    oracle.gmres_stage2(B2, n, ld, k, vi, R2, K, buf2, chunk)
    dB = vh.DeviceArray.from_numpy(basis)
    dR = vh.DeviceArray.zeros(K * K, np.float64)
    dbuf = vh.DeviceArray.zeros(3 * chunk, np.float64)
    dvi = vh.DeviceArray.from_numpy(vi)
    vh.call("vclhip_pipelined_gmres_gram_schmidt_stage2_f64", dB.ptr, n, ld, k, dvi.ptr, dR.ptr, K, dbuf.ptr, chunk,
            None)
    assert np.array_equal(dB.to_numpy(), B2) and np.array_equal(dR.to_numpy(), R2)
    hb = dbuf.to_numpy()
    assert abs(hb[chunk] - buf2[chunk]) <= 1e-13 * buf2[chunk] and not hb[chunk + 1:2 * chunk].any()
    # normalize_vk
    vk = basis[k * ld:k * ld + n].copy()
    R3, rdv = np.zeros(K * K), np.zeros(K * chunk)
    nbuf = np.ascontiguousarray(g["normalize_buf"])
    oracle.gmres_normalize_vk(vk, residual, R3, k * K + k, nbuf, rdv, chunk, k * chunk)
    dv = vh.DeviceArray.from_numpy(basis[k * ld:k * ld + n].copy())
    dR = vh.DeviceArray.zeros(K * K, np.float64)
    drdv = vh.DeviceArray.zeros(K * chunk, np.float64)
    dres = vh.DeviceArray.from_numpy(residual)
    dnbuf = vh.DeviceArray.from_numpy(nbuf)
    vh.call("vclhip_pipelined_gmres_normalize_vk_f64", n, dv.ptr, dres.ptr, dR.ptr, k * K + k, dnbuf.ptr, drdv.ptr,
            chunk, k * chunk, None)
    assert np.array_equal(dv.to_numpy(), vk) and np.array_equal(dR.to_numpy(), R3)
    hr = drdv.to_numpy()
    assert abs(hr[k * chunk] - rdv[k * chunk]) <= 1e-13 * n and not hr[k * chunk + 1:(k + 1) * chunk].any()
    # update_result, also with k = 0 (c[0] applies) and k = 1
    coef = np.ascontiguousarray(g["update_coef"])
    for kk in (K - 1, 1, 0):
        x = np.ascontiguousarray(g["update_x0"]).copy()
        oracle.gmres_update_result(x, residual, basis, n, ld, coef, kk)
        dx = vh.DeviceArray.from_numpy(g["update_x0"])
        dB = vh.DeviceArray.from_numpy(basis)
        dc = vh.DeviceArray.from_numpy(coef)
        vh.call("vclhip_pipelined_gmres_update_result_f64", dx.ptr, dres.ptr, dB.ptr, n, ld, dc.ptr, kk, None)
        assert np.array_equal(dx.to_numpy(), x), kk
    # pipelined_gmres_prod is pipelined_cg_prod
    rp, col, val = vh.convection_diffusion(11, 7, 13)
    pv = vh.random_uniform(3, rp.size - 1)
    A = vh.CsrMatrix(rp, col, val)
    outs = []
    dp = vh.DeviceArray.from_numpy(pv)
    for fn in ("vclhip_pipelined_gmres_prod_csr_f64", "vclhip_pipelined_cg_prod_csr_f64"):
        Ap = vh.DeviceArray.zeros(rp.size - 1, np.float64)
        buf = vh.DeviceArray.zeros(384, np.float64)
        vh.call(fn, C.byref(A.desc), dp.ptr, Ap.ptr, buf.ptr, 128, None)
        outs.append((Ap.to_numpy(), buf.to_numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0], oracle.csr_spmv(rp, col, val, pv))


def test_gmres_convection_diffusion_64(dev, vh, oracle):
    """A 64^3 convection-diffusion problem (n = 262144) against the oracle."""
    rp, col, val = vh.convection_diffusion(64, 64, 64, beta=(0.6, 0.4, 0.2))
    b = vh.random_uniform(1, rp.size - 1)
    for fmt in ("csr", "sell"):
        check_against_oracle(oracle, vh, rp, col, val, b, 1e-8, 600, 20, fmt)


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_gmres_absolute_tolerance_vs_oracle(dev, vh, oracle, fmt):
    """abs_tolerance(): a restart cycle begins with rho_0 = ||b - A x|| and the solve ends there when
    rho_0 < abs_tol (gmres.hpp:234), or at once when rho_0 <= abs_tol (:227-228).  With a relative
    tolerance out of reach the absolute one ends the solve at the oracle's step and cycle counts."""
    rp, col, val = vh.convection_diffusion(24, 20, 12)
    n = rp.size - 1
    b = vh.random_uniform(21, n)
    nb = float(np.linalg.norm(b))
    H = oracle.CsrHost(rp, col, val)
    for kd, rel in ((20, 1e-6), (7, 1e-4)):
        xo, io, eo, _ = oracle.gmres(H, b, 1e-15, 3000, kd, abs_tol=rel * nb)
        x, res = run_gpu(vh, rp, col, val, b, fmt, tol=1e-15, abs_tol=rel * nb, max_iters=3000, krylov_dim=kd)
        print(f"[gmres abs] K={kd} abs={rel:g}*||b||: gpu {res.iters} it {res.restarts} cycles | oracle {io} it",
              flush=True)
        assert res.converged and iters_close(res.iters, io), (res.iters, io)
        assert true_relres(rp, col, val, x, b) < rel
        assert np.linalg.norm(x - xo) <= 10 * rel * np.linalg.norm(xo)
    x, res = run_gpu(vh, rp, col, val, 1e-3 * b, fmt, tol=1e-15, abs_tol=1e-2 * nb, max_iters=3000, krylov_dim=20)
    assert res.iters == 0 and not x.any()
