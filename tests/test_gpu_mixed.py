"""GPU parity of the mixed-precision CG (vclhip_mixed_cg_solve_*).

mixed_mode 0 follows the reference's schedule (mixed_precision_cg.hpp:95-187) with FP32 inner
products accumulated in FP64; its iteration count is checked against the reference build's
fixture (2% bar) and against the oracle restatement with the same accumulation.  mixed_mode 1
(reliable updates, no reference counterpart) is checked against the oracle restatement of the same
algorithm.  Every converged result is verified by the FP64 true residual ||b - A x|| / ||b||
computed on the host, and its distance to the FP64 CG solution is bounded.
"""
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
    r = b - pyoracle.csr_spmv(rp, col, val, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def run_mixed(vh, rp, col, val, b, fmt="csr", precond=0, values32=None, **kw):
    A = vh.CsrMatrix(rp, col, val) if fmt == "csr" else vh.SellMatrix(rp, col, val, 64)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(b.size, np.float64)
    res = vh.mixed_cg(A, db, dx, precond=precond, values32=values32, **kw)
    return dx.to_numpy(), res


def test_reference_schedule_matches_reference_fixture(dev, vh, oracle):
    g = load("solve_lap2d_40x40")
    rp, col, val, b = g["row_ptr"], g["col"], g["val"], g["b"]
    x, res = run_mixed(vh, rp, col, val, b, tol=1e-8, max_iters=2000, mixed_mode=0)
    _, it_o, err_o, upd_o, _ = oracle.mixed_cg2(oracle.CsrHost(rp, col, val), b, 1e-8, 2000, mode=0, acc64=True)
    assert iters_close(res.iters, g["mixed_iters"]), (res.iters, g["mixed_iters"])
    assert iters_close(res.iters, it_o) and abs(res.restarts - upd_o) <= 1
    assert res.converged and res.error < 1e-8
    assert abs(res.error - err_o) <= 1e-3 * err_o
    assert true_relres(rp, col, val, x, b) < 1e-8
    assert np.linalg.norm(x - g["mixed_x"]) / np.linalg.norm(g["mixed_x"]) < 1e-6


@pytest.mark.parametrize("fmt", ["csr", "sell"])
@pytest.mark.parametrize("jacobi", [False, True])
def test_reliable_update_vs_oracle(dev, vh, oracle, fmt, jacobi):
    k = np.exp(vh.random_normal(1, 24 * 20 * 18))
    rp, col, val = vh.diffusion(24, 20, 18, k)
    b = vh.random_uniform(0, rp.size - 1)
    Ah = oracle.CsrHost(rp, col, val)
    diag = oracle.row_info(rp, col, val, 3) if jacobi else None
    x, res = run_mixed(vh, rp, col, val, b, fmt=fmt, precond=1 if jacobi else 0, tol=1e-10, max_iters=3000,
                       mixed_mode=1)
    xo, it_o, err_o, upd_o, _ = oracle.mixed_cg2(Ah, b, 1e-10, 3000, mode=1, diag=diag)
    assert res.converged and res.error < 1e-10
    assert iters_close(res.iters, it_o), (res.iters, it_o)
    assert abs(res.restarts - upd_o) <= 1
    rr = true_relres(rp, col, val, x, b)
    assert rr < 1e-10
    # error against the FP64 CG solution of the same system
    x64, it64, _, _ = oracle.cg_pipelined(Ah, b, 1e-12, 5000)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-7
    # FP32 iterations cost at most a few percent more than FP64 CG to the same tolerance
    _, it64_10, _, _ = oracle.cg_pipelined(Ah, b, 1e-10, 5000)
    if not jacobi:
        assert res.iters <= 1.15 * it64_10


def test_values32_argument_and_graph_are_identical(dev, vh):
    rp, col, val = vh.laplace_3d(30, 26, 22)
    b = vh.random_uniform(0, rp.size - 1)
    # the caller's FP32 values run the CSR inner operator, as inner_format = CSR does
    x0, r0 = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=2000, mixed_mode=1, inner_format=1)
    v32 = vh.DeviceArray.from_numpy(val.astype(np.float32))
    x1, r1 = run_mixed(vh, rp, col, val, b, values32=v32, tol=1e-10, max_iters=2000, mixed_mode=1)
    assert r0.iters == r1.iters and r0.restarts == r1.restarts and np.array_equal(x0, x1)
    # default (DIA inner copy for this stencil), eager and graph-captured
    x2, r2 = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=2000, mixed_mode=1)
    x3, r3 = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=2000, mixed_mode=1, use_graph=True, chunk=8)
    assert r2.iters == r3.iters and r2.restarts == r3.restarts and np.array_equal(x2, x3)
    assert iters_close(r2.iters, r0.iters)
    x4, r4 = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=2000, mixed_mode=0, use_graph=True)
    x5, r5 = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=2000, mixed_mode=0)
    assert r4.iters == r5.iters and np.array_equal(x4, x5)


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_inner_formats_vs_oracle(dev, vh, oracle, fmt):
    """CSR, SELL-64 and SELL-64 with 16-bit column offsets as the inner copy all follow the oracle."""
    k = np.exp(vh.random_normal(5, 20 * 22 * 24))
    rp, col, val = vh.diffusion(20, 22, 24, k)
    b = vh.random_uniform(0, rp.size - 1)
    x, res = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=3000, mixed_mode=1, inner_format=fmt)
    _, it_o, _, upd_o, _ = oracle.mixed_cg2(oracle.CsrHost(rp, col, val), b, 1e-10, 3000, mode=1)
    assert res.converged and iters_close(res.iters, it_o) and abs(res.restarts - upd_o) <= 1
    assert true_relres(rp, col, val, x, b) < 1e-10
    # a random-graph Laplacian (row lengths vary) pads too much for the automatic SELL choice; both
    # formats, forced, solve it
    _, rp2, col2, val2 = vh.graph_laplacian_batch(1, 3000, 3000, 8.0, 0.1, 1.0, 0.01, 3)
    b2 = vh.random_uniform(1, 3000)
    x2, r2 = run_mixed(vh, rp2, col2, val2, b2, tol=1e-10, max_iters=3000, mixed_mode=1, inner_format=fmt)
    assert r2.converged and true_relres(rp2, col2, val2, x2, b2) < 1e-10


@pytest.mark.parametrize("jacobi", [False, True])
def test_dia_inner_copy_equals_sell(dev, vh, jacobi):
    """The DIA inner copy sums each row in one lane in column order, skipping zeros, as the SELL-64
    copy does: both schedules give the same iterations and the same x, bit for bit; the automatic
    choice takes DIA for a 7-point operator."""
    k = np.exp(vh.random_normal(6, 20 * 22 * 24))
    rp, col, val = vh.diffusion(20, 22, 24, k)
    b = vh.random_uniform(0, rp.size - 1)
    pc = 1 if jacobi else 0
    for mode in ((1,) if jacobi else (0, 1)):
        xs, rs = run_mixed(vh, rp, col, val, b, precond=pc, tol=1e-10, max_iters=3000, mixed_mode=mode,
                           inner_format=3)
        xd, rd = run_mixed(vh, rp, col, val, b, precond=pc, tol=1e-10, max_iters=3000, mixed_mode=mode,
                           inner_format=4)
        xa, ra = run_mixed(vh, rp, col, val, b, precond=pc, tol=1e-10, max_iters=3000, mixed_mode=mode)
        for x, r in ((xd, rd), (xa, ra)):
            assert r.iters == rs.iters and r.restarts == rs.restarts and r.error == rs.error, mode
            assert np.array_equal(x, xs), mode
        assert rd.converged and true_relres(rp, col, val, xd, b) < 1e-10


def test_coded_dia_inner_copy_equals_sell(dev, vh):
    """The 72^3 Laplacian (373,248 rows, one value per diagonal): the DIA inner copy is read through
    its value tables (reported as VCLHIP_INNER_DIA_CODED) and gives the SELL-64 inner copy's
    iterations, corrections and x bit for bit."""
    rp, col, val = vh.laplace_3d(72)
    b = vh.random_uniform(31, rp.size - 1)
    xs, rs = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=3000, mixed_mode=1, inner_format=3)
    xd, rd = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=3000, mixed_mode=1)
    assert rd.operator_used == vh.INNER_DIA_CODED and rs.operator_used == vh.INNER_SELL64_D16
    assert rd.converged and rd.iters == rs.iters and rd.restarts == rs.restarts and rd.error == rs.error
    assert np.array_equal(xd, xs)


def test_dia_inner_copy_refuses_what_it_cannot_hold(dev, vh):
    """More than 32 diagonals (a random graph, a band of 33) or more than 25% zeros: DIA, forced,
    is refused; the automatic choice falls back to SELL-64 or CSR and solves."""
    _, rp, col, val = vh.graph_laplacian_batch(1, 3000, 3000, 8.0, 0.1, 1.0, 0.01, 3)
    b = vh.random_uniform(1, 3000)
    with pytest.raises(vh.VclHipError) as e:
        run_mixed(vh, rp, col, val, b, tol=1e-8, max_iters=3000, mixed_mode=1, inner_format=4)
    assert e.value.status == -3  # VCLHIP_ERROR_NOT_SUPPORTED
    x, res = run_mixed(vh, rp, col, val, b, tol=1e-8, max_iters=3000, mixed_mode=1)
    assert res.converged and true_relres(rp, col, val, x, b) < 1e-8
    # a symmetric band of half-width 16 (33 diagonals), diagonally dominant
    n = 2000
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(max(0, i - 16), min(n, i + 17)):
            rows.append(i)
            cols.append(j)
            vals.append(40.0 if i == j else -1.0 / (1 + abs(i - j)))
    rpb = np.zeros(n + 1, np.uint32)
    np.add.at(rpb, np.array(rows) + 1, 1)
    rpb = np.cumsum(rpb).astype(np.uint32)
    colb, valb = np.array(cols, np.uint32), np.array(vals)
    bb = vh.random_uniform(2, n)
    with pytest.raises(vh.VclHipError):
        run_mixed(vh, rpb, colb, valb, bb, tol=1e-8, max_iters=500, mixed_mode=1, inner_format=4)
    x, res = run_mixed(vh, rpb, colb, valb, bb, tol=1e-8, max_iters=500, mixed_mode=1)
    assert res.converged and true_relres(rpb, colb, valb, x, bb) < 1e-8


def test_sell16_refuses_far_columns(dev, vh):
    # a 200 x 200 x 3 grid: the z-neighbours are 40000 rows away, beyond a 16-bit offset
    rp, col, val = vh.laplace_3d(200, 200, 3)
    b = vh.random_uniform(0, rp.size - 1)
    with pytest.raises(vh.VclHipError):
        run_mixed(vh, rp, col, val, b, tol=1e-6, max_iters=50, mixed_mode=1, inner_format=3)
    x, res = run_mixed(vh, rp, col, val, b, tol=1e-6, max_iters=3000, mixed_mode=1)  # automatic: 32-bit columns
    assert res.converged and true_relres(rp, col, val, x, b) < 1e-6


def test_edge_cases(dev, vh):
    rp, col, val = vh.laplace_3d(12, 10, 9)
    n = rp.size - 1
    # b = 0: x = 0 without iterations (mixed_precision_cg.hpp:112-113)
    x, res = run_mixed(vh, rp, col, val, np.zeros(n), tol=1e-10, max_iters=100, mixed_mode=1)
    assert res.iters == 0 and res.converged and not x.any()
    # iteration limit: the last iteration folds x32 into x (mixed_precision_cg.hpp:161)
    b = vh.random_uniform(0, n)
    for mode in (0, 1):
        x, res = run_mixed(vh, rp, col, val, b, tol=1e-14, max_iters=7, mixed_mode=mode)
        assert res.iters == 7 and not res.converged and res.restarts >= 1
        assert abs(true_relres(rp, col, val, x, b) - res.error) <= 1e-6 * res.error
    # mode 0 has no preconditioner; unknown modes are rejected
    with pytest.raises(vh.VclHipError):
        run_mixed(vh, rp, col, val, b, precond=1, mixed_mode=0)
    with pytest.raises(vh.VclHipError):
        run_mixed(vh, rp, col, val, b, mixed_mode=2)


def test_history_is_the_fp32_residual(dev, vh, oracle):
    rp, col, val = vh.laplace_3d(16, 14, 12)
    b = vh.random_uniform(0, rp.size - 1)
    h = np.zeros(400)
    x, res = run_mixed(vh, rp, col, val, b, tol=1e-10, max_iters=400, mixed_mode=1, history=h)
    _, it_o, _, _, ho = oracle.mixed_cg2(oracle.CsrHost(rp, col, val), b, 1e-10, 400, mode=1, history=True)
    m = min(res.iters, it_o)
    assert np.allclose(h[:m], ho[:m], rtol=1e-3, atol=0)
