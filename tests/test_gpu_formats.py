"""GPU parity of coordinate_matrix, ell_matrix and hyb_matrix: SpMV (prod_impl), the fused pipelined
products, the diagonal, and the CG / BiCGStab / GMRES solvers on each format.

Bars: SpMV within 1e-13 relative inf-norm of the reference (FP64) -- ELL, and HYB without a plan,
sum each row in one lane in the reference's order and are bit-exact; HYB with the plan of its CSR
part is bit-exact on every row whose CSR part has at most 32 entries -- and 1e-5 in FP32; the fused
products' Ap likewise, their inner products within 1e-12 relative; solver
iteration counts within 2% of the reference's fixture, with the FP64 true residual below the
tolerance.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
FORMATS = (("coo", 2), ("ell", 3), ("hyb", 4))


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def device_matrix(vh, fmt, rp, col, val, dtype=np.float64):
    if fmt == 4:
        return vh.HybMatrix(rp, col, val, dtype=dtype, plan=True)
    cls = {2: vh.CooMatrix, 3: vh.EllMatrix}[fmt]
    return cls(rp, col, val, dtype=dtype)


def long_csr_part_rows(H):
    """Rows of a HybMatrix whose CSR part is summed by a wavefront (more than 32 entries)."""
    cr = H.host[2].astype(np.int64)
    return np.diff(cr) > 32


def relinf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def iters_close(got, want):
    return abs(int(got) - int(want)) <= max(2, int(np.ceil(0.02 * want)))


def true_relres(rp, col, val, x, b):
    import pyoracle
    r = b - pyoracle.csr_spmv(rp, col, val, np.asarray(x, np.float64))
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_format_spmv_matches_reference(dev, vh, tag, fmt):
    g = load("formats_powerlaw_700")
    rp, col, val = g["row_ptr"], g["col"], g["val"]
    n = rp.size - 1
    A = device_matrix(vh, fmt, rp, col, val)
    x = vh.DeviceArray.from_numpy(g["x"])
    for at, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
        y = vh.DeviceArray.from_numpy(g["y0"])
        A.spmv(x, y, alpha, beta)
        got, want = y.to_numpy(), g[f"{tag}_y_{at}"]
        print(f"[formats] {tag} spmv {at}: relinf {relinf(got, want):.2e}", flush=True)
        if fmt == 3:
            assert np.array_equal(got, want), at  # one lane per row, in the reference's order
        elif fmt == 4:
            short = ~long_csr_part_rows(A)
            assert np.array_equal(got[short], want[short]) and relinf(got, want) <= 1e-13, at
            Hn = vh.HybMatrix(rp, col, val, plan=False)  # one lane per row: the reference's order
            yn = vh.DeviceArray.from_numpy(g["y0"])
            Hn.spmv(x, yn, alpha, beta)
            assert np.array_equal(yn.to_numpy(), want), at
        else:
            # COO runs the CSR kernel: rows of more than 32 entries are summed by a wavefront, and
            # the reference scales every product by alpha
            assert relinf(got, want) <= 1e-13, at
    # strided x and y (start 3, stride 2)
    xs = np.zeros(2 * n + 3)
    xs[3::2] = g["x"]
    ys = np.full(2 * n + 3, 7.0)
    dxs, dys = vh.DeviceArray.from_numpy(xs), vh.DeviceArray.from_numpy(ys)
    A.spmv(dxs, dys, 1.0, 0.0, 3, 2, 3, 2)
    out = dys.to_numpy()
    assert relinf(out[3::2], g[f"{tag}_y_a1b0"]) <= 1e-13 and np.all(out[:3] == 7.0) and np.all(out[4::2] == 7.0)
    # FP32
    A32 = device_matrix(vh, fmt, rp, col, val, np.float32)
    x32 = vh.DeviceArray.from_numpy(g["x"].astype(np.float32))
    y32 = vh.DeviceArray.zeros(n, np.float32)
    A32.spmv(x32, y32)
    assert relinf(y32.to_numpy().astype(np.float64), g[f"{tag}_y_a1b0"]) <= 1e-5


@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_format_pipelined_products(dev, vh, tag, fmt):
    g = load("formats_powerlaw_700")
    rp, col, val = g["row_ptr"], g["col"], g["val"]
    n = rp.size - 1
    A = device_matrix(vh, fmt, rp, col, val)
    m = vh.generic(A)
    p, r0 = vh.DeviceArray.from_numpy(g["x"]), vh.DeviceArray.from_numpy(g["r0"])
    Ap, buf = vh.DeviceArray.zeros(n, np.float64), vh.DeviceArray.zeros(768, np.float64)
    vh.call("vclhip_pipelined_cg_prod_f64", C.byref(m), p.ptr, Ap.ptr, buf.ptr, 256, None)
    assert relinf(Ap.to_numpy(), g[f"{tag}_cgprod_Ap"]) <= (0.0 if fmt == 3 else 1e-13)
    hb, gb = buf.to_numpy(), g[f"{tag}_cgprod_buf"]
    for k in (256, 512):
        assert abs(hb[k] - gb[k]) <= 1e-12 * abs(gb[k]), k
    Ap, buf = vh.DeviceArray.zeros(n, np.float64), vh.DeviceArray.zeros(1536, np.float64)
    vh.call("vclhip_pipelined_bicgstab_prod_f64", C.byref(m), p.ptr, Ap.ptr, r0.ptr, buf.ptr, 256, 768, None)
    assert relinf(Ap.to_numpy(), g[f"{tag}_bprod_Ap"]) <= (0.0 if fmt == 3 else 1e-13)
    hb, gb = buf.to_numpy(), g[f"{tag}_bprod_buf"]
    for k in (256, 512, 768):
        assert abs(hb[k] - gb[k]) <= 1e-12 * abs(gb[k]), k


@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_format_diagonal(dev, vh, oracle, tag, fmt):
    rp, col, val = vh.convection_diffusion(9, 8, 3)
    n = rp.size - 1
    A = device_matrix(vh, fmt, rp, col, val)
    d = vh.DeviceArray.zeros(n, np.float64)
    vh.matrix_diag(A, d)
    assert np.array_equal(d.to_numpy(), oracle.row_info(rp, col, val, 3))


@pytest.mark.parametrize("case", ["lap2d_30x22", "convdiff_12x9x4", "convdiff_12x9x4_gmres"])
def test_format_solvers_match_reference(dev, vh, oracle, case):
    g = load(f"solve_formats_{case}")
    rp, col, val, b = g["row_ptr"], g["col"], g["val"], g["b"]
    kind, tol, mi = int(g["kind"]), float(g["tol"]), int(g["max_iters"])
    n = rp.size - 1
    db = vh.DeviceArray.from_numpy(b)
    for tag, fmt in FORMATS:
        A = device_matrix(vh, fmt, rp, col, val)
        x = vh.DeviceArray.zeros(n, np.float64)
        if kind == 0:
            res, want = vh.cg(A, db, x, tol=tol, max_iters=mi), int(g[f"{tag}_iters"])
        elif kind == 1:
            res, want = vh.bicgstab(A, db, x, tol=tol, max_iters=mi), int(g[f"{tag}_iters"])
        else:
            # the reference's OpenMP GMRES leaves out the tail of its Gram-Schmidt products; the
            # device solver is compared with the oracle's full-product GMRES
            res = vh.gmres(A, db, x, tol=tol, max_iters=mi, krylov_dim=20)
            want = oracle.gmres(oracle.CsrHost(rp, col, val), b, tol, mi, 20)[1]
        rel = true_relres(rp, col, val, x.to_numpy(), b)
        print(f"[formats] {case} {tag}: {res.iters} iterations (reference {want}), true relres {rel:.2e}",
              flush=True)
        assert iters_close(res.iters, want)
        assert res.converged and rel < 10 * tol


@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_format_jacobi_and_fp32(dev, vh, tag, fmt):
    """Jacobi CG on each format matches Jacobi CG on CSR (the same row sums; only the order of the
    per-workgroup partial inner products differs); FP32 CG converges to 1e-5."""
    k = np.exp(vh.random_normal(3, 24 * 20, 0.0, 1.0))
    rp, col, val = vh.diffusion(24, 20, 1, k)
    n = rp.size - 1
    b = vh.random_uniform(4, n)
    db = vh.DeviceArray.from_numpy(b)
    xc, xf = vh.DeviceArray.zeros(n, np.float64), vh.DeviceArray.zeros(n, np.float64)
    rc = vh.cg(vh.CsrMatrix(rp, col, val), db, xc, precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=3000)
    rf = vh.cg(device_matrix(vh, fmt, rp, col, val), db, xf, precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=3000)
    assert iters_close(rf.iters, rc.iters) and rf.converged
    assert np.linalg.norm(xf.to_numpy() - xc.to_numpy()) <= 1e-8 * np.linalg.norm(xc.to_numpy())
    A32 = device_matrix(vh, fmt, rp, col, val, np.float32)
    b32 = vh.DeviceArray.from_numpy(b.astype(np.float32))
    x32 = vh.DeviceArray.zeros(n, np.float32)
    r32 = vh.cg(A32, b32, x32, tol=1e-5, max_iters=3000)
    assert r32.converged and true_relres(rp, col, val, x32.to_numpy(), b) < 1e-4


def test_hyb_plan_long_rows(dev, vh, oracle):
    """HYB with the plan of its CSR part on rows of every kind: empty rows, rows inside the ELL
    width, CSR parts of up to 32 entries (one lane), longer ones (a wavefront) and one row of more
    than 2048 entries (a workgroup of its own).  Compared with the one-lane-per-row HYB, which sums
    in the reference's order, and with the oracle's CSR product."""
    rng = np.random.default_rng(7)
    n = 3000
    lens = rng.integers(0, 7, n)
    lens[rng.choice(n, 60, replace=False)] = rng.integers(20, 400, 60)
    lens[1234] = 2600
    rows = []
    for r in range(n):
        c = np.sort(rng.choice(n, lens[r], replace=False)).astype(np.uint32)
        rows.append(c)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([c.size for c in rows])
    col = np.concatenate(rows)
    val = rng.standard_normal(col.size)
    val[::17] = 0.0  # explicit zeros: skipped in the ELL part, multiplied in the CSR part
    xh = rng.standard_normal(n)
    x = vh.DeviceArray.from_numpy(xh)
    H, Hn = vh.HybMatrix(rp, col, val), vh.HybMatrix(rp, col, val, plan=False)
    assert H.planned and not Hn.planned and long_csr_part_rows(H).sum() > 10
    want = oracle.csr_spmv(rp, col, val, xh)
    for alpha, beta in ((1.0, 0.0), (2.0, 0.5)):
        y0 = rng.standard_normal(n)
        y, yn = vh.DeviceArray.from_numpy(y0), vh.DeviceArray.from_numpy(y0)
        H.spmv(x, y, alpha, beta)
        Hn.spmv(x, yn, alpha, beta)
        got, ref = y.to_numpy(), yn.to_numpy()
        short = ~long_csr_part_rows(H)
        assert np.array_equal(got[short], ref[short])
        assert relinf(got, ref) <= 1e-13
        assert relinf(got, alpha * want + beta * y0) <= 1e-13
    # the fused CG product through the plan: Ap and its inner products
    m, mn = vh.generic(H), vh.generic(Hn)
    Ap, Apn = vh.DeviceArray.zeros(n, np.float64), vh.DeviceArray.zeros(n, np.float64)
    buf, bufn = vh.DeviceArray.zeros(768, np.float64), vh.DeviceArray.zeros(768, np.float64)
    vh.call("vclhip_pipelined_cg_prod_f64", C.byref(m), x.ptr, Ap.ptr, buf.ptr, 256, None)
    vh.call("vclhip_pipelined_cg_prod_f64", C.byref(mn), x.ptr, Apn.ptr, bufn.ptr, 256, None)
    assert relinf(Ap.to_numpy(), Apn.to_numpy()) <= 1e-13
    hb, hn = buf.to_numpy(), bufn.to_numpy()
    for k in (256, 512):
        assert abs(hb[k] - hn[k]) <= 1e-12 * abs(hn[k]), k
    # FP32
    H32 = vh.HybMatrix(rp, col, val, dtype=np.float32)
    x32 = vh.DeviceArray.from_numpy(xh.astype(np.float32))
    y32 = vh.DeviceArray.zeros(n, np.float32)
    H32.spmv(x32, y32)
    assert relinf(y32.to_numpy().astype(np.float64), want) <= 1e-5
