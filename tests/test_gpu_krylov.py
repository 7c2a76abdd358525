"""GPU parity of the pipelined operations and the CG / BiCGStab solvers.

Bars: iteration counts within 2% of the reference (golden fixtures / oracle), the same
stopping quantity below the tag tolerance, and the FP64 true residual ||b - A x|| / ||b|| verified.
Element-wise updates are bit-exact against the reference (no contraction, same expressions);
inner products differ only in summation order.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def close_sum(a, b, scale):
    return abs(a - b) <= 1e-12 * scale


def iters_close(got, want):
    return abs(int(got) - int(want)) <= max(2, int(np.ceil(0.02 * want)))


def iters_in_envelope(got, env, oracle_env=None):
    """Within 2% of the reference's own iteration range when only the summation order of its inner
    products changes: its OpenMP thread counts 1..8 (fixture) and, when given, the oracle's k-block
    reduction orders (the same algorithm, pinned bit-exact to the reference at k = 0)."""
    lo, hi = int(env[0]), int(env[1])
    if oracle_env is not None:
        lo, hi = min(lo, oracle_env[0]), max(hi, oracle_env[1])
    return int(np.floor(0.98 * lo)) <= int(got) <= int(np.ceil(1.02 * hi))


def true_relres(rp, col, val, x, b):
    import pyoracle
    r = b - pyoracle.csr_spmv(rp, col, val, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


def test_pipelined_ops_golden(dev, vh):
    g = load("pipelined_ops_25x17")
    G = [np.ascontiguousarray(v) for v in g["g"]]
    n = G[0].size
    A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
    d = [vh.DeviceArray.from_numpy(v) for v in G]
    # pipelined_cg_vector_update
    x, p, r, Ap = (vh.DeviceArray.from_numpy(G[i]) for i in range(4))
    buf = vh.DeviceArray.zeros(768, np.float64)
    vh.call("vclhip_pipelined_cg_vector_update_f64", n, x.ptr, 0.37, p.ptr, r.ptr, Ap.ptr, -0.61, buf.ptr, 256, None)
    assert np.array_equal(x.to_numpy(), g["cgvu_x"]) and np.array_equal(p.to_numpy(), g["cgvu_p"])
    assert np.array_equal(r.to_numpy(), g["cgvu_r"])
    hb = buf.to_numpy()
    assert close_sum(hb[0], g["cgvu_buf"][0], n) and not hb[1:].any()
    # pipelined_cg_prod
    Ap2 = vh.DeviceArray.zeros(n, np.float64)
    buf2 = vh.DeviceArray.zeros(768, np.float64)
    vh.call("vclhip_pipelined_cg_prod_csr_f64", C.byref(A.desc), d[1].ptr, Ap2.ptr, buf2.ptr, 256, None)
    assert np.array_equal(Ap2.to_numpy(), g["cgprod_Ap"])
    hb, gb = buf2.to_numpy(), g["cgprod_buf"]
    assert close_sum(hb[256], gb[256], 100 * n) and close_sum(hb[512], gb[512], 100 * n)
    assert not hb[257:512].any() and not hb[513:].any()
    # pipelined_bicgstab_prod with offset 768
    Ap3 = vh.DeviceArray.zeros(n, np.float64)
    buf3 = vh.DeviceArray.zeros(1536, np.float64)
    vh.call("vclhip_pipelined_bicgstab_prod_csr_f64", C.byref(A.desc), d[1].ptr, Ap3.ptr, d[4].ptr, buf3.ptr, 256, 768,
            None)
    assert np.array_equal(Ap3.to_numpy(), g["bprod_Ap"])
    hb, gb = buf3.to_numpy(), g["bprod_buf"]
    for k in (256, 512, 768):
        assert close_sum(hb[k], gb[k], 100 * n)
    # pipelined_bicgstab_update_s: alpha = 3.5 / -1.25 from the buffer, on the device
    s = vh.DeviceArray.zeros(n, np.float64)
    b4 = np.zeros(1536)
    b4[0], b4[768] = 3.5, -1.25
    buf4 = vh.DeviceArray.from_numpy(b4)
    vh.call("vclhip_pipelined_bicgstab_update_s_f64", n, s.ptr, d[2].ptr, d[3].ptr, buf4.ptr, 256, 1280, None)
    assert np.array_equal(s.to_numpy(), g["upds_s"])
    assert close_sum(buf4.to_numpy()[1280], g["upds_buf"][1280], 100 * n)
    # pipelined_bicgstab_vector_update
    x5, p5, r5 = (vh.DeviceArray.from_numpy(G[i]) for i in range(3))
    buf5 = vh.DeviceArray.zeros(1536, np.float64)
    vh.call("vclhip_pipelined_bicgstab_vector_update_f64", n, x5.ptr, 0.3, p5.ptr, 0.7, d[5].ptr, r5.ptr, d[6].ptr,
            -0.2, d[7].ptr, d[4].ptr, buf5.ptr, 256, None)
    assert np.array_equal(x5.to_numpy(), g["bvu_x"]) and np.array_equal(p5.to_numpy(), g["bvu_p"])
    assert np.array_equal(r5.to_numpy(), g["bvu_r"])
    assert close_sum(buf5.to_numpy()[0], g["bvu_buf"][0], n)


def solve(vh, kind, A, b, precond=0, **kw):
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(b.size, np.float64)
    fn = vh.cg if kind == "cg" else vh.bicgstab
    res = fn(A, db, dx, precond=precond, **kw)
    return dx.to_numpy(), res


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_cg_lap2d_golden(dev, vh, fmt):
    if fmt == "csr":
        g = load("solve_lap2d_40x40")
        A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
    else:
        g = load("solve_sell_lap2d_37x29")
        A = vh.SellMatrix(g["row_ptr"], g["col"], g["val"], int(g["C"]))
    x, res = solve(vh, "cg", A, g["b"], tol=1e-8, max_iters=2000)
    assert iters_close(res.iters, g["cg_iters"]) and iters_in_envelope(res.iters, g["cg_env"]) and res.converged
    assert res.error < 1e-8
    assert true_relres(g["row_ptr"], g["col"], g["val"], x, g["b"]) < 1.05e-8
    assert np.linalg.norm(x - g["cg_x"]) / np.linalg.norm(g["cg_x"]) < 1e-6


def test_cg_max_iters_golden(dev, vh):
    g = load("solve_lap2d_40x40")
    A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
    x, res = solve(vh, "cg", A, g["b"], tol=1e-8, max_iters=40)
    assert res.iters == 40 and not res.converged
    assert abs(res.error - g["cg40_error"]) <= 1e-9 * g["cg40_error"]
    assert np.linalg.norm(x - g["cg40_x"]) / np.linalg.norm(g["cg40_x"]) < 1e-10


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_bicgstab_golden(dev, vh, oracle, fmt):
    if fmt == "csr":
        g = load("solve_lap2d_40x40")
        A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
        H = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    else:
        g = load("solve_sell_lap2d_37x29")
        A = vh.SellMatrix(g["row_ptr"], g["col"], g["val"], int(g["C"]))
        H = oracle.SellHost(g["row_ptr"], g["col"], g["val"], int(g["C"]))
    x, res = solve(vh, "bicgstab", A, g["b"], tol=1e-8, max_iters=2000)
    oenv = oracle.iteration_envelope(lambda: oracle.bicgstab_pipelined(H, g["b"], 1e-8, 2000))
    assert iters_in_envelope(res.iters, g["bicgstab_env"], oenv), (res.iters, g["bicgstab_env"], oenv)
    assert res.converged and res.error < 1e-8
    # x includes the final update, so the true residual is the tested one
    assert true_relres(g["row_ptr"], g["col"], g["val"], x, g["b"]) < 1.1e-8


def test_jacobi_cg_and_bicgstab_golden(dev, vh, oracle):
    g = load("solve_diffusion2d_30x30")
    A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
    H = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    diag = oracle.row_info(g["row_ptr"], g["col"], g["val"], 3)
    x, res = solve(vh, "cg", A, g["b"], precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=3000)
    assert iters_close(res.iters, g["jcg_iters"]) and res.converged and res.error < 1e-10
    assert true_relres(g["row_ptr"], g["col"], g["val"], x, g["b"]) < 1e-8
    rb = int(g["restart_every"])
    x, res = solve(vh, "bicgstab", A, g["b"], precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=3000,
                   max_iters_before_restart=rb)
    oenv = oracle.iteration_envelope(lambda: oracle.bicgstab_classic(H, g["b"], 1e-10, 3000, diag=diag,
                                                                     max_before_restart=rb))
    assert iters_in_envelope(res.iters, g["jbicgstab_env"], oenv), (res.iters, g["jbicgstab_env"], oenv)
    assert res.converged and res.error < 1e-10
    # the fully pipelined variant (rounding differs from the reference's recurrences) converges too
    x, res = solve(vh, "bicgstab", A, g["b"], precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=3000,
                   max_iters_before_restart=int(g["restart_every"]), variant=1)
    assert res.converged and res.error < 1e-10


def test_powerlaw_bicgstab_golden(dev, vh):
    g = load("solve_powerlaw_2000")
    A = vh.CsrMatrix(g["row_ptr"], g["col"], g["val"])
    x, res = solve(vh, "bicgstab", A, g["b"], tol=1e-10, max_iters=400)
    assert res.iters == g["bicgstab_iters"] == 400 and not res.converged
    assert 0.25 * g["bicgstab_error"] < res.error < 4 * g["bicgstab_error"]


def test_residual_history_tracks_oracle(dev, vh, oracle):
    rp, col, val = vh.laplace_3d(20, 18, 16)
    b = vh.random_uniform(0, rp.size - 1)
    A = vh.CsrMatrix(rp, col, val)
    hist = np.zeros(1000)
    x, res = solve(vh, "cg", A, b, tol=1e-10, max_iters=1000, history=hist)
    _, it, err, ho = oracle.cg_pipelined(oracle.CsrHost(rp, col, val), b, 1e-10, 1000, history=True)
    assert iters_close(res.iters, it)
    m = min(res.iters, it)
    assert np.allclose(hist[:m // 2], ho[:m // 2], rtol=1e-7)


def test_launch_modes_are_deterministic(dev, vh):
    rp, col, val = vh.laplace_3d(24)
    b = vh.random_uniform(0, rp.size - 1)
    A = vh.CsrMatrix(rp, col, val)
    ref_x, ref_res = solve(vh, "cg", A, b, tol=1e-9, max_iters=500, chunk=16)
    for kw in (dict(chunk=1), dict(chunk=7), dict(chunk=16, use_graph=True), dict(chunk=5, use_graph=True)):
        x, res = solve(vh, "cg", A, b, tol=1e-9, max_iters=500, **kw)
        assert res.iters == ref_res.iters and res.error == ref_res.error and np.array_equal(x, ref_x), kw
    bx, bres = solve(vh, "bicgstab", A, b, tol=1e-9, max_iters=500, chunk=16)
    x, res = solve(vh, "bicgstab", A, b, tol=1e-9, max_iters=500, chunk=3, use_graph=True)
    assert res.iters == bres.iters and np.array_equal(x, bx)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["cg", "jacobi_cg", "bicgstab", "jacobi_bicgstab", "gmres"])
def test_dia_operator_equals_csr(dev, vh, kind, dtype):
    """A CSR stencil matrix is solved on a DIA copy by default (operator_format AUTO); the results
    equal those on the CSR matrix as given, bit for bit.  A matrix with more than 32 diagonals keeps
    its CSR operator."""
    k = np.exp(vh.random_normal(8, 18 * 17 * 16))
    rp, col, val = vh.diffusion(18, 17, 16, k)
    n = rp.size - 1
    b = vh.random_uniform(3, n).astype(dtype)
    A = vh.CsrMatrix(rp, col, val, dtype=dtype)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    pc = 1 if kind.startswith("jacobi") else 0
    base = kind.replace("jacobi_", "")
    out = []
    for of in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, dtype)
        if base == "gmres":
            res = vh.gmres(A, db, dx, tol=tol, max_iters=3000, krylov_dim=20, operator_format=of)
        else:
            fn = vh.cg if base == "cg" else vh.bicgstab
            res = fn(A, db, dx, precond=pc, tol=tol, max_iters=3000, operator_format=of)
        out.append((dx.to_numpy(), res))
    (xa, ra), (xg, rg) = out
    assert ra.iters == rg.iters and ra.error == rg.error and ra.converged and np.array_equal(xa, xg)


def stencil_27(nx, ny, nz):
    """27-point stencil on an nx*ny*nz grid (row = x + nx (y + ny z)): -1 to every neighbour in the
    box, 26.5 on the diagonal (symmetric, strictly diagonally dominant: SPD).  27 diagonals, so the
    DIA kernel runs three full groups of eight and a partial fourth."""
    rows, cols = [], []
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = (x + nx * (y + ny * z)).ravel()
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ok = ((x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny) & (z + dz >= 0) & (z + dz < nz)).ravel()
                rows.append(r[ok])
                cols.append(r[ok] + dx + nx * (dy + ny * dz))
    return _csr_from_entries(nx * ny * nz, np.concatenate(rows), np.concatenate(cols), 26.5)


def band(n, offsets):
    """Band matrix with the given diagonals: -1/|o| on diagonal o != 0 (symmetric when the offsets
    are), the diagonal = the row's absolute off-diagonal sum + 0.5."""
    rows, cols, vals = [], [], []
    for o in offsets:
        if o == 0:
            continue
        i = np.arange(max(0, -o), min(n, n - o))
        rows.append(i)
        cols.append(i + o)
        vals.append(np.full(i.size, -1.0 / abs(o)))
    return _csr_from_entries(n, np.concatenate(rows), np.concatenate(cols), None, np.concatenate(vals))


def _csr_from_entries(n, rows, cols, diag, vals=None):
    import scipy.sparse as sp
    if vals is None:
        vals = np.where(rows == cols, diag, -1.0)
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    if diag is None:
        M = M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 0.5)
    M = sp.csr_matrix(M)
    M.sort_indices()
    return M.indptr.astype(np.uint32), M.indices.astype(np.uint32), M.data.astype(np.float64)


DIA_CASES = {
    "band5": (lambda: band(3000, range(-2, 3)), 5, True),
    "lap27": (lambda: stencil_27(14, 13, 12), 27, True),
    "band9": (lambda: band(3000, range(-4, 5)), 9, True),
    "band17": (lambda: band(3000, range(-8, 9)), 17, True),
    "band31": (lambda: band(3000, range(-15, 16)), 31, True),
    "band32": (lambda: band(3000, range(-15, 17)), 32, False),   # the kDiaMax boundary (nonsymmetric)
    "band33": (lambda: band(3000, range(-16, 17)), 33, True),    # one diagonal too many: keeps CSR
}


@pytest.mark.parametrize("case", list(DIA_CASES))
@pytest.mark.parametrize("kind", ["cg", "jacobi_cg", "bicgstab", "gmres"])
def test_dia_operator_diagonal_counts(dev, vh, case, kind):
    """Operators with 5 to 33 diagonals: up to 32 the default solve runs on the DIA copy (reported
    in result.operator_used), across the DIA kernel's full and partial groups of eight diagonals and
    at the 32-diagonal limit of the offset detection; 33 diagonals keep the CSR operator.  The DIA
    products equal the CSR kernel's bit for bit; the inner products fused into them are summed per
    256-row tile on the DIA copy and per row block of the CSR plan on the matrix as given.  With at
    most 8 entries per row every block is a 256-row tile, the partitions coincide and the two solves
    are equal bit for bit; with more, the blocks are shorter, the sums are grouped differently and
    the solves agree to rounding: iterations within 2% and x within 1e-9."""
    make, nd, symmetric = DIA_CASES[case]
    if kind.endswith("cg") and not symmetric:
        pytest.skip("CG needs a symmetric matrix")
    rp, col, val = make()
    n = rp.size - 1
    assert np.unique(col.astype(np.int64) - np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))).size == nd
    b = vh.random_uniform(5, n)
    A = vh.CsrMatrix(rp, col, val)
    pc = 1 if kind.startswith("jacobi") else 0
    base = kind.replace("jacobi_", "")
    out = []
    for of in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
        if base == "gmres":
            res = vh.gmres(A, db, dx, tol=1e-10, max_iters=3000, krylov_dim=30, operator_format=of)
        else:
            fn = vh.cg if base == "cg" else vh.bicgstab
            res = fn(A, db, dx, precond=pc, tol=1e-10, max_iters=3000, operator_format=of)
        out.append((dx.to_numpy(), res))
    (xa, ra), (xg, rg) = out
    assert ra.operator_used == (vh.USED_DIA if nd <= 32 else vh.USED_AS_GIVEN)
    assert rg.operator_used == vh.USED_AS_GIVEN
    assert ra.converged and rg.converged
    if int(np.diff(rp.astype(np.int64)).max()) <= 8:
        assert A.num_blocks == (n + 255) // 256  # the CSR plan's blocks are the 256-row tiles
        assert ra.iters == rg.iters and ra.error == rg.error and np.array_equal(xa, xg)
    else:
        assert iters_close(ra.iters, rg.iters), (ra.iters, rg.iters)
        assert np.linalg.norm(xa - xg) / np.linalg.norm(xg) < 1e-9


@pytest.mark.parametrize("bad", ["zero", "negative_zero", "nan"])
def test_dia_operator_refuses_values_it_would_skip(dev, vh, bad):
    """The DIA kernel skips zero (either sign) and NaN values, where the CSR kernel multiplies every
    stored entry (0 * Inf and NaN reach its sums): a matrix storing such a value keeps the CSR
    operator, so the default solve is the as-given one."""
    rp, col, val = vh.laplace_3d(10, 9, 8)
    n = rp.size - 1
    val = val.copy()
    k = int(rp[57]) + 1  # an off-diagonal entry of row 57
    val[k] = {"zero": 0.0, "negative_zero": -0.0, "nan": np.nan}[bad]
    A = vh.CsrMatrix(rp, col, val)
    b = vh.random_uniform(6, n)
    db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
    res = vh.cg(A, db, dx, tol=1e-10, max_iters=50)
    assert res.operator_used == vh.USED_AS_GIVEN
    d = vh.DiaOperator(A)
    assert not d.built and d.nd == 0
    d.close()


def test_dia_object_reused_across_solves(dev, vh):
    """vclhip_dia_create_f64 detects and copies once; solves given the object (params.dia) run on it
    with the results of a solve that builds its own copy; an object made for other arrays is
    rejected."""
    rp, col, val = stencil_27(12, 11, 10)
    n = rp.size - 1
    A = vh.CsrMatrix(rp, col, val)
    d = vh.DiaOperator(A)
    assert d.built and d.nd == 27
    outs = []
    for seed in (7, 8):
        b = vh.random_uniform(seed, n)
        for dia in (None, d):
            db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
            res = vh.cg(A, db, dx, tol=1e-10, max_iters=2000, dia=dia)
            assert res.operator_used == vh.USED_DIA
            outs.append((dx.to_numpy(), res.iters))
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]
    assert np.array_equal(outs[2][0], outs[3][0]) and outs[2][1] == outs[3][1]
    B = vh.CsrMatrix(rp, col, val)  # same values, other arrays
    db, dx = vh.DeviceArray.from_numpy(vh.random_uniform(7, n)), vh.DeviceArray.zeros(n, np.float64)
    with pytest.raises(vh.VclHipError):
        vh.cg(B, db, dx, tol=1e-10, max_iters=10, dia=d)
    with pytest.raises(vh.VclHipError):
        vh.cg(A, db, dx, tol=1e-10, max_iters=10, operator_format=3)  # not an operator format
    d.close()


def test_dia_operator_skips_unordered_rows(dev, vh):
    """A row whose columns are not strictly increasing (unsorted, or an entry split in two) keeps the
    CSR operator: the default solve equals the as-given one bit for bit and solves the summed matrix."""
    rp, col, val = vh.laplace_3d(14, 13, 12)
    n = rp.size - 1
    b = vh.random_uniform(4, n)
    # row 100: reverse its entries; row 200: split its diagonal into two entries of half the value
    col, val = col.copy(), val.copy()
    a, e = int(rp[100]), int(rp[101])
    col[a:e], val[a:e] = col[a:e][::-1].copy(), val[a:e][::-1].copy()
    a, e = int(rp[200]), int(rp[201])
    d = a + int(np.nonzero(col[a:e] == 200)[0][0])
    col = np.insert(col, d, 200)
    val = np.insert(val, d, val[d] / 2)
    val[d + 1] /= 2
    rp = rp.astype(np.int64)
    rp[201:] += 1
    rp = rp.astype(np.uint32)
    A = vh.CsrMatrix(rp, col, val)
    out = []
    for of in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
        res = vh.cg(A, db, dx, tol=1e-10, max_iters=2000, operator_format=of)
        out.append((dx.to_numpy(), res))
    (xa, ra), (xg, rg) = out
    assert ra.converged and ra.iters == rg.iters and np.array_equal(xa, xg)
    import scipy.sparse as sp
    M = sp.csr_matrix((val, col.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
    assert np.linalg.norm(b - M @ xa) / np.linalg.norm(b) < 1e-9


@pytest.mark.parametrize("precond", [0, 1])
def test_cg_compact_mode_matches_three_kernel_form(dev, vh, precond):
    """Small systems run CG with the scalar step fused into the next update (two kernels per
    iteration).  A monitor, or graph chunks of odd length, select the three-kernel form; every
    form gives the same iterations, error, history and x, bit for bit, also when the iteration
    limit ends the solve."""
    rp, col, val = vh.fdm_laplace_2d(48, 40)
    n = rp.size - 1
    b = vh.random_uniform(5, n)
    A = vh.CsrMatrix(rp, col, val)

    @vh.MONITOR
    def never(xptr, rel, user):
        return 0

    for mi in (1000, 37, 1):
        hc, h3 = np.zeros(mi), np.zeros(mi)
        xc, rc = solve(vh, "cg", A, b, precond=precond, tol=1e-10, max_iters=mi, history=hc)
        x3, r3 = solve(vh, "cg", A, b, precond=precond, tol=1e-10, max_iters=mi, history=h3, monitor=never)
        xg, rg = solve(vh, "cg", A, b, precond=precond, tol=1e-10, max_iters=mi, chunk=5, use_graph=True)
        for x, r in ((x3, r3), (xg, rg)):
            assert r.iters == rc.iters and r.error == rc.error and r.converged == rc.converged, mi
            assert np.array_equal(x, xc), mi
        assert np.array_equal(hc[: rc.iters], h3[: r3.iters]), mi
        assert rc.converged == (mi == 1000) and (mi == 1000 or rc.iters == mi)


def test_edge_cases(dev, vh):
    rp, col, val = vh.fdm_laplace_2d(16, 16)
    A = vh.CsrMatrix(rp, col, val)
    b0 = np.zeros(256)
    x, res = solve(vh, "cg", A, b0, tol=1e-8, max_iters=100)
    assert res.iters == 0 and not x.any()
    x, res = solve(vh, "bicgstab", A, b0, tol=1e-8, max_iters=100)
    assert res.iters == 0 and not x.any()
    b = vh.random_uniform(0, 256)
    x, res = solve(vh, "cg", A, b, tol=1e-8, max_iters=0)
    assert res.iters == 0 and res.error == 1.0 and not x.any()
    # a NaN in b must not hang: the iteration cap ends the solve
    bn = b.copy()
    bn[3] = np.nan
    x, res = solve(vh, "cg", A, bn, tol=1e-8, max_iters=50)
    assert res.iters == 50 and not res.converged
    # zero on the diagonal: Jacobi refuses (zero_on_diagonal_exception)
    d_idx = [k for k in range(rp[5], rp[6]) if col[k] == 5][0]
    val2 = val.copy()
    val2[d_idx] = 0.0
    A2 = vh.CsrMatrix(rp, col, val2)
    with pytest.raises(vh.VclHipError) as e:
        solve(vh, "cg", A2, b, precond=vh.PRECOND_JACOBI, tol=1e-8, max_iters=10)
    assert e.value.status == -2


def test_monitor_stops_and_sees_every_iteration(dev, vh):
    rp, col, val = vh.fdm_laplace_2d(32, 32)
    A = vh.CsrMatrix(rp, col, val)
    b = vh.random_uniform(0, 1024)
    seen = []

    @vh.MONITOR
    def mon(xptr, rel, user):
        seen.append(rel)
        return 1 if len(seen) == 17 else 0

    x, res = solve(vh, "cg", A, b, tol=1e-12, max_iters=1000, monitor=mon)
    assert res.iters == 17 and len(seen) == 17
    assert all(np.diff(seen[:5]) < 1.0)


@pytest.mark.parametrize("precond", [0, 1])
def test_3d_variable_coefficient_vs_oracle(dev, vh, oracle, precond):
    N = 40
    k = np.exp(vh.random_normal(1, N ** 3, 0.0, 1.0))
    rp, col, val = vh.diffusion(N, N, N, k)
    b = vh.random_uniform(0, N ** 3)
    A = vh.CsrMatrix(rp, col, val)
    x, res = solve(vh, "cg", A, b, precond=precond, tol=1e-10, max_iters=5000)
    H = oracle.CsrHost(rp, col, val)
    if precond:
        _, it, err, _ = oracle.cg_classic(H, b, 1e-10, 5000, diag=oracle.row_info(rp, col, val, 3))
    else:
        _, it, err, _ = oracle.cg_pipelined(H, b, 1e-10, 5000)
    assert iters_close(res.iters, it), (res.iters, it)
    assert res.converged and res.error < 1e-10
    assert true_relres(rp, col, val, x, b) < 1e-9


def test_jacobi_bicgstab_powerlaw_vs_oracle(dev, vh, oracle):
    rp, col, val = vh.powerlaw(30000, max_len=512, seed=1)
    b = vh.random_uniform(0, rp.size - 1)
    A = vh.CsrMatrix(rp, col, val)
    x, res = solve(vh, "bicgstab", A, b, precond=vh.PRECOND_JACOBI, tol=1e-10, max_iters=1000)
    _, it, err, _ = oracle.bicgstab_classic(oracle.CsrHost(rp, col, val), b, 1e-10, 1000,
                                            diag=oracle.row_info(rp, col, val, 3))
    assert iters_close(res.iters, it), (res.iters, it)
    assert res.converged and true_relres(rp, col, val, x, b) < 1e-8


FUSED_CASES = {
    "fdm_256": (lambda vh: vh.fdm_laplace_2d(256, 256), 0, dict(tol=1e-8, max_iters=2000)),
    "diffusion_jacobi": (lambda vh: vh.diffusion(90, 70, 1, np.exp(vh.random_normal(3, 90 * 70))), 1,
                         dict(tol=1e-10, max_iters=3000)),
    "lap27": (lambda vh: stencil_27(14, 13, 12), 0, dict(tol=1e-10, max_iters=3000)),
    "graph_chunks": (lambda vh: vh.fdm_laplace_2d(128, 100), 0, dict(tol=1e-10, max_iters=3000, use_graph=True)),
}


@pytest.mark.parametrize("case", list(FUSED_CASES))
def test_cg_one_kernel_per_iteration_equals_two(dev, vh, oracle, case):
    """Small systems on a DIA operator run one kernel per CG iteration (cg_fused_kernel: the scalar
    step of iteration k, the update of k + 1 and the product of k + 1 in one launch, reading r, p and
    A p from one copy and writing the other).  Against the two-kernel compact mode (variant 1): the
    same iteration count, error, per-iteration history and x bit for bit, with and without Jacobi, on
    a 27-diagonal stencil (full and partial groups of eight diagonals) and with hipGraph chunks; and
    the oracle's iteration count (same recurrences) within 2%."""
    make, pc, kw = FUSED_CASES[case]
    rp, col, val = make(vh)
    n = rp.size - 1
    b = vh.random_uniform(7, n)
    A = vh.CsrMatrix(rp, col, val)
    out = []
    for variant in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
        hist = np.zeros(kw["max_iters"])
        res = vh.cg(A, db, dx, precond=pc, variant=variant, history=hist, **kw)
        out.append((dx.to_numpy(), res, hist))
    (xf, rf, hf), (xc, rc, hc) = out
    print(f"{case}: {rf.iters} iterations, kernels per iteration {rf.kernels_per_iter} / {rc.kernels_per_iter}")
    assert rf.operator_used == vh.USED_DIA and rf.kernels_per_iter == 1 and rc.kernels_per_iter == 2
    assert rf.converged and rf.iters == rc.iters and rf.error == rc.error
    assert np.array_equal(xf, xc) and np.array_equal(hf, hc)
    Ah = oracle.CsrHost(rp, col, val)
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    diag = np.zeros(n)
    on = col.astype(np.int64) == rows
    diag[rows[on]] = val[on]
    if pc:
        _, it_o, _, _ = oracle.cg_classic(Ah, b, kw["tol"], kw["max_iters"], diag=diag)
    else:
        _, it_o, _, _ = oracle.cg_pipelined(Ah, b, kw["tol"], kw["max_iters"])
    assert iters_close(rf.iters, it_o), (rf.iters, it_o)
    # Jacobi stops on the D^-1-norm of r, which bounds the 2-norm ratio by sqrt(max d / min d) tol
    bound = 1.05 * kw["tol"] * (np.sqrt(diag.max() / diag.min()) if pc else 1.0)
    r = b - oracle.csr_spmv(rp, col, val, xf)
    assert np.linalg.norm(r) / np.linalg.norm(b) < bound


FUSED_BOUND_CASES = {
    # (operator, kernels per iteration expected): grid x (vector + 2 product partials) per iteration
    "lap2d_330": (lambda vh: vh.fdm_laplace_2d(330, 330), 1),  # 432 x (256 + 864) = 484 K reads
    "lap2d_392": (lambda vh: vh.fdm_laplace_2d(392, 392), 2),  # 608 x 1472 = 895 K
    "lap3d_40": (lambda vh: vh.laplace_3d(40), 1),  # 256 x 768 = 197 K
    "lap3d_50": (lambda vh: vh.laplace_3d(50), 2),  # 496 x 1248 = 619 K
}


@pytest.mark.parametrize("case", list(FUSED_BOUND_CASES))
def test_cg_fused_mode_only_below_its_partial_read_bound(dev, vh, case):
    """One kernel per iteration pays while the partial sums that all its workgroups reduce stay few:
    at most 2^19 reads per iteration (grid x (vector + 2 product partials)), with the grid resident
    at once.  Above it the solve keeps the two-kernel compact mode; either way the results equal
    variant 1's bit for bit."""
    make, want = FUSED_BOUND_CASES[case]
    rp, col, val = make(vh)
    n = rp.size - 1
    b = vh.random_uniform(9, n)
    A = vh.CsrMatrix(rp, col, val)
    out = []
    for variant in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, np.float64)
        res = vh.cg(A, db, dx, tol=1e-8, max_iters=3000, variant=variant)
        out.append((dx.to_numpy(), res))
    (x0, r0), (x1, r1) = out
    assert r0.operator_used == vh.USED_DIA and r0.kernels_per_iter == want and r1.kernels_per_iter == 2
    assert r0.converged and r0.iters == r1.iters and r0.error == r1.error and np.array_equal(x0, x1)


BICG_COMPACT_CASES = {
    "fdm_256": (lambda vh: vh.fdm_laplace_2d(256, 256), np.float64, "csr", dict(tol=1e-8, max_iters=2000)),
    "convdiff_nonsym": (lambda vh: vh.convection_diffusion(40, 36, 30), np.float64, "csr",
                        dict(tol=1e-10, max_iters=2000)),
    "sell_fp32": (lambda vh: vh.laplace_3d(30, 28, 26), np.float32, "sell", dict(tol=1e-5, max_iters=2000)),
    "max_iters": (lambda vh: vh.powerlaw(20000, seed=3), np.float64, "csr", dict(tol=1e-12, max_iters=60)),
    "graph_chunks": (lambda vh: vh.convection_diffusion(30, 30, 30), np.float64, "csr",
                     dict(tol=1e-10, max_iters=2000, use_graph=True, chunk=8)),
}


@pytest.mark.parametrize("case", list(BICG_COMPACT_CASES))
def test_bicgstab_four_kernels_per_iteration_equals_six(dev, vh, oracle, case):
    """Small unpreconditioned BiCGStab solves run four kernels per iteration (the alpha step at the
    head of the s update, the scalar step at the head of the vector update, each workgroup reducing
    the partials itself).  Against the six-kernel form (variant 1): the same iteration count, error,
    residual history and x bit for bit, on a symmetric and a nonsymmetric operator, in FP32 on
    SELL-64, when max_iters ends the solve (the final full update) and with hipGraph chunks; and the
    oracle's pipelined BiCGStab iteration count within its envelope."""
    make, dt, fmt, kw = BICG_COMPACT_CASES[case]
    rp, col, val = make(vh)
    n = rp.size - 1
    b = vh.random_uniform(5, n).astype(dt)
    if fmt == "csr":
        A = vh.CsrMatrix(rp, col, val, dtype=dt)
    else:
        A = vh.SellMatrix(rp, col, val, 64, dtype=dt)
    out = []
    for variant in (0, 1):
        db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(n, dt)
        hist = np.zeros(kw["max_iters"])
        res = vh.bicgstab(A, db, dx, variant=variant, history=hist, **kw)
        out.append((dx.to_numpy(), res, hist))
    (xc, rc, hc), (x6, r6, h6) = out
    print(f"{case}: {rc.iters} iterations, kernels per iteration {rc.kernels_per_iter} / {r6.kernels_per_iter}")
    assert rc.kernels_per_iter == 4 and r6.kernels_per_iter == 6
    assert rc.iters == r6.iters and rc.error == r6.error and rc.converged == r6.converged
    assert np.array_equal(xc, x6) and np.array_equal(hc, h6)
    if case == "max_iters":
        assert rc.iters == kw["max_iters"] and not rc.converged
        return
    assert rc.converged
    if dt == np.float64:
        H = oracle.CsrHost(rp, col, val)
        oenv = oracle.iteration_envelope(lambda: oracle.bicgstab_pipelined(H, b, kw["tol"], kw["max_iters"]))
        assert iters_in_envelope(rc.iters, oenv), (rc.iters, oenv)


def test_bicgstab_compact_mode_only_for_small_systems(dev, vh):
    """At 70^3 rows (343,000) every vector-update workgroup would read more than 2^20 partial sums per
    iteration in all, so the solve keeps six kernels; a monitor keeps compact mode (it sees each
    iteration after its vector update) and stops the solve where it asks to."""
    rp, col, val = vh.laplace_3d(70)
    n = rp.size - 1
    A = vh.CsrMatrix(rp, col, val)
    db = vh.DeviceArray.from_numpy(vh.random_uniform(2, n))
    dx = vh.DeviceArray.zeros(n, np.float64)
    res = vh.bicgstab(A, db, dx, tol=1e-6, max_iters=1000)
    assert res.kernels_per_iter == 6 and res.converged
    rp, col, val = vh.fdm_laplace_2d(128, 128)
    n = rp.size - 1
    A = vh.CsrMatrix(rp, col, val)
    db = vh.DeviceArray.from_numpy(vh.random_uniform(3, n))
    seen = []

    @vh.MONITOR
    def mon(x_dev, rel, user):
        seen.append(rel)
        return 1 if len(seen) == 25 else 0

    dx = vh.DeviceArray.zeros(n, np.float64)
    res = vh.bicgstab(A, db, dx, tol=1e-14, max_iters=1000, monitor=mon)
    assert res.kernels_per_iter == 4 and res.iters == 25 and len(seen) == 25


def test_solvers_refuse_vectors_of_another_precision_or_size(dev, vh):
    """The C ABI takes plain device pointers, so the binding refuses, before any launch, a right-hand
    side or solution whose precision differs from the matrix's or which is shorter than the matrix."""
    rp, col, val = vh.fdm_laplace_2d(16, 16)
    n = rp.size - 1
    A64 = vh.CsrMatrix(rp, col, val)
    S32 = vh.SellMatrix(rp, col, val, 64, dtype=np.float32)
    b32, x32 = vh.DeviceArray.zeros(n, np.float32), vh.DeviceArray.zeros(n, np.float32)
    b64, x64 = vh.DeviceArray.zeros(n, np.float64), vh.DeviceArray.zeros(n, np.float64)
    short = vh.DeviceArray.zeros(n - 1, np.float64)
    for fn in (vh.cg, vh.bicgstab, vh.gmres):
        with pytest.raises(ValueError):
            fn(A64, b32, x32)
        with pytest.raises(ValueError):
            fn(S32, b64, x64)
        with pytest.raises(ValueError):
            fn(A64, b64, short)
    with pytest.raises(ValueError):
        vh.mixed_cg(A64, b32, x32)


def test_concurrent_solves_on_two_streams_match_sequential(dev, vh):
    """The solvers keep their state per call (workspace, scalar record, partial sums on the caller's
    stream), so solves issued from two host threads on two streams run concurrently and give what
    they give one after the other, bit for bit: CG and BiCGStab, a small system (one kernel per CG
    iteration) beside a larger one (three), each thread repeating its solves."""
    import threading

    systems = [vh.fdm_laplace_2d(200, 200), vh.laplace_3d(48)]
    mats = [vh.CsrMatrix(rp, col, val) for rp, col, val in systems]
    sizes = [rp.size - 1 for rp, _, _ in systems]
    rhs = [vh.random_uniform(11 + i, sizes[i]) for i in (0, 1)]
    solvers = (vh.cg, vh.bicgstab)
    reps = 3
    # every array is made and filled before any solve starts (the solves' streams do not wait for
    # the null stream that uploads and fills)
    arrays = {(i, k, r): (vh.DeviceArray.from_numpy(rhs[i]), vh.DeviceArray.zeros(sizes[i], np.float64))
              for i in (0, 1) for k in range(len(solvers)) for r in range(reps + 1)}
    vh.call("vclhip_device_synchronize")

    def run(i, r, stream):
        out = []
        for k, fn in enumerate(solvers):
            db, dx = arrays[(i, k, r)]
            res = fn(mats[i], db, dx, tol=1e-9, max_iters=3000, stream=stream)
            out.append((int(res.iters), float(res.error), dx.to_numpy(stream)))
        return out

    want = [run(0, reps, None), run(1, reps, None)]
    streams = [vh.Stream(), vh.Stream()]
    got = [[], []]
    errors = []

    def worker(i):
        try:
            for r in range(reps):
                got[i].append(run(i, r, streams[i].handle))
        except Exception as e:  # reported below, in the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(i,)) for i in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads) and not errors, errors
    for i in (0, 1):
        assert len(got[i]) == reps
        for rep in got[i]:
            for (it, err, x), (it0, err0, x0) in zip(rep, want[i]):
                assert it == it0 and err == err0 and np.array_equal(x, x0)
    for s in streams:
        s.close()


def tridiag(n):
    """1-D Laplacian (2 on the diagonal, -1 beside it) in CSR."""
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(v)
    rp = np.zeros(n + 1, np.uint32)
    np.add.at(rp, np.array(rows) + 1, 1)
    return np.cumsum(rp).astype(np.uint32), np.array(cols, np.uint32), np.array(vals)


@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 257])
def test_tiny_systems_every_launch_mode(dev, vh, n):
    """Systems of 1 to 257 rows (partial packs, chunks smaller than a workgroup, grids at their
    minimum of 8 workgroups) in every launch mode: CG in one and two kernels per iteration and
    BiCGStab in four and six give the same x bit for bit and solve the system; a NaN in b ends at
    the iteration cap in every mode.  When s vanishes exactly (a one- or two-row system), the
    pipelined BiCGStab's omega is 0/0, as in the reference (bicgstab.hpp:193-195); there only the
    agreement of the modes is checked."""
    rp, col, val = tridiag(n)
    A = vh.CsrMatrix(rp, col, val)
    b = vh.random_uniform(3, n)
    for kind, modes in (("cg", (0, 1)), ("bicgstab", (0, 1))):
        exact_breakdown = kind == "bicgstab" and n <= 2
        xs = []
        for variant in modes:
            x, res = solve(vh, kind, A, b, tol=1e-10, max_iters=4 * n + 20, variant=variant)
            assert res.converged or exact_breakdown, (kind, variant, res.iters)
            xs.append((x, res.iters))
        assert np.array_equal(xs[0][0], xs[1][0], equal_nan=True) and xs[0][1] == xs[1][1], kind
        if not exact_breakdown:
            assert true_relres(rp, col, val, xs[0][0], b) < 1e-8
        bn = b.copy()
        bn[n // 2] = np.nan
        for variant in modes:
            _, res = solve(vh, kind, A, bn, tol=1e-10, max_iters=30, variant=variant)
            assert res.iters == 30 and not res.converged, (kind, variant)


_ABS_ORACLE = {}


def abs_tol_oracle(oracle, H, key, b, abs_tol, jacobi_diag):
    """The oracle's CG, Jacobi-CG and BiCGStab iteration counts at abs_tol (the last as its range
    over reduction orders), computed once per matrix."""
    if key not in _ABS_ORACLE:
        cg_it = oracle.cg_pipelined(H, b, 1e-15, 5000, abs_tol=abs_tol)[1]
        jac_it = None
        if jacobi_diag is not None:
            jac_it = oracle.cg_classic(H, b, 1e-15, 5000, diag=jacobi_diag, abs_tol=abs_tol)[1]
        env = oracle.iteration_envelope(lambda: oracle.bicgstab_pipelined(H, b, 1e-15, 5000, abs_tol=abs_tol))
        hist = oracle.bicgstab_pipelined(H, b, 1e-15, 5000, abs_tol=abs_tol, history=True)[3]
        _ABS_ORACLE[key] = (cg_it, jac_it, env, hist)
    return _ABS_ORACLE[key]


@pytest.mark.parametrize("case", ["fdm_256_fused", "fdm_256_two_kernels", "lap3d_96_dia", "lap3d_96_as_given",
                                  "lap3d_40_sell"])
def test_absolute_tolerance_stop_vs_oracle(dev, vh, oracle, case):
    """abs_tolerance(): CG stops when |<r,r>| < abs_tol^2 (cg.hpp:176, classic form :309) and
    BiCGStab when ||r|| < abs_tol (bicgstab.hpp:198, :465), whichever of the two tests comes first.
    With a relative tolerance out of reach (1e-15) the absolute one ends the solve at the oracle's
    iteration count in every launch mode (one fused kernel, two, three on the 96^3 operator; DIA
    copy, CSR as given, SELL-64), and ||b|| below abs_tol returns x = 0 at once (cg.hpp:151,
    bicgstab.hpp:140)."""
    m, kw = {"fdm_256_fused": (256, {}), "fdm_256_two_kernels": (256, {"variant": 1}),
             "lap3d_96_dia": (96, {}), "lap3d_96_as_given": (96, {"operator_format": vh.OPERATOR_AS_GIVEN}),
             "lap3d_40_sell": (40, {})}[case]
    rp, col, val = vh.fdm_laplace_2d(m, m) if case.startswith("fdm") else vh.laplace_3d(m)
    n = rp.size - 1
    b = vh.random_uniform(11, n)
    sell = case.endswith("sell")
    if sell:
        A, H = vh.SellMatrix(rp, col, val, 64), oracle.SellHost(rp, col, val, 64)
    else:
        A, H = vh.CsrMatrix(rp, col, val), oracle.CsrHost(rp, col, val)
    nb = float(np.linalg.norm(b))
    abs_tol = 1e-5 * nb
    diag = None if sell else oracle.row_info(rp, col, val, 3)
    cg_it, jac_it, env, bicg_hist = abs_tol_oracle(oracle, H, (case.split("_")[0], m, sell), b, abs_tol, diag)
    x, res = solve(vh, "cg", A, b, tol=1e-15, abs_tol=abs_tol, max_iters=5000, **kw)
    assert res.converged and iters_close(res.iters, cg_it), (res.iters, cg_it)
    assert true_relres(rp, col, val, x, b) < 1.1e-5
    if not sell:
        _, res = solve(vh, "cg", A, b, precond=vh.PRECOND_JACOBI, tol=1e-15, abs_tol=abs_tol, max_iters=5000, **kw)
        assert res.converged and iters_close(res.iters, jac_it), (res.iters, jac_it)
    # BiCGStab's residual is jagged at this level: the oracle's own count moves by up to 13% with
    # the summation order (114-129 on the 96^3 case), and the reference's, whose OpenMP reductions
    # combine thread partials in arrival order, from run to run (345-411 over four runs of the 256^2
    # case on 8 threads).  The bar is the oracle's early residual history, the stop at the first
    # step below abs_tol, and the count within 10% of the oracle's range.
    hist = np.zeros(5000)
    x, res = solve(vh, "bicgstab", A, b, tol=1e-15, abs_tol=abs_tol, max_iters=5000, history=hist, **kw)
    h = hist[: res.iters]
    assert res.converged and h[-1] < abs_tol and (h[:-1] >= abs_tol).all()  # ||r|| per step
    assert np.allclose(h[:8], bicg_hist[:8], rtol=1e-6, atol=0)
    assert 0.9 * env[0] <= res.iters <= 1.1 * env[1], (res.iters, env)
    assert true_relres(rp, col, val, x, b) < 1.1e-5
    for kind in ("cg", "bicgstab"):
        x, res = solve(vh, kind, A, 1e-3 * b, tol=1e-15, abs_tol=1e-2 * nb, max_iters=5000, **kw)
        assert res.iters == 0 and not x.any(), kind
