"""Pins the oracle (oracle/vcl_oracle.c) to the real ViennaCL host backend.

Two sources: the committed golden fixtures (tests/golden/*.npz, written by make_golden.py from the
reference build), and -- where oracle/_ref is built -- direct calls into the reference build.
Single-threaded results must agree bit for bit.
"""
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


@pytest.mark.parametrize("case", ["lap2d_30x20", "powerlaw_3000", "empty_rows_400"])
def test_csr_spmv_matches_reference(oracle, case):
    g = load(f"spmv_{case}")
    for tag, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
        y = oracle.csr_spmv(g["row_ptr"], g["col"], g["val"], g["x"], alpha, g["y0"], beta)
        assert np.array_equal(y, g[f"y_{tag}"]), tag


@pytest.mark.parametrize("case", ["lap2d_37x29_c32", "powerlaw_1001_c64"])
def test_sell_arrays_and_spmv_match_reference(oracle, vh, case):
    g = load(f"sell_{case}")
    C_ = int(g["C"])
    S = oracle.SellHost(g["row_ptr"], g["col"], g["val"], C_)
    assert np.array_equal(S.cpb, g["columns_per_block"])
    assert np.array_equal(S.bs, g["block_start"])
    assert np.array_equal(S.ci, g["column_indices"])
    assert np.array_equal(S.el, g["elements"])
    # the library's host-side converter builds the same arrays
    cpb, ci, bs, el = vh.sell_from_csr(g["row_ptr"], g["col"], g["val"], C_)
    assert np.array_equal(cpb, g["columns_per_block"]) and np.array_equal(bs, g["block_start"])
    assert np.array_equal(ci, g["column_indices"]) and np.array_equal(el, g["elements"])
    y = oracle.sell_spmv(S, g["x"], float(g["alpha"]), g["y0"], float(g["beta"]))
    assert np.array_equal(y, g["y"])


def test_row_blocks_and_diag_match_reference(oracle, vh):
    g = load("row_blocks_powerlaw_5000")
    rb = np.zeros(g["row_ptr"].size, np.uint32)
    nb = oracle.orc().orc_row_blocks(g["row_ptr"].size - 1, oracle.p(oracle.cu32(g["row_ptr"])), oracle.p(rb))
    assert np.array_equal(rb[: nb + 1], g["row_blocks"])
    assert np.array_equal(vh.reference_row_blocks(g["row_ptr"]), g["row_blocks"])
    d = oracle.row_info(g["row_ptr"], g["col"], g["val"], 3)
    assert np.array_equal(d, g["diag"])


def test_fdm_laplace_generator_matches_reference(vh):
    g = load("fdm_laplace_12x9")
    rp, col, val = vh.fdm_laplace_2d(int(g["px"]), int(g["py"]))
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(col, g["col"]) and np.array_equal(val, g["val"])


def test_pipelined_ops_match_reference(oracle):
    g = load("pipelined_ops_25x17")
    o = oracle.orc()
    p = oracle.p
    rp, col, val = oracle.cu32(g["row_ptr"]), oracle.cu32(g["col"]), oracle.c64(g["val"])
    G = [np.ascontiguousarray(v) for v in g["g"]]
    n = G[0].size
    x, pv, r, Ap = G[0].copy(), G[1].copy(), G[2].copy(), G[3].copy()
    buf = np.zeros(768)
    o.orc_pipelined_cg_vector_update_f64(n, p(x), 0.37, p(pv), p(r), p(Ap), -0.61, p(buf))
    assert np.array_equal(x, g["cgvu_x"]) and np.array_equal(pv, g["cgvu_p"]) and np.array_equal(r, g["cgvu_r"])
    assert np.array_equal(buf, g["cgvu_buf"])
    Ap2, buf2 = np.zeros(n), np.zeros(768)
    o.orc_pipelined_prod_csr_f64(n, p(rp), p(col), p(val), p(G[1]), p(Ap2), None, p(buf2), 256, 0)
    assert np.array_equal(Ap2, g["cgprod_Ap"]) and np.array_equal(buf2, g["cgprod_buf"])
    Ap3, buf3 = np.zeros(n), np.zeros(1536)
    o.orc_pipelined_prod_csr_f64(n, p(rp), p(col), p(val), p(G[1]), p(Ap3), p(G[4]), p(buf3), 256, 768)
    assert np.array_equal(Ap3, g["bprod_Ap"]) and np.array_equal(buf3, g["bprod_buf"])
    s, buf4 = np.zeros(n), np.zeros(1536)
    buf4[0], buf4[768] = 3.5, -1.25
    o.orc_pipelined_bicgstab_update_s_f64(n, p(s), p(G[2]), p(G[3]), p(buf4), 256, 1280)
    assert np.array_equal(s, g["upds_s"]) and np.array_equal(buf4, g["upds_buf"])
    x5, p5, r5, buf5 = G[0].copy(), G[1].copy(), G[2].copy(), np.zeros(1536)
    o.orc_pipelined_bicgstab_vector_update_f64(n, p(x5), 0.3, p(p5), 0.7, p(G[5]), p(r5), p(G[6]), -0.2, p(G[7]),
                                               p(G[4]), p(buf5))
    assert np.array_equal(x5, g["bvu_x"]) and np.array_equal(p5, g["bvu_p"]) and np.array_equal(r5, g["bvu_r"])
    assert np.array_equal(buf5, g["bvu_buf"])


def test_solvers_match_reference_lap2d(oracle):
    g = load("solve_lap2d_40x40")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    x, it, err, _ = oracle.cg_pipelined(A, g["b"], 1e-8, 2000)
    assert it == g["cg_iters"] and err == g["cg_error"] and np.array_equal(x, g["cg_x"])
    x, it, err, _ = oracle.cg_pipelined(A, g["b"], 1e-8, 40)
    assert it == g["cg40_iters"] == 40 and err == g["cg40_error"] and np.array_equal(x, g["cg40_x"])
    x, it, err, _ = oracle.bicgstab_pipelined(A, g["b"], 1e-8, 2000)
    assert it == g["bicgstab_iters"] and err == g["bicgstab_error"] and np.array_equal(x, g["bicgstab_x"])
    x, it, err, _ = oracle.mixed_cg(A, g["b"], 1e-8, 2000)
    assert it == g["mixed_iters"] and err == g["mixed_error"] and np.array_equal(x, g["mixed_x"])


def test_preconditioned_solvers_match_reference(oracle):
    g = load("solve_diffusion2d_30x30")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    d = oracle.row_info(g["row_ptr"], g["col"], g["val"], 3)
    x, it, err, _ = oracle.cg_classic(A, g["b"], 1e-10, 3000, diag=d)
    assert it == g["jcg_iters"] and err == g["jcg_error"] and np.array_equal(x, g["jcg_x"])
    x, it, err, _ = oracle.bicgstab_classic(A, g["b"], 1e-10, 3000, diag=d, max_before_restart=int(g["restart_every"]))
    assert it == g["jbicgstab_iters"] and err == g["jbicgstab_error"] and np.array_equal(x, g["jbicgstab_x"])


def test_sell_solvers_match_reference(oracle):
    g = load("solve_sell_lap2d_37x29")
    S = oracle.SellHost(g["row_ptr"], g["col"], g["val"], int(g["C"]))
    x, it, err, _ = oracle.cg_pipelined(S, g["b"], 1e-8, 2000)
    assert it == g["cg_iters"] and err == g["cg_error"] and np.array_equal(x, g["cg_x"])
    x, it, err, _ = oracle.bicgstab_pipelined(S, g["b"], 1e-8, 2000)
    assert it == g["bicgstab_iters"] and err == g["bicgstab_error"] and np.array_equal(x, g["bicgstab_x"])


def test_nonsymmetric_bicgstab_matches_reference(oracle):
    g = load("solve_powerlaw_2000")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    x, it, err, _ = oracle.bicgstab_pipelined(A, g["b"], 1e-10, 400)
    assert it == g["bicgstab_iters"] and err == g["bicgstab_error"] and np.array_equal(x, g["bicgstab_x"])


def test_gmres_ops_match_reference(oracle):
    """pipelined_gmres_* (host_based/iterative_operations.hpp:734-937), single-threaded."""
    g = load("gmres_ops_1001")
    n, ld, K, k, chunk = (int(g[q]) for q in ("n", "ld", "K", "k", "chunk"))
    basis = np.ascontiguousarray(g["basis"])
    s1 = np.zeros(K * chunk)
    oracle.gmres_stage1(basis, n, ld, k, s1, chunk, skip_tail=True)
    assert np.array_equal(s1, g["stage1_out"])
    b2, R2, buf2 = basis.copy(), np.zeros(K * K), np.zeros(3 * chunk)
    oracle.gmres_stage2(b2, n, ld, k, np.ascontiguousarray(g["vi_in_vk"]), R2, K, buf2, chunk)
    assert np.array_equal(b2, g["stage2_basis"]) and np.array_equal(R2, g["stage2_R"])
    assert np.array_equal(buf2, g["stage2_buf"])
    vk = basis[k * ld:k * ld + n].copy()
    R3, rdv = np.zeros(K * K), np.zeros(K * chunk)
    oracle.gmres_normalize_vk(vk, np.ascontiguousarray(g["residual"]), R3, k * K + k,
                              np.ascontiguousarray(g["normalize_buf"]), rdv, chunk, k * chunk)
    assert np.array_equal(vk, g["normalize_vk"]) and np.array_equal(R3, g["normalize_R"])
    assert np.array_equal(rdv, g["normalize_rdv"])
    x = np.ascontiguousarray(g["update_x0"]).copy()
    oracle.gmres_update_result(x, np.ascontiguousarray(g["residual"]), basis, n, ld,
                               np.ascontiguousarray(g["update_coef"]), K - 1)
    assert np.array_equal(x, g["update_x"])


@pytest.mark.parametrize("case", ["convdiff_20x15x3", "lap2d_40x40"])
def test_gmres_matches_reference(oracle, case):
    """solve(A, b, gmres_tag) (gmres.hpp:182-370), single-threaded: x, tag.iters(), tag.error()."""
    g = load(f"solve_gmres_{case}")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    for i, (tol, mi, kd) in enumerate(g["cases"]):
        x, it, err, _ = oracle.gmres(A, g["b"], tol, int(mi), int(kd), gs_skip_tail=True)
        assert it == g["iters"][i] and err == g["error"][i] and np.array_equal(x, g["x"][i]), (tol, mi, kd)
        lo, hi = g["env"][i]
        assert lo <= it <= hi


FORMATS = (("coo", 2), ("ell", 3), ("hyb", 4))


def fixture_arrays(g, tag):
    return {k[len(tag) + 1:]: g[k] for k in g if k.startswith(tag + "_") and not k.startswith(tag + "_y_")
            and "prod" not in k}


@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_format_arrays_spmv_and_products_match_reference(oracle, vh, tag, fmt):
    """copy() into coordinate_matrix / ell_matrix / hyb_matrix, prod_impl and the pipelined products
    (host_based/sparse_matrix_operations.hpp:1221-1929, iterative_operations.hpp:113-365)."""
    g = load("formats_powerlaw_700")
    rp, col, val = g["row_ptr"], g["col"], g["val"]
    arrays = fixture_arrays(g, tag)
    # the library's host conversions build the reference's arrays
    if fmt == 3:
        ld, w, c, e = vh.ell_from_csr(rp, col, val)
        assert (ld, w) == (int(arrays["ld"]), int(arrays["width"]))
        assert np.array_equal(c, arrays["coords"]) and np.array_equal(e, arrays["elements"])
    elif fmt == 4:
        ld, w, ec, ee, cr, cc, ce = vh.hyb_from_csr(rp, col, val)
        assert (ld, w) == (int(arrays["ld"]), int(arrays["width"]))
        assert np.array_equal(ec, arrays["coords"]) and np.array_equal(ee, arrays["elements"])
        assert np.array_equal(cr, arrays["csr_rows"]) and np.array_equal(cc, arrays["csr_cols"])
        assert np.array_equal(ce, arrays["csr_elements"])
    else:
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp.astype(np.int64)))
        assert np.array_equal(arrays["coords"][0::2], rows) and np.array_equal(arrays["coords"][1::2], col)
    A = oracle.FmtHost(fmt, rp, col, val, arrays)
    for at, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
        assert np.array_equal(oracle.fmt_spmv(A, g["x"], alpha, g["y0"], beta), g[f"{tag}_y_{at}"]), at
    Ap, buf = oracle.fmt_pipelined_prod(A, g["x"])
    assert np.array_equal(Ap, g[f"{tag}_cgprod_Ap"]) and np.array_equal(buf[:768], g[f"{tag}_cgprod_buf"])
    Ap, buf = oracle.fmt_pipelined_prod(A, g["x"], r0=g["r0"], offset=768)
    assert np.array_equal(Ap, g[f"{tag}_bprod_Ap"]) and np.array_equal(buf, g[f"{tag}_bprod_buf"])


def library_format_arrays(vh, fmt, rp, col, val):
    """The format's arrays as the library's host conversions build them."""
    if fmt == 2:
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp.astype(np.int64))).astype(np.uint32)
        coords = np.empty(2 * col.size, np.uint32)
        coords[0::2], coords[1::2] = rows, col
        return {"coords": coords, "elements": np.asarray(val, np.float64)}
    if fmt == 3:
        ld, w, c, e = vh.ell_from_csr(rp, col, val)
        return {"ld": ld, "width": w, "coords": c, "elements": e}
    ld, w, ec, ee, cr, cc, ce = vh.hyb_from_csr(rp, col, val)
    return {"ld": ld, "width": w, "coords": ec, "elements": ee, "csr_rows": cr, "csr_cols": cc, "csr_elements": ce}


@pytest.mark.parametrize("case", ["lap2d_30x22", "convdiff_12x9x4", "convdiff_12x9x4_gmres"])
def test_format_solvers_match_reference(oracle, vh, case):
    """solve() with coordinate_matrix / ell_matrix / hyb_matrix (cg.hpp:204-254, bicgstab.hpp:231-285,
    gmres.hpp:384-432): the oracle's pipelined solvers on the format reproduce x, iterations and error.
    GMRES on these formats is the CSR restatement (the products sum each row in the same order)."""
    g = load(f"solve_formats_{case}")
    rp, col, val, b = g["row_ptr"], g["col"], g["val"], g["b"]
    kind, tol, mi = int(g["kind"]), float(g["tol"]), int(g["max_iters"])
    for tag, fmt in FORMATS:
        if kind == 2:
            x, it, err, _ = oracle.gmres(oracle.CsrHost(rp, col, val), b, tol, mi, 20, gs_skip_tail=True)
        else:
            A = oracle.FmtHost(fmt, rp, col, val, library_format_arrays(vh, fmt, rp, col, val))
            run = oracle.cg_pipelined_fmt if kind == 0 else oracle.bicgstab_pipelined_fmt
            x, it, err = run(A, b, tol, mi)
        assert it == g[f"{tag}_iters"] and err == g[f"{tag}_error"] and np.array_equal(x, g[f"{tag}_x"]), tag


needs_ref = pytest.mark.skipif(not (pathlib.Path(__file__).resolve().parents[1] / "oracle/_ref/libvclref.so").exists(),
                               reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_equals_reference_build_on_fresh_inputs(oracle, vh, seed):
    """Fresh seeded inputs beyond the fixtures, straight against the reference build."""
    rp, col, val = vh.powerlaw(1500, max_len=100, diag_factor=1.3, seed=seed)
    b = vh.random_uniform(seed, rp.size - 1)
    A = oracle.CsrHost(rp, col, val)
    x, it, err, _ = oracle.bicgstab_pipelined(A, b, 1e-9, 300)
    xr, itr, errr = oracle.ref_bicgstab(rp, col, val, b, 1e-9, 300)
    assert (it, err) == (itr, errr) and np.array_equal(x, xr)
    k = np.exp(vh.random_normal(seed, 12 * 11 * 7, 0.0, 1.0))
    rp, col, val = vh.diffusion(12, 11, 7, k)
    b = vh.random_uniform(seed, rp.size - 1)
    A = oracle.CsrHost(rp, col, val)
    x, it, err, _ = oracle.cg_pipelined(A, b, 1e-10, 3000)
    xr, itr, errr = oracle.ref_cg(rp, col, val, b, 1e-10, 3000)
    assert (it, err) == (itr, errr) and np.array_equal(x, xr)
    x, it, err, _ = oracle.mixed_cg(A, b, 1e-10, 3000)
    xr, itr, errr = oracle.ref_mixed_cg(rp, col, val, b, 1e-10, 3000)
    assert (it, err) == (itr, errr) and np.array_equal(x, xr)
    d = oracle.row_info(rp, col, val, 3)
    x, it, err, _ = oracle.bicgstab_classic(A, b, 1e-10, 3000, diag=d, max_before_restart=15)
    xr, itr, errr = oracle.ref_bicgstab(rp, col, val, b, 1e-10, 3000, jacobi=True, max_before_restart=15)
    assert (it, err) == (itr, errr) and np.array_equal(x, xr)
    rp, col, val = vh.convection_diffusion(13, 9, 4, beta=(0.2 * seed, 0.1, 0.3))
    b = vh.random_uniform(seed, rp.size - 1)
    A = oracle.CsrHost(rp, col, val)
    for tol, mi, kd in ((1e-9, 400, 25), (1e-7, 90, 7)):
        x, it, err, _ = oracle.gmres(A, b, tol, mi, kd, gs_skip_tail=True)
        xr, itr, errr = oracle.ref_gmres(rp, col, val, b, tol, mi, kd)
        assert (it, err) == (itr, errr) and np.array_equal(x, xr)


def test_blocked_reduction_orders(oracle, vh):
    """k-block reduction orders: k = 0 is the reference order (bit-exact with the fixture); other k
    change only rounding, so CG moves by at most a couple of iterations while the restarted
    BiCGStab of this fixture spreads much wider (which is why its GPU bar is an envelope)."""
    g = load("solve_lap2d_40x40")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    lo, hi = oracle.iteration_envelope(lambda: oracle.cg_pipelined(A, g["b"], 1e-8, 2000))
    assert lo <= int(g["cg_iters"]) <= hi and hi - lo <= 2
    with oracle.reduction_blocks(0):
        assert oracle.cg_pipelined(A, g["b"], 1e-8, 2000)[1] == g["cg_iters"]
    with oracle.reduction_blocks(1):  # one block == sequential order
        x, it, err, _ = oracle.cg_pipelined(A, g["b"], 1e-8, 2000)
        assert it == g["cg_iters"] and err == g["cg_error"] and np.array_equal(x, g["cg_x"])
    g = load("solve_diffusion2d_30x30")
    A = oracle.CsrHost(g["row_ptr"], g["col"], g["val"])
    d = oracle.row_info(g["row_ptr"], g["col"], g["val"], 3)
    lo, hi = oracle.iteration_envelope(lambda: oracle.bicgstab_classic(A, g["b"], 1e-10, 3000, diag=d,
                                                                        max_before_restart=20))
    assert lo <= int(g["jbicgstab_iters"]) <= hi


@needs_ref
@pytest.mark.parametrize("tag,fmt", FORMATS)
def test_formats_equal_reference_build_on_fresh_inputs(oracle, vh, tag, fmt):
    rp, col, val = vh.convection_diffusion(11, 6, 5, beta=(0.7, 0.1, 0.4))
    n = rp.size - 1
    x, y0 = vh.random_uniform(41, n), vh.random_uniform(42, n)
    ref_arrays = oracle.ref_format_arrays(fmt, rp, col, val)
    mine = library_format_arrays(vh, fmt, rp, col, val)
    for k in ref_arrays:
        assert np.array_equal(np.asarray(ref_arrays[k]), np.asarray(mine[k])), k
    A = oracle.FmtHost(fmt, rp, col, val, mine)
    y = y0.copy()
    oracle.ref().ref_fmt_spmv_f64(fmt, n, oracle.p(oracle.cu32(rp)), oracle.p(oracle.cu32(col)),
                                  oracle.p(oracle.c64(val)), oracle.p(x), 1.5, oracle.p(y), -0.5)
    assert np.array_equal(oracle.fmt_spmv(A, x, 1.5, y0, -0.5), y)
    b = vh.random_uniform(43, n)
    xo, io, eo = oracle.bicgstab_pipelined_fmt(A, b, 1e-10, 500)
    xr, ir, er = oracle.ref_fmt_solve(fmt, 1, rp, col, val, b, 1e-10, 500)
    assert (io, eo) == (ir, er) and np.array_equal(xo, xr)
