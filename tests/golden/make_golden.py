"""Writes the golden fixtures of tests/golden/*.npz from the REAL ViennaCL 1.7.1 host backend.

The reference is compiled in place by oracle/Makefile (target `ref`, output oracle/_ref/) and
called single-threaded through its public API (oracle/ref_shim.cpp).  Each fixture stores the
inputs and the reference's outputs, so the tests pin the oracle and the HIP path without the
reference being present.  The reference's own SpMV test matrix (examples/testdata/mat65k.mtx) is
not available, so these seeded problems stand in for it.

Run from the repository root:  python tests/golden/make_golden.py  (add `gmres`, `formats` or
`c2_envelope` to write only those fixtures)
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "viennacl-dev_amd"))

import pyoracle as po  # noqa: E402
import vclhip as vh  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent


def save(name, **arrays):
    np.savez_compressed(OUT / f"{name}.npz", **arrays)
    print("wrote", name, {k: getattr(v, "shape", v) for k, v in arrays.items()})


def with_empty_rows(rp, col, val, rows):
    """Removes all entries of the given rows (empty CSR rows)."""
    keep = np.ones(col.size, bool)
    for r in rows:
        keep[rp[r]:rp[r + 1]] = False
    lens = np.diff(rp.astype(np.int64))
    for r in rows:
        lens[r] = 0
    nrp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    return nrp, col[keep], val[keep]


def main():
    po.set_threads(1)
    R = po.ref()
    p = po.p

    # generate_fdm_laplace (tools/matrix_generation.hpp:47-86)
    px, py = 12, 9
    n = px * py
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(5 * n, np.uint32)
    val = np.zeros(5 * n)
    nnz = R.ref_fdm_laplace(px, py, p(rp), p(col), p(val))
    save("fdm_laplace_12x9", px=px, py=py, row_ptr=rp, col=col[:nnz], val=val[:nnz])

    # SpMV y = alpha A x + beta y (linalg::prod_impl)
    cases = {
        "lap2d_30x20": vh.fdm_laplace_2d(30, 20),
        "powerlaw_3000": vh.powerlaw(3000, max_len=300, seed=5),
    }
    rp0, col0, val0 = vh.fdm_laplace_2d(20, 20)
    cases["empty_rows_400"] = with_empty_rows(rp0, col0, val0, [0, 7, 133, 399])
    for name, (rp, col, val) in cases.items():
        n = rp.size - 1
        x = vh.random_uniform(11, n)
        y0 = vh.random_uniform(12, n)
        outs = {}
        for tag, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
            y = y0.copy()
            R.ref_csr_spmv_f64(n, p(rp), p(col), p(val), p(x), alpha, p(y), beta)
            outs[f"y_{tag}"] = y
        save(f"spmv_{name}", row_ptr=rp, col=col, val=val, x=x, y0=y0, **outs)

    # SELL-C arrays and SpMV (sliced_ell_matrix.hpp:140-214, host_based :1796-1868); n % C != 0
    for name, C_, (rp, col, val) in (("lap2d_37x29_c32", 32, vh.fdm_laplace_2d(37, 29)),
                                     ("powerlaw_1001_c64", 64, vh.powerlaw(1001, max_len=200, seed=6))):
        n = rp.size - 1
        ns = (n + C_ - 1) // C_
        S = po.SellHost(rp, col, val, C_)
        stored = S.el.size
        cpb = np.zeros(ns, np.uint32)
        bs = np.zeros(ns, np.uint32)
        ci = np.zeros(stored, np.uint32)
        el = np.zeros(stored)
        R.ref_sell_arrays(n, C_, p(rp), p(col), p(val), p(cpb), p(ci), p(bs), p(el), stored)
        x = vh.random_uniform(13, n)
        y0 = vh.random_uniform(14, n)
        y = y0.copy()
        R.ref_sell_spmv_f64(n, C_, p(rp), p(col), p(val), p(x), 1.5, p(y), -0.25)
        save(f"sell_{name}", C=C_, row_ptr=rp, col=col, val=val, columns_per_block=cpb, column_indices=ci,
             block_start=bs, elements=el, x=x, y0=y0, y=y, alpha=1.5, beta=-0.25)

    # row blocks (compressed_matrix.hpp:1152-1188) and the Jacobi diagonal (row_info)
    rp, col, val = vh.powerlaw(5000, max_len=1500, seed=7)
    n = rp.size - 1
    out = np.zeros(n + 1, np.uint32)
    nb = R.ref_row_blocks(n, p(rp), p(col), p(val), p(out))
    d = np.zeros(n)
    R.ref_row_info_diag(n, p(rp), p(col), p(val), p(d))
    save("row_blocks_powerlaw_5000", row_ptr=rp, col=col, val=val, row_blocks=out[: nb + 1], diag=d)

    # pipelined kernels (iterative_operations.hpp), one call each
    rp, col, val = vh.fdm_laplace_2d(25, 17)
    n = rp.size - 1
    g = [vh.random_uniform(20 + i, n) for i in range(8)]
    x, pv, r, Ap = g[0].copy(), g[1].copy(), g[2].copy(), g[3].copy()
    buf = np.zeros(768)
    R.ref_pipelined_cg_vector_update(n, p(x), 0.37, p(pv), p(r), p(Ap), -0.61, p(buf))
    Ap2 = np.zeros(n)
    buf2 = np.zeros(768)
    R.ref_pipelined_cg_prod_csr(n, p(rp), p(col), p(val), p(g[1]), p(Ap2), p(buf2))
    Ap3 = np.zeros(n)
    buf3 = np.zeros(1536)
    R.ref_pipelined_bicgstab_prod_csr(n, p(rp), p(col), p(val), p(g[1]), p(Ap3), p(g[4]), p(buf3), 768)
    s = np.zeros(n)
    buf4 = np.zeros(1536)
    buf4[0] = 3.5
    buf4[768] = -1.25
    R.ref_pipelined_bicgstab_update_s(n, p(s), p(g[2]), p(g[3]), p(buf4), 1280)
    x5, p5, r5 = g[0].copy(), g[1].copy(), g[2].copy()
    buf5 = np.zeros(1536)
    R.ref_pipelined_bicgstab_vector_update(n, p(x5), 0.3, p(p5), 0.7, p(g[5]), p(r5), p(g[6]), -0.2, p(g[7]),
                                           p(g[4]), p(buf5))
    save("pipelined_ops_25x17", row_ptr=rp, col=col, val=val, g=np.stack(g),
         cgvu_x=x, cgvu_p=pv, cgvu_r=r, cgvu_buf=buf, cgprod_Ap=Ap2, cgprod_buf=buf2,
         bprod_Ap=Ap3, bprod_buf=buf3, upds_s=s, upds_buf=buf4, bvu_x=x5, bvu_p=p5, bvu_r=r5, bvu_buf=buf5)

    # solvers through linalg::solve.  Besides the single-threaded run, each case records the
    # envelope [min, max] of the reference's iteration count over OpenMP thread counts 1..8: the
    # reference's own spread when only the summation order of its inner products changes.
    def solver_case(name, rp, col, val, b, **meta):
        save(name, row_ptr=rp, col=col, val=val, b=b, **meta)

    def envelope(fn):
        its = []
        for t in range(1, 9):
            po.set_threads(t)
            its.append(fn()[1])
        po.set_threads(1)
        return np.array([min(its), max(its)], np.int64)

    rp, col, val = vh.fdm_laplace_2d(40, 40)
    b = vh.random_uniform(0, rp.size - 1)
    x, it, err = po.ref_cg(rp, col, val, b, 1e-8, 2000)
    xb, itb, errb = po.ref_bicgstab(rp, col, val, b, 1e-8, 2000)
    xm, itm, errm = po.ref_mixed_cg(rp, col, val, b, 1e-8, 2000)
    x2, it2, err2 = po.ref_cg(rp, col, val, b, 1e-8, 40)  # stops at max_iters
    solver_case("solve_lap2d_40x40", rp, col, val, b, cg_x=x, cg_iters=it, cg_error=err,
                cg_env=envelope(lambda: po.ref_cg(rp, col, val, b, 1e-8, 2000)),
                bicgstab_x=xb, bicgstab_iters=itb, bicgstab_error=errb,
                bicgstab_env=envelope(lambda: po.ref_bicgstab(rp, col, val, b, 1e-8, 2000)),
                mixed_x=xm, mixed_iters=itm, mixed_error=errm,
                mixed_env=envelope(lambda: po.ref_mixed_cg(rp, col, val, b, 1e-8, 2000)),
                cg40_x=x2, cg40_iters=it2, cg40_error=err2)

    k = np.exp(vh.random_normal(1, 30 * 30, 0.0, 1.0))
    rp, col, val = vh.diffusion(30, 30, 1, k)
    b = vh.random_uniform(0, rp.size - 1)
    x, it, err = po.ref_cg(rp, col, val, b, 1e-10, 3000, jacobi=True)
    xb, itb, errb = po.ref_bicgstab(rp, col, val, b, 1e-10, 3000, jacobi=True, max_before_restart=20)
    solver_case("solve_diffusion2d_30x30", rp, col, val, b, jcg_x=x, jcg_iters=it, jcg_error=err,
                jcg_env=envelope(lambda: po.ref_cg(rp, col, val, b, 1e-10, 3000, jacobi=True)),
                jbicgstab_x=xb, jbicgstab_iters=itb, jbicgstab_error=errb, restart_every=20,
                jbicgstab_env=envelope(lambda: po.ref_bicgstab(rp, col, val, b, 1e-10, 3000, jacobi=True,
                                                               max_before_restart=20)))

    rp, col, val = vh.fdm_laplace_2d(37, 29)
    b = vh.random_uniform(0, rp.size - 1)
    x, it, err = po.ref_cg_sell(rp, col, val, 32, b, 1e-8, 2000)
    xb, itb, errb = po.ref_bicgstab_sell(rp, col, val, 32, b, 1e-8, 2000)
    solver_case("solve_sell_lap2d_37x29", rp, col, val, b, C=32, cg_x=x, cg_iters=it, cg_error=err,
                cg_env=envelope(lambda: po.ref_cg_sell(rp, col, val, 32, b, 1e-8, 2000)),
                bicgstab_x=xb, bicgstab_iters=itb, bicgstab_error=errb,
                bicgstab_env=envelope(lambda: po.ref_bicgstab_sell(rp, col, val, 32, b, 1e-8, 2000)))

    rp, col, val = vh.powerlaw(2000, max_len=512, seed=1)
    b = vh.random_uniform(0, rp.size - 1)
    xb, itb, errb = po.ref_bicgstab(rp, col, val, b, 1e-10, 400)
    solver_case("solve_powerlaw_2000", rp, col, val, b, bicgstab_x=xb, bicgstab_iters=itb, bicgstab_error=errb)


def gmres_main():
    """Pipelined GMRES (gmres.hpp:182-370) and its operations (iterative_operations.hpp:247-418)."""
    po.set_threads(1)
    R = po.ref()
    p = po.p

    # the four operations on a random basis; n = 1001 with ld = 1024, K = 6, k = 3
    rng = np.random.default_rng(11)
    n, ld, K, k, chunk = 1001, 1024, 6, 3, 128
    basis = rng.standard_normal(K * ld)
    residual = rng.standard_normal(n)
    s1 = np.zeros(K * chunk)
    R.ref_pipelined_gmres_stage1(p(basis), n, ld, K, k, p(s1), chunk)
    vi = rng.standard_normal(K * chunk)
    b2, R2, buf2 = basis.copy(), np.zeros(K * K), np.zeros(3 * chunk)
    R.ref_pipelined_gmres_stage2(p(b2), n, ld, K, k, p(vi), p(R2), p(buf2), chunk)
    vk = basis[k * ld:k * ld + n].copy()
    R3, buf3, rdv = np.zeros(K * K), np.abs(rng.standard_normal(3 * chunk)), np.zeros(K * chunk)
    buf3_in = buf3.copy()
    R.ref_pipelined_gmres_normalize_vk(n, p(vk), p(residual), p(R3), K * K, k * K + k, p(buf3), p(rdv), K * chunk,
                                       chunk, k * chunk)
    x4, coef = rng.standard_normal(n), rng.standard_normal(K)
    x4_in = x4.copy()
    R.ref_pipelined_gmres_update_result(p(x4), p(residual), p(basis), n, ld, K, p(coef), K - 1)
    save("gmres_ops_1001", n=n, ld=ld, K=K, k=k, chunk=chunk, basis=basis, residual=residual,
         stage1_out=s1, vi_in_vk=vi, stage2_basis=b2, stage2_R=R2, stage2_buf=buf2,
         normalize_buf=buf3_in, normalize_vk=vk, normalize_R=R3, normalize_rdv=rdv,
         update_x0=x4_in, update_coef=coef, update_x=x4)

    # solves: nonsymmetric convection-diffusion and the 2-D Laplacian.  Cases (tol, max_iters,
    # krylov_dim) include a run that stops at max_iters and one with max_iters not a multiple of
    # krylov_dim.
    for name, (rp, col, val) in (("convdiff_20x15x3", vh.convection_diffusion(20, 15, 3)),
                                 ("lap2d_40x40", vh.fdm_laplace_2d(40, 40))):
        nn = rp.size - 1
        b = vh.random_uniform(5, nn)
        cases = np.array([(1e-8, 300, 20), (1e-10, 500, 30), (1e-6, 100, 10), (1e-12, 45, 20), (1e-8, 60, 1)])
        xs, its, errs, env = [], [], [], []
        for tol, mi, kd in cases:
            x, it, err = po.ref_gmres(rp, col, val, b, tol, int(mi), int(kd))
            xs.append(x)
            its.append(it)
            errs.append(err)
            e = []
            for t in range(1, 9):
                po.set_threads(t)
                e.append(po.ref_gmres(rp, col, val, b, tol, int(mi), int(kd))[1])
            po.set_threads(1)
            env.append((min(e), max(e)))
        save(f"solve_gmres_{name}", row_ptr=rp, col=col, val=val, b=b, cases=cases, x=np.stack(xs),
             iters=np.array(its, np.int64), error=np.array(errs), env=np.array(env, np.int64))


def formats_main():
    """coordinate_matrix / ell_matrix / hyb_matrix: the arrays copy() writes, prod_impl, the fused
    pipelined products and the solvers (iterative_operations.hpp:113-365; cg.hpp:204-254,
    bicgstab.hpp:231-285, gmres.hpp:384-432)."""
    po.set_threads(1)
    R = po.ref()
    p = po.p
    rp, col, val = vh.powerlaw(700, max_len=60, seed=4)
    # rows 0, 5 and 699 empty
    keep = np.ones(col.size, bool)
    for r in (0, 5, 699):
        keep[rp[r]:rp[r + 1]] = False
    lens = np.diff(rp.astype(np.int64))
    lens[[0, 5, 699]] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    col, val = col[keep], val[keep]
    n = rp.size - 1
    x, y0, r0 = vh.random_uniform(31, n), vh.random_uniform(32, n), vh.random_uniform(33, n)
    out = {"row_ptr": rp, "col": col, "val": val, "x": x, "y0": y0, "r0": r0}
    for fmt, tag in ((po.FORMAT_COO, "coo"), (po.FORMAT_ELL, "ell"), (po.FORMAT_HYB, "hyb")):
        for k, v in po.ref_format_arrays(fmt, rp, col, val).items():
            out[f"{tag}_{k}"] = v
        for at, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
            y = y0.copy()
            R.ref_fmt_spmv_f64(fmt, n, p(rp), p(col), p(val), p(x), alpha, p(y), beta)
            out[f"{tag}_y_{at}"] = y
        Ap, buf = np.zeros(n), np.zeros(768)
        R.ref_fmt_pipelined_prod(fmt, n, p(rp), p(col), p(val), p(x), p(Ap), None, p(buf), 0)
        out[f"{tag}_cgprod_Ap"], out[f"{tag}_cgprod_buf"] = Ap, buf
        Ap, buf = np.zeros(n), np.zeros(1536)
        R.ref_fmt_pipelined_prod(fmt, n, p(rp), p(col), p(val), p(x), p(Ap), p(r0), p(buf), 768)
        out[f"{tag}_bprod_Ap"], out[f"{tag}_bprod_buf"] = Ap, buf
    save("formats_powerlaw_700", **out)

    # solvers through linalg::solve for each format: CG on the 2-D Laplacian, BiCGStab and GMRES on
    # convection-diffusion
    for name, (rp, col, val), kind, tol, mi in (("lap2d_30x22", vh.fdm_laplace_2d(30, 22), 0, 1e-9, 2000),
                                                ("convdiff_12x9x4", vh.convection_diffusion(12, 9, 4), 1, 1e-9, 2000),
                                                ("convdiff_12x9x4_gmres", vh.convection_diffusion(12, 9, 4), 2, 1e-9,
                                                 400)):
        n = rp.size - 1
        b = vh.random_uniform(7, n)
        out = {"row_ptr": rp, "col": col, "val": val, "b": b, "kind": kind, "tol": tol, "max_iters": mi}
        for fmt, tag in ((po.FORMAT_COO, "coo"), (po.FORMAT_ELL, "ell"), (po.FORMAT_HYB, "hyb")):
            x, it, err = po.ref_fmt_solve(fmt, kind, rp, col, val, b, tol, mi, 20)
            out[f"{tag}_x"], out[f"{tag}_iters"], out[f"{tag}_error"] = x, it, err
        save(f"solve_formats_{name}", **out)


def c2_envelope_main():
    """BiCGStab iteration counts on the 128^3 Laplacian (b = U[-1,1] seed 0, tol 1e-10), plain and
    Jacobi: the reference's own solve() at 1, 2, 4, 8 and 16 OpenMP threads, and the oracle's
    restatement over its k-block reduction orders.  Only the summation order of the inner products
    differs between these runs; the spread of the counts is the reference's own."""
    rp, col, val = vh.laplace_3d(128)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    d = np.full(n, 6.0)
    out = {"threads": np.array([1, 2, 4, 8, 16], np.int32), "oracle_blocks": np.array(po.ENVELOPE_BLOCKS, np.int32)}
    ref_p, ref_j = [], []
    for t in out["threads"]:
        po.set_threads(int(t))
        ref_p.append(po.ref_bicgstab(rp, col, val, b, 1e-10, 5000)[1])
        ref_j.append(po.ref_bicgstab(rp, col, val, b, 1e-10, 5000, jacobi=True)[1])
        print("threads", t, ref_p[-1], ref_j[-1], flush=True)
    po.set_threads(8)
    A = po.CsrHost(rp, col, val)
    orc_p, orc_j = [], []
    for k in po.ENVELOPE_BLOCKS:
        with po.reduction_blocks(k):
            orc_p.append(po.bicgstab_pipelined(A, b, 1e-10, 5000)[1])
            orc_j.append(po.bicgstab_classic(A, b, 1e-10, 5000, diag=d)[1])
        print("blocks", k, orc_p[-1], orc_j[-1], flush=True)
    po.set_threads(1)
    save("c2_bicgstab_lap3d_128_iters", ref_bicgstab=np.array(ref_p, np.int32),
         ref_jacobi_bicgstab=np.array(ref_j, np.int32), oracle_bicgstab=np.array(orc_p, np.int32),
         oracle_jacobi_bicgstab=np.array(orc_j, np.int32), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["gmres"]:
        gmres_main()
    elif sys.argv[1:] == ["formats"]:
        formats_main()
    elif sys.argv[1:] == ["c2_envelope"]:
        c2_envelope_main()
    else:
        main()
        gmres_main()
        formats_main()
        c2_envelope_main()
