"""GPU parity of the batched CG (vclhip_batched_cg_solve_*, persistent kernel).

Each system follows the pipelined CG of cg.hpp:128-187; the oracle runs the same recurrences per
system (orc_batched_cg2_csr, pinned by dense solves in test_oracle_math.py).  Bars: per-system
iteration count within 2% (at least 2) of the oracle's, the recursive residual below the tolerance,
the FP64 true residual of every system verified on the host, and x close to the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def iters_close(got, want):
    got = np.asarray(got, np.int64)
    want = np.asarray(want, np.int64)
    return np.abs(got - want) <= np.maximum(2, np.ceil(0.02 * want))


def per_system_relres(oracle, sys_ptr, rp, col, val, x, b):
    r = b - oracle.csr_spmv(rp, col, val, x)
    out = []
    for s in range(sys_ptr.size - 1):
        a, e = sys_ptr[s], sys_ptr[s + 1]
        out.append(np.linalg.norm(r[a:e]) / np.linalg.norm(b[a:e]))
    return np.array(out)


def small_batch(vh, ns=48, n_max=1500, seed=3):
    return vh.graph_laplacian_batch(ns, 16, n_max, 8.0, 0.1, 1.0, 0.01, seed)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("staged", [True, False])
def test_batched_vs_oracle(dev, vh, oracle, single, staged):
    """staged: the plan has the systems' row blocks (the staged SpMV, the untuned default); without
    the row pointers at plan time every system takes the lanes-per-row SpMV."""
    sys_ptr, rp, col, val = small_batch(vh)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    dtype = np.float32 if single else np.float64
    tol = 1e-5 if single else 1e-8
    A = vh.CsrMatrix(rp, col, val, dtype=dtype)
    plan = vh.BatchPlan(sys_ptr, rp if staged else None)
    db = vh.DeviceArray.from_numpy(b.astype(dtype))
    dx = vh.DeviceArray.zeros(n, dtype)
    it, err, conv = plan.solve(A, db, dx, tol=tol, max_iters=3000)
    xo, ito, erro = oracle.batched_cg(sys_ptr, oracle.CsrHost(rp, col, val), b, tol, 3000, single=single)
    assert (conv == 1).all()
    assert iters_close(it, ito).all(), np.c_[it, ito][~iters_close(it, ito)]
    assert (err < tol).all()
    x = dx.to_numpy().astype(np.float64)
    rr = per_system_relres(oracle, sys_ptr, rp, col, val, x, b.astype(dtype).astype(np.float64))
    assert (rr < (1e-3 if single else 1e-7)).all(), rr.max()
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < (1e-3 if single else 1e-6)


def test_autotune_and_graph_replay(dev, vh, oracle):
    sys_ptr, rp, col, val = small_batch(vh, ns=64, n_max=4096, seed=5)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    A = vh.CsrMatrix(rp, col, val)
    plan = vh.BatchPlan(sys_ptr, rp)
    plan.autotune(A)
    lanes = plan.lanes()
    sizes = np.diff(sys_ptr)
    classes = set(int(np.floor(np.log2(s))) for s in sizes)
    for c in range(16):
        if c in classes:
            assert 0 <= lanes[c] <= 5  # 2^l lanes per row, or 5: staged over row blocks
        else:
            assert lanes[c] == -1
    _, ito, _ = oracle.batched_cg(sys_ptr, oracle.CsrHost(rp, col, val), b, 1e-8, 3000)
    stream = vh.Stream()
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, np.float64)
    ns = sys_ptr.size - 1
    it_d = vh.DeviceArray.zeros(ns, np.uint32)
    er_d = vh.DeviceArray.zeros(ns, np.float64)
    cv_d = vh.DeviceArray.zeros(ns, np.uint32)
    plan.solve_into(A, db, dx, it_d, er_d, cv_d, stream.handle, tol=1e-8, max_iters=3000)
    stream.synchronize()
    x_eager, it_eager = dx.to_numpy(), it_d.to_numpy()
    assert (cv_d.to_numpy() == 1).all() and iters_close(it_eager, ito).all()
    for rep in range(3):
        vh.call("vclhip_fill_zero", dx.ptr, C.c_size_t(dx.nbytes), stream.handle)
        vh.call("vclhip_fill_zero", it_d.ptr, C.c_size_t(it_d.nbytes), stream.handle)
        plan.solve_into(A, db, dx, it_d, er_d, cv_d, stream.handle, tol=1e-8, max_iters=3000, use_graph=True)
        stream.synchronize()
        assert np.array_equal(dx.to_numpy(), x_eager), rep
        assert np.array_equal(it_d.to_numpy(), it_eager), rep


def test_edge_cases(dev, vh):
    sys_ptr, rp, col, val = small_batch(vh, ns=6, n_max=300, seed=7)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    b[sys_ptr[2]:sys_ptr[3]] = 0.0  # system 2 has b = 0: x = 0 without iterations
    A = vh.CsrMatrix(rp, col, val)
    plan = vh.BatchPlan(sys_ptr, rp)
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(n, np.float64)
    it, err, conv = plan.solve(A, db, dx, tol=1e-8, max_iters=3000)
    assert it[2] == 0 and conv[2] == 1 and not dx.to_numpy()[sys_ptr[2]:sys_ptr[3]].any()
    assert (conv == 1).all()
    # iteration limit
    it, err, conv = plan.solve(A, db, dx, tol=1e-8, max_iters=5)
    assert (it[[0, 1, 3, 4, 5]] == 5).all() and (conv[[0, 1, 3, 4, 5]] == 0).all()
    # a column outside its system's block is reported per system, and nothing is read out of LDS range
    col_bad = col.copy()
    k = rp[sys_ptr[4]]  # first entry of system 4 points into system 0
    col_bad[k] = 0
    Ab = vh.CsrMatrix(rp, col_bad, val)
    for p in (plan, vh.BatchPlan(sys_ptr)):  # staged and lanes-per-row SpMV
        it, err, conv = p.solve(Ab, db, dx, tol=1e-8, max_iters=50)
        assert conv[4] == 2 and (conv[[0, 1, 2, 3, 5]] != 2).all()
    # a system larger than one workgroup holds is refused at plan time
    with pytest.raises(vh.VclHipError):
        vh.BatchPlan(np.array([0, 10, 4107], np.uint32))



def long_row_batch(seed=11):
    """Four systems: 300 rows, an empty one, 2600 rows with two dense rows of more than 2048 entries,
    and 900 rows with one denser row; each SPD (symmetric, diagonally dominant)."""
    rng = np.random.default_rng(seed)
    sizes = [300, 0, 2600, 900]
    sys_ptr = np.zeros(len(sizes) + 1, np.uint32)
    sys_ptr[1:] = np.cumsum(sizes)
    ri, ci, vv = [], [], []
    for s, n in enumerate(sizes):
        if n == 0:
            continue
        a = int(sys_ptr[s])
        M = {}

        def add(i, j, w):
            M[(i, j)] = M.get((i, j), 0.0) - w
            M[(j, i)] = M.get((j, i), 0.0) - w

        for i in range(n):
            for j in rng.choice(n, 4, replace=False):
                if i != j:
                    add(i, int(j), rng.uniform(0.1, 1.0))
        dense, frac = ((5, n // 2), 0.9) if n > 2100 else ((3,), 0.3)
        for d in dense:
            for j in np.nonzero(rng.random(n) < frac)[0]:
                if j != d:
                    add(d, int(j), rng.uniform(0.01, 0.1))
        diag = np.zeros(n)
        for (i, _), v in M.items():
            diag[i] -= v
        for i in range(n):
            M[(i, i)] = diag[i] + 0.05
        for (i, j) in sorted(M):
            ri.append(a + i)
            ci.append(a + j)
            vv.append(M[(i, j)])
    ntot = int(sys_ptr[-1])
    rp = np.zeros(ntot + 1, np.int64)
    np.add.at(rp, np.array(ri) + 1, 1)
    return sys_ptr, np.cumsum(rp).astype(np.uint32), np.array(ci, np.uint32), np.array(vv)


def test_staged_long_rows_and_empty_system(dev, vh, oracle):
    """Staged SpMV over every block kind: rows longer than 32 entries (a wavefront), dense rows of
    more than 2048 entries (a block of their own), and an empty system."""
    sys_ptr, rp, col, val = long_row_batch()
    ntot = rp.size - 1
    assert np.diff(rp.astype(np.int64)).max() > 2048
    b = vh.random_uniform(2, ntot)
    A = vh.CsrMatrix(rp, col, val)
    plan = vh.BatchPlan(sys_ptr, rp)
    db, dx = vh.DeviceArray.from_numpy(b), vh.DeviceArray.zeros(ntot, np.float64)
    it, err, conv = plan.solve(A, db, dx, tol=1e-8, max_iters=3000)
    xo, ito, _ = oracle.batched_cg(sys_ptr, oracle.CsrHost(rp, col, val), b, 1e-8, 3000)
    assert (conv == 1).all() and it[1] == 0
    assert iters_close(it, ito).all(), np.c_[it, ito]
    x = dx.to_numpy()
    r = b - oracle.csr_spmv(rp, col, val, x)
    for s in (0, 2, 3):
        a, e = int(sys_ptr[s]), int(sys_ptr[s + 1])
        assert np.linalg.norm(r[a:e]) / np.linalg.norm(b[a:e]) < 1e-7
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6
