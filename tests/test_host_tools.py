"""Host-side parts of libvclhip that need no GPU: the exported ABI, the CSR block planner,
the generators and the project RNG."""
import ctypes
import subprocess

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(vh):
    from vclhip import _abi
    declared = _abi.declared_functions()
    assert len(declared) > 50
    L = vh.lib()
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding, and nothing extra
    assert sorted(_abi._SIGS) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T vclhip_" in line}
    assert set(declared) == exported


def test_status_strings_and_errors_without_device(vh):
    L = vh.lib()
    assert L.vclhip_status_string(0) == b"success"
    assert b"zero" in L.vclhip_status_string(-2)
    # argument checks fail before touching a device
    with pytest.raises(vh.VclHipError) as e:
        vh.call("vclhip_csr_spmv_f64", None, None, 0, 1, 1.0, None, 0, 1, 0.0, None)
    assert e.value.status == -1
    # row pointers that decrease would make the kernels read outside the arrays: the planner refuses
    with pytest.raises(vh.VclHipError) as e:
        vh.plan_blocks(np.array([0, 3, 2, 5], np.uint32))
    assert e.value.status == -1 and b"row 1" in vh.lib().vclhip_last_error()


def check_plan(rp, blocks):
    """Blocks of <= 256 rows and <= 2048 entries (mode 1 when a row is longer than 32, else 0); a row
    longer than 2048 is a block of its own (mode 7); blocks are maximal (greedy)."""
    n = rp.size - 1
    assert blocks[0] & ((1 << 28) - 1) == 0 and blocks[-1] == n
    rows = blocks & ((1 << 28) - 1)
    modes = blocks >> 28
    assert np.all(np.diff(rows.astype(np.int64)) > 0)
    lens_all = np.diff(rp.astype(np.int64))
    for b in range(blocks.size - 1):
        r0, r1, m = rows[b], rows[b + 1], modes[b]
        lens = lens_all[r0:r1]
        if m == 7:
            assert r1 == r0 + 1 and lens[0] > 2048
            continue
        assert m in (0, 1) and (m == 1) == bool((lens > 32).any())
        assert lens.sum() <= 2048 and r1 - r0 <= 256 and (lens <= 2048).all()
        if r1 < n:  # greedy: the next row did not fit
            nxt = lens_all[r1]
            assert r1 - r0 == 256 or nxt > 2048 or lens.sum() + nxt > 2048


@pytest.mark.parametrize("kind", ["lap3d", "powerlaw", "long_rows", "empty"])
def test_block_plan_invariants(vh, kind):
    if kind == "lap3d":
        rp, _, _ = vh.laplace_3d(20, 17, 9)
    elif kind == "powerlaw":
        rp, _, _ = vh.powerlaw(20000, max_len=512, seed=1)
    elif kind == "long_rows":
        lens = np.array([3, 5000, 1, 0, 2049, 2048, 7, 0, 0, 300] * 50)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    else:
        rp = np.zeros(1001, np.uint32)
    blocks = vh.plan_blocks(rp)
    check_plan(rp, blocks)
    if kind == "lap3d":
        # 7-point rows: short rows only, 256 rows per block
        assert np.all(blocks[:-1] >> 28 == 0)
        assert np.median(np.diff(blocks & ((1 << 28) - 1))) == 256
    if kind == "powerlaw":
        # long rows no longer shrink the blocks: most blocks are full (256 rows or 2048 entries)
        sizes = np.diff(blocks & ((1 << 28) - 1))
        assert np.median(sizes) >= 200


def test_rng_is_the_documented_mt19937_64(vh):
    # std::mt19937_64 with the default seed 5489 yields 14514284786278117030 as its first draw
    u = vh.random_uniform(5489, 3, 0.0, 1.0)
    assert u[0] == (14514284786278117030 >> 11) * 2.0 ** -53
    z = vh.random_normal(7, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    assert np.array_equal(vh.random_uniform(0, 10), vh.random_uniform(0, 10))


def test_laplace_3d_structure(vh):
    rp, col, val = vh.laplace_3d(128, 128, 128)
    n = 128 ** 3
    assert rp.size == n + 1 and rp[-1] == 14581760
    assert np.all(np.diff(rp) <= 7) and np.all(np.diff(rp) >= 4)


def test_diffusion_reduces_to_laplacian_and_is_symmetric(vh):
    rp, col, val = vh.diffusion(7, 6, 5, np.ones(210))
    rp2, col2, val2 = vh.laplace_3d(7, 6, 5)
    assert np.array_equal(rp, rp2) and np.array_equal(col, col2) and np.allclose(val, val2)
    k = np.exp(vh.random_normal(1, 210, 0.0, 1.0))
    rp, col, val = vh.diffusion(7, 6, 5, k)
    A = {(i, int(col[j])): val[j] for i in range(210) for j in range(rp[i], rp[i + 1])}
    for (i, j), v in A.items():
        assert A[(j, i)] == v
    # interior rows sum to zero, boundary rows to the boundary face coefficients
    rs = np.array([val[rp[i]:rp[i + 1]].sum() for i in range(210)])
    assert np.all(rs > -1e-12)
    assert np.allclose(rs[5 + 7 * (3 + 6 * 2)], 0, atol=1e-12)


def test_powerlaw_statistics(vh):
    rp, col, val = vh.powerlaw(200000, max_len=512, alpha=2.0, diag_factor=1.1, seed=1)
    lens = np.diff(rp.astype(np.int64)) - 1
    assert lens.min() >= 1 and lens.max() <= 512
    assert abs(lens.mean() - 4.15) < 0.1
    for i in range(0, 200000, 997):
        seg = slice(rp[i], rp[i + 1])
        c, v = col[seg], val[seg]
        assert np.all(np.diff(c.astype(np.int64)) > 0)
        d = v[c == i][0]
        assert np.isclose(d, 1.1 * np.abs(v[c != i]).sum())


def test_graph_laplacian_batch(vh):
    sys_ptr, rp, col, val = vh.graph_laplacian_batch(40, 64, 4096, seed=3)
    sizes = np.diff(sys_ptr.astype(np.int64))
    assert sizes.min() >= 64 and sizes.max() <= 4096
    for s in range(0, 40, 7):
        r0, r1 = sys_ptr[s], sys_ptr[s + 1]
        for i in range(r0, min(r1, r0 + 50)):
            c = col[rp[i]:rp[i + 1]]
            v = val[rp[i]:rp[i + 1]]
            assert np.all((c >= r0) & (c < r1))
            assert np.isclose(v.sum(), 0.01)
            assert np.all(v[c != i] < 0)


def test_coo_plan_rows_and_order(vh):
    """vclhip_coo_plan: row pointers of the row order and the CSR plan; unordered rows are refused."""
    rp, col, val = vh.powerlaw(3000, max_len=300, seed=9)
    n, nnz = rp.size - 1, int(rp[-1])
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp.astype(np.int64)))
    coords = np.empty(2 * nnz, np.uint32)
    coords[0::2], coords[1::2] = rows, col
    prp, blocks, nb = np.zeros(n + 1, np.uint32), np.zeros(n + 1, np.uint32), ctypes.c_uint32()
    vh.call("vclhip_coo_plan", _abi_ptr(coords), nnz, n, _abi_ptr(prp), _abi_ptr(blocks), ctypes.byref(nb))
    assert np.array_equal(prp, rp)
    ref_blocks = vh.plan_blocks(rp)
    assert np.array_equal(blocks[: nb.value + 1], ref_blocks)
    bad = coords.copy()
    bad[0], bad[2 * (nnz - 1)] = bad[2 * (nnz - 1)], bad[0]
    with pytest.raises(vh.VclHipError):
        vh.call("vclhip_coo_plan", _abi_ptr(bad), nnz, n, _abi_ptr(prp), _abi_ptr(blocks), ctypes.byref(nb))


def _abi_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)
