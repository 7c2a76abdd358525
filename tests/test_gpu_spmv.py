"""GPU parity of the CSR / SELL SpMV kernels against the oracle and the golden fixtures.

Bar: relative inf-norm error <= 1e-13 for FP64, bit-exact where a row is reduced by a
single lane (the 5-/7-point stencils: one lane per row, same summation order as the host loop),
<= 1e-5 relative for FP32.  Full-size (128^3) checks use size-independent properties.
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def relinf(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def run_csr(vh, rp, col, val, x, alpha=1.0, y0=None, beta=0.0, dtype=np.float64):
    A = vh.CsrMatrix(rp, col, val, dtype=dtype)
    dx = vh.DeviceArray.from_numpy(np.asarray(x, dtype))
    dy = vh.DeviceArray.from_numpy(np.asarray(y0 if y0 is not None else np.zeros(A.n), dtype))
    A.spmv(dx, dy, alpha, beta)
    return dy.to_numpy()


@pytest.mark.parametrize("case", ["lap2d_30x20", "powerlaw_3000", "empty_rows_400"])
def test_csr_golden(dev, vh, case):
    g = dict(np.load(GOLDEN / f"spmv_{case}.npz"))
    for tag, alpha, beta in (("a1b0", 1.0, 0.0), ("a2b05", 2.0, 0.5), ("am1b1", -1.0, 1.0)):
        y = run_csr(vh, g["row_ptr"], g["col"], g["val"], g["x"], alpha, g["y0"], beta)
        if case == "powerlaw_3000":
            assert relinf(y, g[f"y_{tag}"]) <= 1e-13, tag
        else:
            assert np.array_equal(y, g[f"y_{tag}"]), tag


@pytest.mark.parametrize("case", ["lap2d_37x29_c32", "powerlaw_1001_c64"])
def test_sell_golden(dev, vh, case):
    g = dict(np.load(GOLDEN / f"sell_{case}.npz"))
    S = vh.SellMatrix(g["row_ptr"], g["col"], g["val"], int(g["C"]))
    dx = vh.DeviceArray.from_numpy(g["x"])
    dy = vh.DeviceArray.from_numpy(g["y0"])
    S.spmv(dx, dy, float(g["alpha"]), float(g["beta"]))
    # one lane per row, sequential over the slice's columns: bit-exact
    assert np.array_equal(dy.to_numpy(), g["y"])


@pytest.mark.parametrize("shape", ["lap3d_32", "powerlaw_20000", "long_rows", "tiny"])
def test_csr_against_oracle(dev, vh, oracle, shape):
    if shape == "lap3d_32":
        rp, col, val = vh.laplace_3d(32)
    elif shape == "powerlaw_20000":
        rp, col, val = vh.powerlaw(20000, max_len=512, seed=2)
    elif shape == "long_rows":
        # rows longer than the LDS staging buffer (direct mode), empty rows, full-wavefront rows
        lens = np.array([3, 5000, 1, 0, 2049, 2048, 7, 0, 0, 300, 64, 65] * 20)
        n = lens.size
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
        rng = np.random.default_rng(0)
        col = np.concatenate([np.sort(rng.choice(n, size=min(L, n), replace=False)) if L <= n else
                              rng.integers(0, n, L) for L in lens]).astype(np.uint32)
        val = rng.standard_normal(rp[-1])
    else:
        rp, col, val = vh.fdm_laplace_2d(3, 1)
    n = rp.size - 1
    x = vh.random_uniform(1, n)
    y0 = vh.random_uniform(2, n)
    for alpha, beta in ((1.0, 0.0), (0.5, -2.0)):
        y = run_csr(vh, rp, col, val, x, alpha, y0, beta)
        yo = oracle.csr_spmv(rp, col, val, x, alpha, y0, beta)
        assert relinf(y, yo) <= 1e-13
        if shape in ("lap3d_32", "tiny"):
            assert np.array_equal(y, yo)


def test_csr_strided_vectors(dev, vh, oracle):
    rp, col, val = vh.powerlaw(5000, max_len=100, seed=4)
    n = rp.size - 1
    x = vh.random_uniform(1, 3 * n + 2)
    y0 = vh.random_uniform(2, 2 * n + 7)
    A = vh.CsrMatrix(rp, col, val)
    dx = vh.DeviceArray.from_numpy(x)
    dy = vh.DeviceArray.from_numpy(y0)
    A.spmv(dx, dy, 1.25, 0.5, xs=2, xi=3, ys=7, yi=2)
    y = dy.to_numpy()
    yo = oracle.csr_spmv(rp, col, val, x, 1.25, y0, 0.5, xs=2, xi=3, ys=7, yi=2)
    assert relinf(y, yo) <= 1e-13
    assert np.array_equal(y[:7], y0[:7]) and np.array_equal(y[8::2], y0[8::2])


def test_csr_and_sell_fp32(dev, vh, oracle):
    rp, col, val = vh.laplace_3d(24, 20, 16)
    n = rp.size - 1
    x = vh.random_uniform(3, n).astype(np.float32)
    y = run_csr(vh, rp, col, val, x, dtype=np.float32)
    yo = oracle.csr_spmv(rp, col, val.astype(np.float32), x, dtype=np.float32)
    assert np.array_equal(y, yo)
    S = vh.SellMatrix(rp, col, val, 64, dtype=np.float32)
    dx = vh.DeviceArray.from_numpy(x)
    dy = vh.DeviceArray.zeros(n, np.float32)
    S.spmv(dx, dy)
    assert np.array_equal(dy.to_numpy(), yo)


@pytest.mark.parametrize("C_", [32, 64, 128, 256])
def test_sell_against_oracle(dev, vh, oracle, C_):
    rp, col, val = vh.powerlaw(3000 + C_ // 2, max_len=200, seed=C_)
    n = rp.size - 1
    S = vh.SellMatrix(rp, col, val, C_)
    So = oracle.SellHost(rp, col, val, C_)
    x = vh.random_uniform(7, n)
    y0 = vh.random_uniform(8, n)
    dx, dy = vh.DeviceArray.from_numpy(x), vh.DeviceArray.from_numpy(y0)
    S.spmv(dx, dy, -1.0, 0.25)
    assert np.array_equal(dy.to_numpy(), oracle.sell_spmv(So, x, -1.0, y0, 0.25))


def test_row_info(dev, vh, oracle):
    rp, col, val = vh.powerlaw(6000, max_len=300, seed=3)
    A = vh.CsrMatrix(rp, col, val)
    out = vh.DeviceArray.zeros(A.n, np.float64)
    for info in range(4):
        A.row_info(out, info)
        assert np.allclose(out.to_numpy(), oracle.row_info(rp, col, val, info), rtol=1e-14, atol=0)


def test_fullsize_3d_laplacian_properties(dev, vh):
    """128^3 (n = 2,097,152, nnz = 14,581,760): eigenpair, row sums, linearity."""
    N = 128
    rp, col, val = vh.laplace_3d(N)
    A = vh.CsrMatrix(rp, col, val)
    S = vh.SellMatrix(rp, col, val, 64)
    n = A.n
    i = np.arange(N) + 1
    a, b, c = 1, 2, 3
    sx, sy, sz = (np.sin(i * m * np.pi / (N + 1)) for m in (a, b, c))
    v = (sz[:, None, None] * sy[None, :, None] * sx[None, None, :]).ravel()
    lam = sum(2 - 2 * np.cos(m * np.pi / (N + 1)) for m in (a, b, c))
    dv = vh.DeviceArray.from_numpy(v)
    dy = vh.DeviceArray.zeros(n, np.float64)
    A.spmv(dv, dy)
    y = dy.to_numpy()
    assert np.max(np.abs(y - lam * v)) < 1e-13
    S.spmv(dv, dy)
    assert np.array_equal(dy.to_numpy(), y)
    ones = vh.DeviceArray.from_numpy(np.ones(n))
    A.spmv(ones, dy)
    rs = dy.to_numpy().reshape(N, N, N)
    assert rs[5, 6, 7] == 0 and rs[0, 0, 0] == 3 and rs[0, 5, 5] == 1 and rs.sum() == 6 * N * N
    # linearity: A(2u - 3w) = 2Au - 3Aw (bitwise not required; same summation order gives 1 ulp)
    u = vh.random_uniform(1, n)
    w = vh.random_uniform(2, n)
    du, dw, dz = (vh.DeviceArray.from_numpy(t) for t in (u, w, 2 * u - 3 * w))
    yu, yw, yz = (vh.DeviceArray.zeros(n, np.float64) for _ in range(3))
    A.spmv(du, yu)
    A.spmv(dw, yw)
    A.spmv(dz, yz)
    assert relinf(yz.to_numpy(), 2 * yu.to_numpy() - 3 * yw.to_numpy()) < 1e-14


def test_rejects_bad_arguments(dev, vh):
    rp, col, val = vh.fdm_laplace_2d(4, 4)
    A = vh.CsrMatrix(rp, col, val)
    dx = vh.DeviceArray.zeros(16, np.float64)
    with pytest.raises(vh.VclHipError):
        A.spmv(dx, dx)  # aliasing x and y
    with pytest.raises(vh.VclHipError):
        A.spmv(dx, vh.DeviceArray.zeros(16, np.float64), xi=0)
