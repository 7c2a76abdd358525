"""GPU parity of the vector entry points (inner_prod, norm_2, avbv, fill, element_div, convert)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev_arr(vh, a):
    return vh.DeviceArray.from_numpy(a)


@pytest.mark.parametrize("n", [1, 255, 256, 1000, 65536, 1_000_003])
def test_dot_and_norm(dev, vh, oracle, n):
    x = vh.random_uniform(1, n)
    y = vh.random_uniform(2, n)
    dx, dy = dev_arr(vh, x), dev_arr(vh, y)
    r = C.c_double()
    vh.call("vclhip_dot_f64", n, dx.ptr, 0, 1, dy.ptr, 0, 1, C.byref(r), None)
    ref = oracle.dot(x, y)
    # different summation order: bound by n * eps * sum|x_i y_i|
    assert abs(r.value - ref) <= 4 * n * 2.2e-16 * np.sum(np.abs(x * y)) + 1e-300
    vh.call("vclhip_norm2_f64", n, dx.ptr, 0, 1, C.byref(r), None)
    assert abs(r.value - oracle.norm2(x)) <= 1e-13 * oracle.norm2(x)


def test_dot_strided_and_mixed_precision(dev, vh):
    n = 30000
    x = vh.random_uniform(3, 3 * n)
    y = vh.random_uniform(4, 2 * n + 5)
    dx, dy = dev_arr(vh, x), dev_arr(vh, y)
    r = C.c_double()
    vh.call("vclhip_dot_f64", n, dx.ptr, 1, 3, dy.ptr, 5, 2, C.byref(r), None)
    assert np.isclose(r.value, np.dot(x[1::3][:n], y[5::2][:n]), rtol=1e-12)
    xf, yf = x[:n].astype(np.float32), y[:n].astype(np.float32)
    dxf, dyf = dev_arr(vh, xf), dev_arr(vh, yf)
    vh.call("vclhip_dot_f32_acc64", n, dxf.ptr, dyf.ptr, C.byref(r), None)
    exact = np.dot(xf.astype(np.float64), yf.astype(np.float64))
    assert abs(r.value - exact) <= 1e-12 * np.sum(np.abs(xf.astype(np.float64) * yf))
    rf = C.c_float()
    vh.call("vclhip_dot_f32", n, dxf.ptr, 0, 1, dyf.ptr, 0, 1, C.byref(rf), None)
    assert abs(rf.value - exact) <= 1e-4 * np.sum(np.abs(xf.astype(np.float64) * yf))


def test_axpby_fill_ediv_convert(dev, vh):
    n = 100_001
    x = vh.random_uniform(5, n)
    y = vh.random_uniform(6, n) + 2.0
    dx, dy = dev_arr(vh, x), dev_arr(vh, y)
    dz = vh.DeviceArray.zeros(n, np.float64)
    vh.call("vclhip_axpby_f64", n, 0.3, dx.ptr, -1.7, dy.ptr, dz.ptr, None)
    assert np.array_equal(dz.to_numpy(), x * 0.3 + y * -1.7)  # avbv form: v2*alpha + v3*beta
    vh.call("vclhip_axpby_f64", n, 2.0, dx.ptr, 1.0, dz.ptr, dz.ptr, None)  # z may alias
    assert np.array_equal(dz.to_numpy(), x * 2.0 + (x * 0.3 + y * -1.7))
    vh.call("vclhip_element_div_f64", n, dx.ptr, dy.ptr, dz.ptr, None)
    assert np.array_equal(dz.to_numpy(), x / y)
    vh.call("vclhip_fill_f64", n // 3, dz.ptr, 1, 3, 4.5, None)
    out = dz.to_numpy()
    assert np.all(out[1::3][: n // 3] == 4.5) and np.array_equal(out[0::3], (x / y)[0::3])
    df = vh.DeviceArray.zeros(n, np.float32)
    vh.call("vclhip_convert_f64_to_f32", n, dx.ptr, df.ptr, None)
    assert np.array_equal(df.to_numpy(), x.astype(np.float32))
    vh.call("vclhip_convert_f32_to_f64", n, df.ptr, dz.ptr, None)
    assert np.array_equal(dz.to_numpy(), x.astype(np.float32).astype(np.float64))
    # strided avbv: v1[2i+1] = v2[i]*a + v3[3i]*b
    v1 = vh.DeviceArray.zeros(2 * 1000 + 1, np.float64)
    vh.call("vclhip_avbv_f64", 1000, v1.ptr, 1, 2, dx.ptr, 0, 1, 1.5, dy.ptr, 0, 3, -0.5, None)
    assert np.array_equal(v1.to_numpy()[1::2], x[:1000] * 1.5 + y[0:3000:3] * -0.5)


def test_fill_zero_and_copies(dev, vh):
    for nbytes in (1, 7, 16, 17, 4099, 1 << 20):
        a = vh.DeviceArray.from_numpy(np.full(nbytes, 0xAB, np.uint8))
        vh.call("vclhip_fill_zero", a.ptr, nbytes, None)
        assert not a.to_numpy().any()
    src = vh.DeviceArray.from_numpy(np.arange(1000, dtype=np.float64))
    dst = vh.DeviceArray.zeros(1000, np.float64)
    vh.call("vclhip_memcpy_d2d", dst.ptr, src.ptr, 8000, None)
    assert np.array_equal(dst.to_numpy(), np.arange(1000.0))


@pytest.mark.parametrize("n", [1, 257, 100_003])
def test_asum_iamax(dev, vh, n):
    x = vh.random_uniform(11, 2 * n + 1) - 0.5
    dx = dev_arr(vh, x)
    r, rf, idx = C.c_double(), C.c_float(), C.c_uint32()
    sub = x[1::2][:n]
    vh.call("vclhip_asum_f64", n, dx.ptr, 1, 2, C.byref(r), None)
    assert abs(r.value - np.sum(np.abs(sub))) <= 4 * n * 2.2e-16 * np.sum(np.abs(sub))
    vh.call("vclhip_iamax_f64", n, dx.ptr, 1, 2, C.byref(idx), None)
    assert idx.value == int(np.argmax(np.abs(sub)))  # numpy also returns the first maximum
    xf = sub.astype(np.float32)
    dxf = dev_arr(vh, xf)
    vh.call("vclhip_asum_f32", n, dxf.ptr, 0, 1, C.byref(rf), None)
    assert abs(rf.value - np.sum(np.abs(xf.astype(np.float64)))) <= 1e-5 * np.sum(np.abs(xf))
    vh.call("vclhip_iamax_f32", n, dxf.ptr, 0, 1, C.byref(idx), None)
    assert idx.value == int(np.argmax(np.abs(xf)))


def test_iamax_ties_and_zero_vector(dev, vh):
    n = 70_000
    x = np.zeros(n)
    x[[5, 40_000, 69_999]] = [-3.0, 3.0, 3.0]  # ties across workgroups: the first one wins
    dx, idx = dev_arr(vh, x), C.c_uint32()
    vh.call("vclhip_iamax_f64", n, dx.ptr, 0, 1, C.byref(idx), None)
    assert idx.value == 5
    z = dev_arr(vh, np.zeros(n))
    vh.call("vclhip_iamax_f64", n - 7, z.ptr, 7, 1, C.byref(idx), None)
    assert idx.value == 7  # no entry above zero: the vector's start offset


def test_scal_copy_swap_rot(dev, vh):
    n = 50_001
    x = vh.random_uniform(12, 3 * n)
    y = vh.random_uniform(13, 2 * n + 4)
    dx, dy = dev_arr(vh, x), dev_arr(vh, y)
    vh.call("vclhip_scal_f64", n, dx.ptr, 2, 3, -1.25, None)
    want_x = x.copy()
    want_x[2::3] *= -1.25
    assert np.array_equal(dx.to_numpy(), want_x)
    vh.call("vclhip_copy_f64", n, dx.ptr, 2, 3, dy.ptr, 4, 2, None)
    want_y = y.copy()
    want_y[4::2] = want_x[2::3]
    assert np.array_equal(dy.to_numpy(), want_y)
    dx, dy = dev_arr(vh, x), dev_arr(vh, y)
    vh.call("vclhip_swap_f64", n, dx.ptr, 0, 3, dy.ptr, 4, 2, None)
    sx, sy = x.copy(), y.copy()
    sx[0::3], sy[4::2] = y[4::2], x[0::3]
    assert np.array_equal(dx.to_numpy(), sx) and np.array_equal(dy.to_numpy(), sy)
    c, s = 0.6, 0.8
    vh.call("vclhip_rot_f64", n, dx.ptr, 0, 3, dy.ptr, 4, 2, c, s, None)
    a, b = sx[0::3].copy(), sy[4::2].copy()
    rx, ry = sx.copy(), sy.copy()
    rx[0::3], ry[4::2] = c * a + s * b, c * b - s * a
    assert np.allclose(dx.to_numpy(), rx, rtol=0, atol=1e-15) and np.allclose(dy.to_numpy(), ry, rtol=0, atol=1e-15)
    # FP32 forms and the strided f32 avbv
    xf, yf = x[:n].astype(np.float32), y[:n].astype(np.float32)
    dxf, dyf = dev_arr(vh, xf), dev_arr(vh, yf)
    vh.call("vclhip_rot_f32", n, dxf.ptr, 0, 1, dyf.ptr, 0, 1, 0.6, 0.8, None)
    c32, s32 = np.float32(0.6), np.float32(0.8)
    assert np.allclose(dxf.to_numpy(), c32 * xf + s32 * yf, rtol=1e-6, atol=1e-6)
    assert np.allclose(dyf.to_numpy(), c32 * yf - s32 * xf, rtol=1e-6, atol=1e-6)
    dxf = dev_arr(vh, xf)
    vh.call("vclhip_scal_f32", n, dxf.ptr, 0, 1, 2.0, None)
    assert np.array_equal(dxf.to_numpy(), xf * 2)
    v1 = vh.DeviceArray.zeros(2 * 1000 + 1, np.float32)
    vh.call("vclhip_avbv_f32", 1000, v1.ptr, 1, 2, dxf.ptr, 0, 1, 1.5, dyf.ptr, 0, 3, -0.5, None)
    got = v1.to_numpy()[1::2]
    want = (xf[:1000] * 2) * np.float32(1.5) + dyf.to_numpy()[0:3000:3] * np.float32(-0.5)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)


def test_latency_probes(dev, vh):
    """The measurement kernels behind the bench's latency reference: empty launches run, the
    device-wide synchronisation completes every step without a workgroup giving up (every arrival
    flag ends at `steps`, the go flag too), and a grid larger than the device holds at once, or
    beyond 4096 workgroups, is refused before anything is launched."""
    flags = vh.DeviceArray.zeros(4200, np.uint32)
    assert vh.call("vclhip_empty_launch", 256, None) == 0
    for g, steps in ((8, 50), (256, 200)):
        assert vh.call("vclhip_grid_sync_probe", g, steps, flags.ptr, None) == 0
        f = flags.to_numpy()
        assert np.all(f[:g] == steps) and f[g] == steps and f[g + 1] == 0
    with pytest.raises(vh.VclHipError) as e:
        vh.call("vclhip_grid_sync_probe", 4097, 1, flags.ptr, None)
    assert e.value.status == vh.ERROR_INVALID_VALUE


def test_gather_reference(dev, vh):
    """vclhip_gather_read_f64 (the power-law SpMV's gather floor in the bench): workgroup w sums
    x[idx[i]] over the w-th contiguous chunk of idx; an index at or beyond n reads x[0]."""
    n, m, g = 5000, 123457, 64
    rng = np.random.default_rng(3)
    idx = rng.integers(0, n, m).astype(np.uint32)
    idx[::1000] = n + 7  # out of range: x[0]
    x = vh.random_uniform(4, n)
    di, dx = vh.DeviceArray.from_numpy(idx), vh.DeviceArray.from_numpy(x)
    sums = vh.DeviceArray.zeros(g, np.float64)
    vh.call("vclhip_gather_read_f64", m, di.ptr, dx.ptr, n, sums.ptr, g, None)
    got = sums.to_numpy()
    safe = np.where(idx < n, idx, 0)
    bounds = [m * w // g for w in range(g + 1)]
    want = np.array([x[safe[bounds[w]:bounds[w + 1]]].sum() for w in range(g)])
    scale = np.array([np.abs(x[safe[bounds[w]:bounds[w + 1]]]).sum() for w in range(g)])
    # sums[w] is chunk w's sum (different summation order: bounded by the absolute sum)
    assert np.all(np.abs(got - want) <= 1e-12 * scale)
