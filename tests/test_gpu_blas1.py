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
