"""GPU tests of the coded DIA operator: a DIA copy of 2^18 rows or more whose diagonals each hold
at most 15 distinct values is read as one byte per entry into per-diagonal value tables.

The table entry a code selects is the stored value itself, so every product, and with it every
solve, equals the uncoded copy's bit for bit (VCLHIP_OPERATOR_DIA_UNCODED), which equals the CSR
kernel's (tests/test_gpu_krylov.py).  Checked here: which operator a solve reports, the 15/16
distinct-value boundary of a diagonal, the 2^18-row threshold, x, iterations, errors and residual
histories equal across the two copies for CG, Jacobi-CG, BiCGStab and GMRES in FP64 and FP32, and
the copy kept across solves (vclhip_dia_create_*)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(vh, kind, A, b, dtype=np.float64, **kw):
    db = vh.DeviceArray.from_numpy(b.astype(dtype))
    dx = vh.DeviceArray.zeros(b.size, dtype)
    hist = np.zeros(3000)
    if kind == "gmres":
        res = vh.gmres(A, db, dx, tol=1e-10 if dtype == np.float64 else 1e-5, max_iters=3000, krylov_dim=20,
                       history=hist, **kw)
    else:
        fn = vh.cg if kind.endswith("cg") else vh.bicgstab
        res = fn(A, db, dx, precond=1 if kind.startswith("jacobi") else 0,
                 tol=1e-10 if dtype == np.float64 else 1e-5, max_iters=3000, history=hist, **kw)
    x = dx.to_numpy()
    db.free()
    dx.free()
    return x, res, hist[: res.iters]


def with_diagonal_values(rp, col, val, distinct):
    """The matrix with its main diagonal set to 4 + 0.01 (row mod distinct): `distinct` values there,
    still diagonally dominant for the 5-point stencil."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    val = val.copy()
    on = col.astype(np.int64) == rows
    val[on] = 4.0 + 0.01 * (rows[on] % distinct)
    return val


@pytest.mark.parametrize("kind", ["cg", "jacobi_cg", "bicgstab", "gmres"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_coded_operator_equals_uncoded(dev, vh, kind, dtype):
    """The 72^3 Laplacian (373,248 rows; one value per diagonal): the default solve runs on the coded
    copy and gives the uncoded copy's x, iterations, error and residual history bit for bit."""
    rp, col, val = vh.laplace_3d(72)
    n = rp.size - 1
    b = vh.random_uniform(17, n)
    A = vh.CsrMatrix(rp, col, val, dtype=dtype)
    xc, rc, hc = run(vh, kind, A, b, dtype)
    xu, ru, hu = run(vh, kind, A, b, dtype, operator_format=vh.OPERATOR_DIA_UNCODED)
    print(f"[dia codes] {kind} {np.dtype(dtype).name}: {rc.iters} iterations, error {rc.error:.3e}", flush=True)
    assert rc.operator_used == vh.USED_DIA_CODED and ru.operator_used == vh.USED_DIA
    assert rc.converged and rc.iters == ru.iters and rc.error == ru.error
    assert np.array_equal(xc, xu) and np.array_equal(hc, hu)


@pytest.mark.parametrize("distinct", [15, 16])
def test_distinct_value_boundary(dev, vh, distinct):
    """512^2 5-point Laplacian (2^18 rows, the threshold) with 15 or 16 distinct values on its main
    diagonal: 15 fit a table (code 0 is the absent entry), 16 keep the uncoded copy; either way the
    solve equals the uncoded one bit for bit."""
    rp, col, val = vh.fdm_laplace_2d(512, 512)
    val = with_diagonal_values(rp, col, val, distinct)
    n = rp.size - 1
    assert n == 1 << 18
    b = vh.random_uniform(19, n)
    A = vh.CsrMatrix(rp, col, val)
    for kind in ("cg", "bicgstab"):
        x, r, h = run(vh, kind, A, b)
        xu, ru, hu = run(vh, kind, A, b, operator_format=vh.OPERATOR_DIA_UNCODED)
        assert r.operator_used == (vh.USED_DIA_CODED if distinct <= 15 else vh.USED_DIA), kind
        assert ru.operator_used == vh.USED_DIA
        assert r.converged and r.iters == ru.iters and np.array_equal(x, xu) and np.array_equal(h, hu), kind


def test_row_threshold_and_variable_coefficients(dev, vh):
    """Below 2^18 rows (511 x 512) the copy stays uncoded; a variable-coefficient operator at
    2^18 rows and more (64^3 diffusion, log-normal coefficients) has too many distinct values and
    stays uncoded too."""
    rp, col, val = vh.fdm_laplace_2d(511, 512)
    A = vh.CsrMatrix(rp, col, val)
    _, r, _ = run(vh, "cg", A, vh.random_uniform(3, rp.size - 1))
    assert r.converged and r.operator_used == vh.USED_DIA
    k = np.exp(0.5 * vh.random_normal(4, 64 ** 3))
    rp, col, val = vh.diffusion(64, 64, 64, k)
    A = vh.CsrMatrix(rp, col, val)
    _, r, _ = run(vh, "cg", A, vh.random_uniform(5, rp.size - 1))
    assert r.converged and r.operator_used == vh.USED_DIA


def test_kept_copy_is_coded(dev, vh):
    """A DIA copy made once (vclhip_dia_create_f64) carries the value tables: solves with it run
    coded and equal the per-solve copy; with VCLHIP_OPERATOR_DIA_UNCODED the same kept copy is
    read as stored, with the same result."""
    rp, col, val = vh.laplace_3d(64, 64, 70)
    n = rp.size - 1
    A = vh.CsrMatrix(rp, col, val)
    d = vh.DiaOperator(A)
    assert d.built and d.nd == 7
    b = vh.random_uniform(23, n)
    x0, r0, _ = run(vh, "cg", A, b)
    x1, r1, _ = run(vh, "cg", A, b, dia=d)
    x2, r2, _ = run(vh, "cg", A, b, dia=d, operator_format=vh.OPERATOR_DIA_UNCODED)
    d.close()
    assert r0.operator_used == r1.operator_used == vh.USED_DIA_CODED and r2.operator_used == vh.USED_DIA
    assert r0.iters == r1.iters == r2.iters and np.array_equal(x0, x1) and np.array_equal(x0, x2)


def test_coded_27_point_stencil(dev, vh):
    """The 27-point stencil on 64^3 (2^18 rows, 27 diagonals of values -1 and 26.5: three full
    groups of eight diagonals in the product and a partial fourth, 27 value tables) runs coded and
    equals the uncoded copy bit for bit."""
    from test_gpu_krylov import stencil_27
    rp, col, val = stencil_27(64, 64, 64)
    n = rp.size - 1
    A = vh.CsrMatrix(rp, col, val)
    b = vh.random_uniform(29, n)
    for kind in ("cg", "bicgstab"):
        x, r, h = run(vh, kind, A, b)
        xu, ru, hu = run(vh, kind, A, b, operator_format=vh.OPERATOR_DIA_UNCODED)
        assert r.operator_used == vh.USED_DIA_CODED and ru.operator_used == vh.USED_DIA, kind
        assert r.converged and r.iters == ru.iters and np.array_equal(x, xu) and np.array_equal(h, hu), kind
