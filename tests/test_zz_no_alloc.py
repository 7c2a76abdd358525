"""vclhip_cg_solve_csr_f64_noalloc (the CG solve that allocates nothing, its memory a 32 MiB
scratch array per device) against vclhip_cg_solve_csr_f64, through the C ABI: the same iterations,
error, operator and solution bit for bit, on two calls in a row with the second system smaller
(the scratch array keeps the first call's contents), with and without Jacobi, with a history; a
system too large for the array and a graph request are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def solve(vh, A, b, entry, precond=0, **kw):
    db = vh.DeviceArray.from_numpy(b)
    dx = vh.DeviceArray.zeros(b.size, np.float64)
    hist = np.zeros(kw.pop("history_len", 0))
    params = vh.solver_params(history=hist if hist.size else None, **kw)
    res = vh.SolverResult()
    vh.call(entry, C.byref(A.desc), db.ptr, dx.ptr, precond, C.byref(params), C.byref(res), None)
    return dx.to_numpy(), res, hist


def same_solve(vh, A, b, precond=0, **kw):
    x0, r0, h0 = solve(vh, A, b, "vclhip_cg_solve_csr_f64", precond, **dict(kw))
    x1, r1, h1 = solve(vh, A, b, "vclhip_cg_solve_csr_f64_noalloc", precond, **dict(kw))
    print(f"n={b.size} precond={precond}: {r1.iters} iterations, error {r1.error:.3e}, operator {r1.operator_used}",
          flush=True)
    assert r1.converged == r0.converged and r1.iters == r0.iters and r1.error == r0.error
    assert r1.operator_used == r0.operator_used
    assert np.array_equal(x1, x0)
    assert np.array_equal(h1, h0)
    return r1


def test_noalloc_cg_equals_cg_on_two_calls(dev, vh):
    # first call: the 256 x 256 Poisson matrix (n = 65,536), solved on its DIA copy in the array
    rp, col, val = vh.fdm_laplace_2d(256, 256)
    A = vh.CsrMatrix(rp, col, val)
    r = same_solve(vh, A, vh.random_uniform(0, rp.size - 1), tol=1e-8, max_iters=2000, history_len=2000)
    assert r.converged and r.operator_used == vh.USED_DIA
    # second call, smaller: a 2-D variable-coefficient diffusion operator with Jacobi
    k = np.exp(vh.random_normal(3, 90 * 70))
    rp, col, val = vh.diffusion(90, 70, 1, k)
    A = vh.CsrMatrix(rp, col, val)
    r = same_solve(vh, A, vh.random_uniform(1, rp.size - 1), precond=1, tol=1e-10, max_iters=3000)
    assert r.converged
    # and a third on the matrix as given (no DIA copy), after a larger call left its contents
    same_solve(vh, A, vh.random_uniform(2, rp.size - 1), tol=1e-10, max_iters=3000, operator_format=1)


def test_noalloc_cg_refuses_what_it_cannot_hold(dev, vh):
    n = 1500 * 1000  # 24 bytes per row: more than the 32 MiB array holds
    rp = np.arange(n + 1, dtype=np.uint32)
    col = np.arange(n, dtype=np.uint32)
    A = vh.CsrMatrix(rp, col, np.full(n, 2.0))
    b = np.ones(n)
    with pytest.raises(vh.VclHipError) as e:
        solve(vh, A, b, "vclhip_cg_solve_csr_f64_noalloc", tol=1e-8, max_iters=10)
    assert e.value.status == vh.ERROR_TOO_LARGE
    rp, col, val = vh.fdm_laplace_2d(20, 20)
    A = vh.CsrMatrix(rp, col, val)
    with pytest.raises(vh.VclHipError) as e:
        solve(vh, A, np.ones(400), "vclhip_cg_solve_csr_f64_noalloc", tol=1e-8, max_iters=10, use_graph=True)
    assert e.value.status == vh.ERROR_INVALID_VALUE
    # the refusals left nothing behind: a solve right after them still equals the allocating one
    same_solve(vh, A, vh.random_uniform(4, 400), tol=1e-10, max_iters=500)


def test_noalloc_cg_on_a_coded_dia_copy(dev, vh):
    """2^18 rows (the 512 x 512 Poisson matrix): the DIA copy, its value dictionary and the
    dictionary's temporaries all come from the scratch array, and the solve equals the allocating
    one bit for bit on the coded copy; a smaller call after it still does."""
    rp, col, val = vh.fdm_laplace_2d(512, 512)
    A = vh.CsrMatrix(rp, col, val)
    r = same_solve(vh, A, vh.random_uniform(8, rp.size - 1), tol=1e-8, max_iters=3000)
    assert r.converged and r.operator_used == vh.USED_DIA_CODED
    rp, col, val = vh.fdm_laplace_2d(100, 90)
    A = vh.CsrMatrix(rp, col, val)
    r = same_solve(vh, A, vh.random_uniform(9, rp.size - 1), tol=1e-8, max_iters=3000)
    assert r.converged and r.operator_used == vh.USED_DIA
