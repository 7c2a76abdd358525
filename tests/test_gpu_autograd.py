"""GPU checks of the differentiable solve (vclhip.torch_solve.SparseSolve).

The adjoint kernel dA[k] = -lambda[row(k)] * x[col(k)] is checked bit-exact against the oracle.
dL/dtheta of L = 0.5 ||x(theta) - x_t||^2, x = A(theta)^{-1} b on the 2-D variable-coefficient
diffusion operator, is checked against central finite differences (h = 1e-6) and against the
gradient formed from the FP64 oracle (host CG for x and lambda, host adjoint product) chained
through the same torch assembly on the CPU.

torch is imported inside the tests, after libvclhip has been loaded, so that both use one HIP
runtime.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_adjoint_grad_kernel_bit_exact(dev, vh, oracle):
    # power-law rows reach the long-row blocks (> 2048 entries) and empty-offdiagonal rows
    rp, col, val = vh.powerlaw(20000, 4000, 1.6, 1.1, 7)
    n = rp.size - 1
    lam = vh.random_uniform(3, n)
    x = vh.random_uniform(4, n)
    A = vh.CsrMatrix(rp, col, val)
    dl = vh.DeviceArray.from_numpy(lam)
    dx = vh.DeviceArray.from_numpy(x)
    dA = vh.DeviceArray.zeros(rp[-1], np.float64)
    vh.call("vclhip_csr_adjoint_grad_f64", vh._abi.C.byref(A.desc), dl.ptr, dx.ptr, dA.ptr, None)
    assert np.array_equal(dA.to_numpy(), oracle.adjoint_grad(rp, col, lam, x))


def _problem(vh, nx, ny):
    import torch
    from vclhip import torch_solve as ts
    theta = np.exp(0.5 * vh.random_normal(1, nx * ny))
    b = vh.random_uniform(0, nx * ny)
    xt = vh.random_uniform(2, nx * ny)
    rp, col, pos, masks = ts.diffusion_pattern_2d(nx, ny)
    return theta, b, xt, rp, col, pos, masks


def test_gradient_vs_finite_differences_and_host(dev, vh, oracle):
    import torch
    from vclhip import torch_solve as ts
    nx, ny = 48, 40
    theta, b, xt, rp, col, pos, masks = _problem(vh, nx, ny)
    pat = ts.CsrPattern(rp, col)
    dev_t = torch.device("cuda")
    th = torch.tensor(theta.reshape(ny, nx), dtype=torch.float64, device=dev_t, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device=dev_t, requires_grad=True)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev_t)

    def loss_of(theta_t, tol=1e-13):
        vals = ts.diffusion_values_2d(theta_t, nx, ny, pos, masks)
        x = ts.sparse_solve(vals, bt, pat, tol=tol, max_iters=20000)
        return 0.5 * torch.sum((x - xtt) ** 2), x

    L, x = loss_of(th)
    L.backward()
    g = th.grad.detach().cpu().numpy().ravel()
    gb = bt.grad.detach().cpu().numpy()

    # FP64 host gradient: oracle CG for x and lambda, oracle adjoint product, chained on the CPU
    vals_h = ts.diffusion_values_2d(torch.tensor(theta.reshape(ny, nx)), nx, ny, pos, masks).numpy()
    Ah = oracle.CsrHost(rp, col, vals_h)
    xh, _, _, _ = oracle.cg_pipelined(Ah, b, 1e-13, 20000)
    lamh, _, _, _ = oracle.cg_pipelined(Ah, xh - xt, 1e-13, 20000)
    dAh = oracle.adjoint_grad(rp, col, lamh, xh)
    th_c = torch.tensor(theta.reshape(ny, nx), requires_grad=True)
    ts.diffusion_values_2d(th_c, nx, ny, pos, masks).backward(torch.tensor(dAh))
    gh = th_c.grad.numpy().ravel()
    assert np.linalg.norm(g - gh) / np.linalg.norm(gh) < 1e-8
    assert np.linalg.norm(gb - lamh) / np.linalg.norm(lamh) < 1e-8

    # central finite differences on 16 random entries
    rng = np.random.default_rng(11)
    idx = rng.choice(nx * ny, 16, replace=False)
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        with torch.no_grad():
            Lp, _ = loss_of(torch.tensor(tp.reshape(ny, nx), device=dev_t))
            Lm, _ = loss_of(torch.tensor(tm.reshape(ny, nx), device=dev_t))
        fd = (Lp.item() - Lm.item()) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(g[i]), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_solve_input_checks(dev, vh):
    import torch
    from vclhip import torch_solve as ts
    rp, col, pos, masks = ts.diffusion_pattern_2d(8, 6)
    pat = ts.CsrPattern(rp, col)
    vals = torch.ones(int(rp[-1]), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ts.sparse_solve(vals.float(), torch.ones(48, dtype=torch.float64, device="cuda"), pat)
    with pytest.raises(ValueError):
        ts.sparse_solve(vals, torch.ones(47, dtype=torch.float64, device="cuda"), pat)
