"""Checks of the differentiable solve, run in a process of their own by test_gpu_autograd.py.

torch is imported before libvclhip is loaded, so that the process has one HIP runtime (torch's).
Each check prints one JSON line and exits nonzero on failure.
"""
import json
import pathlib
import sys

import torch  # noqa: F401  (first: its HIP runtime is the one libvclhip binds to)

ROOT = pathlib.Path(__file__).resolve().parent.parent
for sub in ("viennacl-dev_amd", "oracle"):
    sys.path.insert(0, str(ROOT / sub))

import numpy as np  # noqa: E402

import pyoracle as oracle  # noqa: E402
import vclhip as vh  # noqa: E402
from vclhip import torch_solve as ts  # noqa: E402


def gradients(nx=48, ny=40, n_fd=16):
    theta = np.exp(0.5 * vh.random_normal(1, nx * ny))
    b = vh.random_uniform(0, nx * ny)
    xt = vh.random_uniform(2, nx * ny)
    rp, col, pos, masks = ts.diffusion_pattern_2d(nx, ny)
    pat = ts.CsrPattern(rp, col)
    dev = torch.device("cuda")
    th = torch.tensor(theta.reshape(ny, nx), dtype=torch.float64, device=dev, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev)

    def loss_of(theta_t, tol=1e-13):
        vals = ts.diffusion_values_2d(theta_t, nx, ny, pos, masks)
        x = ts.sparse_solve(vals, bt, pat, tol=tol, max_iters=20000)
        return 0.5 * torch.sum((x - xtt) ** 2)

    L = loss_of(th)
    L.backward()
    g = th.grad.detach().cpu().numpy().ravel()
    gb = bt.grad.detach().cpu().numpy()

    # FP64 host gradient: oracle CG for x and lambda, oracle adjoint product, chained on the CPU
    oracle.set_threads(1)
    vals_h = ts.diffusion_values_2d(torch.tensor(theta.reshape(ny, nx)), nx, ny, pos, masks).numpy()
    Ah = oracle.CsrHost(rp, col, vals_h)
    xh, _, _, _ = oracle.cg_pipelined(Ah, b, 1e-13, 20000)
    lamh, _, _, _ = oracle.cg_pipelined(Ah, xh - xt, 1e-13, 20000)
    dAh = oracle.adjoint_grad(rp, col, lamh, xh)
    th_c = torch.tensor(theta.reshape(ny, nx), requires_grad=True)
    ts.diffusion_values_2d(th_c, nx, ny, pos, masks).backward(torch.tensor(dAh))
    gh = th_c.grad.numpy().ravel()
    host_rel = float(np.linalg.norm(g - gh) / np.linalg.norm(gh))
    db_rel = float(np.linalg.norm(gb - lamh) / np.linalg.norm(lamh))

    rng = np.random.default_rng(11)
    idx = rng.choice(nx * ny, n_fd, replace=False)
    h = 1e-6
    fd_bad = []
    worst = 0.0
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        with torch.no_grad():
            Lp = loss_of(torch.tensor(tp.reshape(ny, nx), device=dev)).item()
            Lm = loss_of(torch.tensor(tm.reshape(ny, nx), device=dev)).item()
        fd = (Lp - Lm) / (2 * h)
        err = abs(fd - g[i])
        scale = max(abs(g[i]), 1e-3 * np.abs(g).max())
        worst = max(worst, err / scale)
        if err > 1e-5 * scale:
            fd_bad.append([int(i), fd, float(g[i])])
    ok = host_rel < 1e-8 and db_rel < 1e-8 and not fd_bad
    return ok, {"host_rel": host_rel, "db_rel": db_rel, "fd_worst_rel": worst, "fd_bad": fd_bad}


def gradients_full(nx=512, n_fd=64):
    """Config C5 at its size: the 512 x 512 variable-coefficient diffusion operator (n = 262,144),
    theta = exp(N(0, 0.5)) per cell, L = 0.5 ||x - x_t||^2 with x_t ~ U[-1, 1].  dL/dtheta from the
    device (forward CG, adjoint CG, fused dA kernel, torch chain rule) against
    (1) the FP64 host gradient: oracle CG for x and lambda (16 threads), oracle dA, the same chain
        rule on the CPU.  Both are solutions of the same systems to the same tolerance (1e-13), so
        they agree to about the condition number times 1e-13: relative 2-norm 1e-7 asserted;
    (2) central differences with h = 1e-6 on 64 random entries, (L(theta + h e_i) - L(theta - h e_i))
        / 2h, formed two ways from the device solves:
        * in solution space: with x_- the solve at theta - h e_i, d = x_+ - x_- solves
          A(theta + h e_i) d = -(A(theta + h e_i) - A(theta - h e_i)) x_-, so the difference of the
          losses is 0.5 d . (2 x_- + d - 2 x_t) exactly, with no large number subtracted.  What is
          left is the rounding of the perturbed matrix itself: each value that depends on theta_i
          (the stencil of cell i and its neighbours' diagonals) carries a relative error of at most
          8 eps (the harmonic mean and the four-term diagonal), and dL/da_k = dA_k, so the quotient
          moves by at most noise_i = 8 eps sum_k |dA_k| |a_k| / h over those values -- about 1e-1
          where |dA| reaches 7e6 (dA_k = -lambda_r x_c is large while dL/dtheta_i is a difference of
          such terms).  |fd_i - g_i| <= noise_i asserted (first order; the errors seen are 1/30 to
          1/100 of it);
        * plainly, from the two losses.  L is about 2.5e7 here, so its rounding adds about
          2.5e7 * 1.1e-16 / 2e-6 = 1.4e-3, and the two solves' errors (relative residual 1e-13,
          condition number about 1e5) add (x - x_t) . dx / h, a few 1e-2:
          |fd_i - g_i| <= 1e-3 max(|g_i|, 1e-2 |g|_inf) asserted."""
    ny = nx
    n = nx * ny
    theta = np.exp(0.5 * vh.random_normal(1, n))
    b = vh.random_uniform(0, n)
    xt = vh.random_uniform(2, n)
    rp, col, pos, masks = ts.diffusion_pattern_2d(nx, ny)
    pat = ts.CsrPattern(rp, col)
    dev = torch.device("cuda")
    th = torch.tensor(theta.reshape(ny, nx), dtype=torch.float64, device=dev, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev)
    tol = 1e-13

    def loss_of(theta_t):
        vals = ts.diffusion_values_2d(theta_t, nx, ny, pos, masks)
        x = ts.sparse_solve(vals, bt, pat, tol=tol, max_iters=50000)
        return 0.5 * torch.sum((x - xtt) ** 2)

    L = loss_of(th)
    L.backward()
    g = th.grad.detach().cpu().numpy().ravel()
    gb = bt.grad.detach().cpu().numpy()

    oracle.set_threads(16)
    vals_h = ts.diffusion_values_2d(torch.tensor(theta.reshape(ny, nx)), nx, ny, pos, masks).numpy()
    Ah = oracle.CsrHost(rp, col, vals_h)
    xh, _, _, _ = oracle.cg_pipelined(Ah, b, tol, 50000)
    lamh, _, _, _ = oracle.cg_pipelined(Ah, xh - xt, tol, 50000)
    oracle.set_threads(1)
    dAh = oracle.adjoint_grad(rp, col, lamh, xh)
    th_c = torch.tensor(theta.reshape(ny, nx), requires_grad=True)
    ts.diffusion_values_2d(th_c, nx, ny, pos, masks).backward(torch.tensor(dAh))
    gh = th_c.grad.numpy().ravel()
    host_rel = float(np.linalg.norm(g - gh) / np.linalg.norm(gh))
    db_rel = float(np.linalg.norm(gb - lamh) / np.linalg.norm(lamh))

    import scipy.sparse as sp

    rng = np.random.default_rng(11)
    idx = rng.choice(n, n_fd, replace=False)
    h = 1e-6
    g_inf = float(np.abs(g).max())
    eps = np.finfo(np.float64).eps
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    cols = col.astype(np.int64)

    def rounding_noise(i):
        # the values that depend on theta_i: entries (r, c) with r and c in cell i's stencil
        y0, x0 = divmod(int(i), nx)
        cells = [y0 * nx + x0] + [(y0 + dy) * nx + x0 + dx for dy, dx in ((-1, 0), (1, 0), (0, -1), (0, 1))
                                  if 0 <= y0 + dy < ny and 0 <= x0 + dx < nx]
        m = np.isin(rows, cells) & np.isin(cols, cells)
        return 8 * eps * float(np.sum(np.abs(dAh[m]) * np.abs(vals_h[m]))) / h

    fd_bad, worst, worst_sol = [], 0.0, 0.0
    with torch.no_grad():
        for i in idx:
            tp, tm = theta.copy(), theta.copy()
            tp[i] += h
            tm[i] -= h
            vp = ts.diffusion_values_2d(torch.tensor(tp.reshape(ny, nx), device=dev), nx, ny, pos, masks)
            vm = ts.diffusion_values_2d(torch.tensor(tm.reshape(ny, nx), device=dev), nx, ny, pos, masks)
            xp = ts.sparse_solve(vp, bt.detach(), pat, tol=tol, max_iters=50000)
            xm = ts.sparse_solve(vm, bt.detach(), pat, tol=tol, max_iters=50000)
            Lp = (0.5 * torch.sum((xp - xtt) ** 2)).item()
            Lm = (0.5 * torch.sum((xm - xtt) ** 2)).item()
            fd = (Lp - Lm) / (2 * h)
            # solution space: A_+ d = -(A_+ - A_-) x_-, the matrix difference formed entrywise
            dv = (vp - vm).cpu().numpy()
            xm_h = xm.cpu().numpy()
            rhs = -(sp.csr_matrix((dv, col, rp), shape=(n, n)) @ xm_h)
            d = ts.sparse_solve(vp, torch.tensor(rhs, device=dev), pat, tol=tol, max_iters=50000).cpu().numpy()
            fd_sol = float(0.5 * np.dot(d, 2 * xm_h + d - 2 * xt)) / (2 * h)
            scale = max(abs(g[i]), 1e-2 * g_inf)
            err, err_sol = abs(fd - g[i]) / scale, abs(fd_sol - g[i]) / rounding_noise(i)
            worst, worst_sol = max(worst, err), max(worst_sol, err_sol)
            if err > 1e-3 or err_sol > 1.0:
                fd_bad.append([int(i), fd, fd_sol, float(g[i]), rounding_noise(i)])
    ok = host_rel < 1e-7 and db_rel < 1e-7 and not fd_bad
    return ok, {"n": n, "host_rel": host_rel, "db_rel": db_rel, "fd_entries": int(n_fd), "fd_worst_scaled": worst,
                "fd_solution_space_worst_over_noise": worst_sol, "g_inf": g_inf, "fd_bad": fd_bad}


def input_checks():
    rp, col, pos, masks = ts.diffusion_pattern_2d(8, 6)
    pat = ts.CsrPattern(rp, col)
    vals = torch.ones(int(rp[-1]), dtype=torch.float64, device="cuda")
    raised = []
    nonsym = ts.diffusion_values_2d(torch.ones(6, 8, dtype=torch.float64, device="cuda"), 8, 6, pos, masks)
    nonsym[int(pos["xp"][3])] -= 0.25  # one entry above the diagonal without its mirror
    for v, b in ((vals.float(), torch.ones(48, dtype=torch.float64, device="cuda")),
                 (vals, torch.ones(47, dtype=torch.float64, device="cuda")),
                 (nonsym, torch.ones(48, dtype=torch.float64, device="cuda"))):
        try:
            ts.sparse_solve(v, b, pat)
            raised.append(False)
        except ValueError:
            raised.append(True)
    return all(raised), {"raised": raised}


if __name__ == "__main__":
    check = {"gradients": gradients, "gradients_full": gradients_full, "input_checks": input_checks}[sys.argv[1]]
    ok, info = check()
    print(json.dumps({"ok": ok, **info}), flush=True)
    sys.exit(0 if ok else 1)
