"""GPU checks of the differentiable solve (vclhip.torch_solve.SparseSolve).

The adjoint kernel dA[k] = -lambda[row(k)] * x[col(k)] is checked bit-exact against the oracle
here.  The torch checks run in a child process (tests/autograd_checks.py) that imports torch
before libvclhip, so that it holds a single HIP runtime: dL/dtheta of L = 0.5 ||x(theta) - x_t||^2,
x = A(theta)^{-1} b on the 2-D variable-coefficient diffusion operator, against central finite
differences (h = 1e-6) and against the gradient formed from the FP64 oracle (host CG for x and
lambda, host adjoint product) chained through the same torch assembly on the CPU, at 48 x 40 and at
the full size of config C5 (512 x 512, 64 entries); and that a nonsymmetric A is rejected.
"""
import ctypes as C
import json
import pathlib
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHECKS = pathlib.Path(__file__).resolve().parent / "autograd_checks.py"


def test_adjoint_grad_kernel_bit_exact(dev, vh, oracle):
    # power-law rows reach the long-row blocks (> 2048 entries)
    rp, col, val = vh.powerlaw(20000, 4000, 1.6, 1.1, 7)
    n = rp.size - 1
    lam = vh.random_uniform(3, n)
    x = vh.random_uniform(4, n)
    A = vh.CsrMatrix(rp, col, val)
    dl = vh.DeviceArray.from_numpy(lam)
    dx = vh.DeviceArray.from_numpy(x)
    dA = vh.DeviceArray.zeros(rp[-1], np.float64)
    vh.call("vclhip_csr_adjoint_grad_f64", C.byref(A.desc), dl.ptr, dx.ptr, dA.ptr, None)
    assert np.array_equal(dA.to_numpy(), oracle.adjoint_grad(rp, col, lam, x))


@pytest.mark.parametrize("check", ["gradients", "gradients_full", "input_checks"])
def test_torch_checks(dev, check):
    print(f"running {CHECKS.name} {check}", flush=True)
    p = subprocess.run([sys.executable, str(CHECKS), check], capture_output=True, text=True, timeout=300)
    print(p.stdout[-3000:], p.stderr[-3000:])
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    assert json.loads(p.stdout.strip().splitlines()[-1])["ok"]
