"""Differentiable sparse solve: x = A(values)^{-1} b as a torch.autograd.Function.

The forward pass runs the device CG of libvclhip (vclhip_cg_solve_csr_f64) on torch's current
stream; the backward pass solves A^T lambda = dL/dx with the same CG, returns dL/db = lambda, and
forms dL/dA_ij = -lambda_i x_j on A's pattern with one kernel (vclhip_csr_adjoint_grad_f64).  CG
needs A symmetric positive definite, so A^T = A: the pattern's symmetry is checked once when the
CsrPattern is made and the values' symmetry in every forward pass (a nonsymmetric A raises
ValueError instead of returning a wrong gradient).  Both solves share one DIA detection and copy of
A (vclhip_dia_create_f64), made in the forward pass.  Gradients with respect to whatever produced
the values (a coefficient field, say) then follow from torch's autograd.  torch supplies device
memory and streams only; every solve and gradient runs in libvclhip.
"""
import ctypes as C

import numpy as np
import torch

from ._abi import Csr, SolverResult, call
from . import plan_blocks, solver_params, PRECOND_NONE, PRECOND_JACOBI


class CsrPattern:
    """Pattern of a square CSR matrix on the device: row_ptr, col and the row-block plan."""

    def __init__(self, row_ptr, col, device="cuda"):
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        cl = np.ascontiguousarray(col, dtype=np.uint32)
        self.n = rp.size - 1
        self.nnz = int(rp[-1])
        blocks = plan_blocks(rp)
        self.num_blocks = blocks.size - 1
        self.row_ptr_host = rp
        self.col_host = cl
        self.row_ptr = torch.from_numpy(rp.view(np.int32).copy()).to(device)
        self.col = torch.from_numpy((cl if cl.size else np.zeros(1, np.uint32)).view(np.int32).copy()).to(device)
        self.blocks = torch.from_numpy(blocks.view(np.int32).copy()).to(device)
        # transpose[k] = the entry (col(k), row(k)); None when some entry has no mirror
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(rp.astype(np.int64)))
        key = rows * self.n + cl.astype(np.int64)
        key_t = cl.astype(np.int64) * self.n + rows
        order = np.argsort(key, kind="stable")
        at = np.searchsorted(key, key_t, sorter=order) if key.size else np.zeros(0, np.int64)
        at = np.minimum(at, max(key.size - 1, 0))
        found = key[order[at]] == key_t if key.size else np.ones(0, bool)
        self.transpose = torch.from_numpy(order[at]).to(device) if bool(np.all(found)) else None

    def descriptor(self, values):
        return Csr(self.n, self.n, self.nnz, self.num_blocks, self.row_ptr.data_ptr(), self.col.data_ptr(),
                   values.data_ptr(), self.blocks.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(values, vec, pattern):
    if not (values.is_cuda and vec.is_cuda):
        raise ValueError("values and vectors must be device tensors")
    if values.dtype != torch.float64 or vec.dtype != torch.float64:
        raise ValueError("the differentiable solve runs in float64")
    if values.numel() != pattern.nnz or vec.numel() != pattern.n:
        raise ValueError("sizes do not match the pattern")


def check_symmetric(pattern, values):
    """Raises ValueError unless A = (pattern, values) equals its transpose."""
    if pattern.transpose is None:
        raise ValueError("the CG solve needs a symmetric matrix: the pattern is not symmetric")
    if not torch.equal(values, values[pattern.transpose]):
        raise ValueError("the CG solve needs a symmetric matrix: values[i, j] != values[j, i]")


class DiaCache:
    """The DIA detection and copy of A = (pattern, values) made once (vclhip_dia_create_f64) and
    shared by the solves with these values."""

    def __init__(self, pattern, values):
        self.handle = C.c_void_p()
        if not values.is_cuda or values.dtype != torch.float64 or values.numel() != pattern.nnz:
            raise ValueError("the DIA copy needs the float64 device values of the pattern")
        if not values.is_contiguous():
            raise ValueError("values must be contiguous")
        self.desc = pattern.descriptor(values)
        call("vclhip_dia_create_f64", C.byref(self.desc), C.byref(self.handle), _stream())

    def close(self):
        if self.handle:
            call("vclhip_dia_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cg_solve(pattern, values, b, tol=1e-10, max_iters=10000, precond=PRECOND_NONE, dia=None):
    """x = A^{-1} b by the device CG; returns (x, SolverResult).  dia: an optional DiaCache of these
    values (otherwise the solve detects and copies the diagonals itself)."""
    _check(values, b, pattern)
    values = values.contiguous()
    b = b.contiguous()
    x = torch.empty_like(b)
    desc = pattern.descriptor(values)
    params = solver_params(tol=tol, max_iters=max_iters)
    if dia is not None:
        params.dia = dia.handle
    res = SolverResult()
    call("vclhip_cg_solve_csr_f64", C.byref(desc), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), precond,
         C.byref(params), C.byref(res), _stream())
    return x, res


def adjoint_grad(pattern, values, lam, x):
    """dA[k] = -lam[row(k)] * x[col(k)] on the pattern (device kernel)."""
    _check(values, lam, pattern)
    dA = torch.empty_like(values)
    desc = pattern.descriptor(values)
    call("vclhip_csr_adjoint_grad_f64", C.byref(desc), C.c_void_p(lam.contiguous().data_ptr()),
         C.c_void_p(x.contiguous().data_ptr()), C.c_void_p(dA.data_ptr()), _stream())
    return dA


class SparseSolve(torch.autograd.Function):
    """x = A^{-1} b for a symmetric positive definite A given by (pattern, values)."""

    @staticmethod
    def forward(ctx, values, b, pattern, tol, max_iters, precond, check):
        _check(values, b, pattern)  # types and sizes before any device work on these arrays
        values = values.contiguous()
        if check:
            check_symmetric(pattern, values)
        dia = DiaCache(pattern, values)
        x, res = cg_solve(pattern, values, b, tol, max_iters, precond, dia)
        if not res.converged:
            dia.close()
            raise RuntimeError(f"forward CG did not reach {tol} in {max_iters} iterations (error {res.error:.3e})")
        ctx.save_for_backward(values, x)
        ctx.pattern, ctx.tol, ctx.max_iters, ctx.precond = pattern, tol, max_iters, precond
        ctx.forward_iters = res.iters
        ctx.dia = dia
        return x

    @staticmethod
    def backward(ctx, gx):
        values, x = ctx.saved_tensors
        lam, res = cg_solve(ctx.pattern, values, gx.contiguous(), ctx.tol, ctx.max_iters, ctx.precond, ctx.dia)
        ctx.dia.close()
        if not res.converged:
            raise RuntimeError(f"adjoint CG did not reach {ctx.tol} in {ctx.max_iters} iterations")
        dA = adjoint_grad(ctx.pattern, values, lam, x) if ctx.needs_input_grad[0] else None
        return dA, lam, None, None, None, None, None


def sparse_solve(values, b, pattern, tol=1e-10, max_iters=10000, jacobi=False, check_symmetric=True):
    """Differentiable x = A(values)^{-1} b (A symmetric positive definite; check_symmetric=False skips
    the per-call value check for callers that build A symmetric by construction)."""
    return SparseSolve.apply(values, b, pattern, tol, max_iters, PRECOND_JACOBI if jacobi else PRECOND_NONE,
                             bool(check_symmetric))


# ------------------------------------------------------------------ 2-D diffusion operator in torch

def diffusion_pattern_2d(nx, ny):
    """CSR pattern of the 2-D 5-point diffusion operator (the order of vclhip_gen_diffusion with
    nz = 1: row = x + nx*y; entries y-1, x-1, diagonal, x+1, y+1) and the value positions of each
    entry kind."""
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    has_ym, has_xm = y > 0, x > 0
    has_xp, has_yp = x < nx - 1, y < ny - 1
    cnt = 1 + has_ym.astype(np.int64) + has_xm + has_xp + has_yp
    rp = np.zeros(nx * ny + 1, np.int64)
    rp[1:] = np.cumsum(cnt.ravel())
    start = rp[:-1].reshape(ny, nx)
    pos = {}
    pos["d"] = start + has_ym + has_xm
    pos["ym"] = start[has_ym]
    pos["xm"] = (start + has_ym)[has_xm]
    pos["xp"] = (pos["d"] + 1)[has_xp]
    pos["yp"] = (pos["d"] + 1 + has_xp)[has_yp]
    pos["d"] = pos["d"].ravel()
    col = np.zeros(rp[-1], np.int64)
    row = (x + nx * y)
    col[pos["d"]] = row.ravel()
    col[pos["ym"]] = (row - nx)[has_ym]
    col[pos["xm"]] = (row - 1)[has_xm]
    col[pos["xp"]] = (row + 1)[has_xp]
    col[pos["yp"]] = (row + nx)[has_yp]
    masks = {"ym": has_ym, "xm": has_xm, "xp": has_xp, "yp": has_yp}
    return rp.astype(np.uint32), col.astype(np.uint32), pos, masks


def diffusion_values_2d(theta, nx, ny, pos, masks):
    """Values of -div(theta grad u) (harmonic-mean faces, Dirichlet boundary faces add the cell's
    theta to the diagonal) on the pattern of diffusion_pattern_2d, differentiable in theta
    (shape (ny, nx)).  Same arithmetic order as vclhip_gen_diffusion."""
    dev = theta.device
    t = theta.reshape(ny, nx)

    def face(a, b):
        return 2.0 * a * b / (a + b)

    def shifted(dy, dx):
        # theta of the neighbour (y+dy, x+dx), or the cell's own theta where it does not exist
        nb = torch.roll(t, shifts=(-dy, -dx), dims=(0, 1))
        return nb

    t_ym = torch.where(torch.as_tensor(masks["ym"], device=dev), face(t, shifted(-1, 0)), t)
    t_xm = torch.where(torch.as_tensor(masks["xm"], device=dev), face(t, shifted(0, -1)), t)
    t_xp = torch.where(torch.as_tensor(masks["xp"], device=dev), face(t, shifted(0, 1)), t)
    t_yp = torch.where(torch.as_tensor(masks["yp"], device=dev), face(t, shifted(1, 0)), t)
    diag = torch.zeros_like(t) + t_ym
    diag = diag + t_xm
    diag = diag + t_xp
    diag = diag + t_yp
    nnz = int(pos["d"].size + pos["ym"].size + pos["xm"].size + pos["xp"].size + pos["yp"].size)
    vals = torch.zeros(nnz, dtype=theta.dtype, device=dev)

    def put(v, key, src):
        idx = torch.as_tensor(pos[key], device=dev, dtype=torch.int64)
        return v.index_put((idx,), src)

    vals = put(vals, "d", diag.reshape(-1))
    for key, tv in (("ym", t_ym), ("xm", t_xm), ("xp", t_xp), ("yp", t_yp)):
        m = torch.as_tensor(masks[key], device=dev)
        vals = put(vals, key, -tv[m])
    return vals
