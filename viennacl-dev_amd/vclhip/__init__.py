"""Python binding of libvclhip (the C ABI in include/vclhip.h).

Thin layer for tests, benchmarks and the torch autograd hook: device buffers, CSR/SELL matrices
on the device, and the SpMV / vector / solver entry points.  Every compute call goes through the
HIP library; nothing here computes on the CPU.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import (Csr, Sell, Coo, Ell, Hyb, Matrix, SolverParams, SolverResult, MONITOR, VclHipError, call, lib,
                   PRECOND_NONE, PRECOND_JACOBI, LIB_PATH, declared_functions, OPERATOR_AUTO, OPERATOR_AS_GIVEN,
                   OPERATOR_DIA_UNCODED, USED_AS_GIVEN, USED_DIA, USED_DIA_CODED, INNER_AUTO, INNER_CSR, INNER_SELL64, INNER_SELL64_D16, INNER_DIA, INNER_DIA_CODED,
                   FORMAT_CSR, FORMAT_SELL, FORMAT_COO, FORMAT_ELL, FORMAT_HYB, ERROR_INVALID_VALUE,
                   ERROR_ZERO_DIAGONAL, ERROR_NOT_SUPPORTED, ERROR_NO_DEVICE, ERROR_TOO_LARGE)

__all__ = ["DeviceArray", "Stream", "CsrMatrix", "DiaOperator", "SellMatrix", "CooMatrix", "EllMatrix", "HybMatrix", "BatchPlan",
           "lib", "call", "VclHipError",
           "device_count", "PRECOND_NONE", "PRECOND_JACOBI", "SolverParams", "SolverResult"]


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def device_count():
    c = C.c_int(0)
    call("vclhip_device_count", C.byref(c))
    return c.value


class Stream:
    """A non-blocking HIP stream owned by the caller (None = null stream)."""

    def __init__(self):
        h = C.c_void_p()
        call("vclhip_stream_create", C.byref(h))
        self.handle = h

    def synchronize(self):
        call("vclhip_stream_synchronize", self.handle)

    def close(self):
        if self.handle:
            call("vclhip_stream_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation holding `size` elements of `dtype`."""

    def __init__(self, size, dtype):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        call("vclhip_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        call("vclhip_memcpy_h2d", d.ptr, _ptr(a), C.c_size_t(d.nbytes), stream)
        return d

    @classmethod
    def zeros(cls, size, dtype, stream=None):
        d = cls(size, dtype)
        call("vclhip_fill_zero", d.ptr, C.c_size_t(d.nbytes), stream)
        return d

    def to_numpy(self, stream=None):
        out = np.empty(self.size, self.dtype)
        call("vclhip_memcpy_d2h", _ptr(out), self.ptr, C.c_size_t(self.nbytes), stream)
        return out

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.size
        call("vclhip_memcpy_h2d", self.ptr, _ptr(a), C.c_size_t(self.nbytes), stream)

    def free(self):
        if self.ptr:
            call("vclhip_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def plan_blocks(row_ptr):
    """Row-block plan of the CSR kernels (host)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    n = row_ptr.size - 1
    blocks = np.zeros(n + 1, np.uint32)
    nb = C.c_uint32()
    call("vclhip_csr_plan_blocks", _ptr(row_ptr), n, _ptr(blocks), C.byref(nb))
    return blocks[: nb.value + 1].copy()


def reference_row_blocks(row_ptr):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    n = row_ptr.size - 1
    out = np.zeros(n + 1, np.uint32)
    nb = C.c_uint32()
    call("vclhip_csr_reference_row_blocks", _ptr(row_ptr), n, _ptr(out), C.byref(nb))
    return out[: nb.value + 1].copy()


def sell_from_csr(row_ptr, col, val, C_):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    n = row_ptr.size - 1
    ns = C.c_uint32()
    stored = C.c_uint64()
    call("vclhip_sell_from_csr_f64", n, C_, _ptr(row_ptr), _ptr(col), _ptr(val), C.byref(ns), C.byref(stored),
         None, None, None, None)
    cpb = np.zeros(ns.value, np.uint32)
    bs = np.zeros(ns.value, np.uint32)
    ci = np.zeros(max(stored.value, 1), np.uint32)
    el = np.zeros(max(stored.value, 1), np.float64)
    call("vclhip_sell_from_csr_f64", n, C_, _ptr(row_ptr), _ptr(col), _ptr(val), C.byref(ns), C.byref(stored),
         _ptr(cpb), _ptr(ci), _ptr(bs), _ptr(el))
    return cpb, ci[: stored.value], bs, el[: stored.value]


class CsrMatrix:
    """compressed_matrix on the device: row_ptr, col, values (float64 or float32) and the block plan."""

    def __init__(self, row_ptr, col, val, dtype=np.float64, n_cols=None, stream=None):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=dtype)
        self.n = row_ptr.size - 1
        self.n_cols = self.n if n_cols is None else n_cols
        self.nnz = int(row_ptr[-1])
        self.dtype = np.dtype(dtype)
        self.blocks_host = plan_blocks(row_ptr)
        self.d_row_ptr = DeviceArray.from_numpy(row_ptr, stream)
        self.d_col = DeviceArray.from_numpy(col if col.size else np.zeros(1, np.uint32), stream)
        self.d_val = DeviceArray.from_numpy(val if val.size else np.zeros(1, dtype), stream)
        self.d_blocks = DeviceArray.from_numpy(self.blocks_host, stream)
        self.desc = Csr(self.n, self.n_cols, self.nnz, self.blocks_host.size - 1, self.d_row_ptr.ptr,
                        self.d_col.ptr, self.d_val.ptr, self.d_blocks.ptr)

    @property
    def num_blocks(self):
        return self.blocks_host.size - 1

    def free(self):
        for a in (self.d_row_ptr, self.d_col, self.d_val, self.d_blocks):
            a.free()

    def spmv(self, x, y, alpha=1.0, beta=0.0, xs=0, xi=1, ys=0, yi=1, stream=None):
        fn = "vclhip_csr_spmv_f64" if self.dtype == np.float64 else "vclhip_csr_spmv_f32"
        call(fn, C.byref(self.desc), x.ptr, xs, xi, alpha, y.ptr, ys, yi, beta, stream)

    def row_info(self, out, info, stream=None):
        fn = "vclhip_csr_row_info_f64" if self.dtype == np.float64 else "vclhip_csr_row_info_f32"
        call(fn, C.byref(self.desc), out.ptr, info, stream)


class SellMatrix:
    """sliced_ell_matrix (sigma = 1) on the device, built from CSR arrays."""

    def __init__(self, row_ptr, col, val, C_=64, dtype=np.float64, stream=None):
        cpb, ci, bs, el = sell_from_csr(row_ptr, col, val, C_)
        self.n = len(row_ptr) - 1
        self.C = C_
        self.dtype = np.dtype(dtype)
        self.stored = el.size
        self.host = (cpb, ci, bs, el)
        self.d_cpb = DeviceArray.from_numpy(cpb, stream)
        self.d_ci = DeviceArray.from_numpy(ci if ci.size else np.zeros(1, np.uint32), stream)
        self.d_bs = DeviceArray.from_numpy(bs, stream)
        self.d_el = DeviceArray.from_numpy((el if el.size else np.zeros(1)).astype(dtype), stream)
        self.desc = Sell(self.n, self.n, C_, cpb.size, self.d_cpb.ptr, self.d_ci.ptr, self.d_bs.ptr, self.d_el.ptr)

    def spmv(self, x, y, alpha=1.0, beta=0.0, xs=0, xi=1, ys=0, yi=1, stream=None):
        fn = "vclhip_sell_spmv_f64" if self.dtype == np.float64 else "vclhip_sell_spmv_f32"
        call(fn, C.byref(self.desc), x.ptr, xs, xi, alpha, y.ptr, ys, yi, beta, stream)

    def free(self):
        for a in (self.d_cpb, self.d_ci, self.d_bs, self.d_el):
            a.free()


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class CooMatrix:
    """coordinate_matrix on the device: (row, col) pairs in row order, values, and the row pointers
    and CSR-kernel plan of that order (vclhip_coo_plan)."""

    format = FORMAT_COO

    def __init__(self, row_ptr, col, val, dtype=np.float64, n_cols=None, stream=None):
        rp = _u32(row_ptr)
        n = rp.size - 1
        self.n, self.nnz, self.dtype = n, int(rp[-1]), np.dtype(dtype)
        rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp.astype(np.int64)))
        coords = np.empty(2 * self.nnz, np.uint32)
        coords[0::2], coords[1::2] = rows, _u32(col)
        self.coords_host = coords
        prp = np.zeros(n + 1, np.uint32)
        blocks = np.zeros(n + 1, np.uint32)
        nb = C.c_uint32()
        call("vclhip_coo_plan", _ptr(coords), self.nnz, n, _ptr(prp), _ptr(blocks), C.byref(nb))
        self.d_coords = DeviceArray.from_numpy(coords if coords.size else np.zeros(2, np.uint32), stream)
        vv = np.ascontiguousarray(val, dtype=dtype)
        self.d_el = DeviceArray.from_numpy(vv if vv.size else np.zeros(1, dtype), stream)
        self.d_rp = DeviceArray.from_numpy(prp, stream)
        self.d_blocks = DeviceArray.from_numpy(blocks[: nb.value + 1], stream)
        self.desc = Coo(n, n if n_cols is None else n_cols, self.nnz, nb.value, self.d_coords.ptr, self.d_el.ptr,
                        self.d_rp.ptr, self.d_blocks.ptr)

    def spmv(self, x, y, alpha=1.0, beta=0.0, xs=0, xi=1, ys=0, yi=1, stream=None):
        fn = "vclhip_coo_spmv_f64" if self.dtype == np.float64 else "vclhip_coo_spmv_f32"
        call(fn, C.byref(self.desc), x.ptr, xs, xi, alpha, y.ptr, ys, yi, beta, stream)

    def free(self):
        for a in (self.d_coords, self.d_el, self.d_rp, self.d_blocks):
            a.free()


def ell_from_csr(row_ptr, col, val, align=1):
    """Host CSR -> the ell_matrix arrays (ld, width, coords, elements)."""
    rp, cl, vl = _u32(row_ptr), _u32(col), np.ascontiguousarray(val, dtype=np.float64)
    n = rp.size - 1
    w = C.c_uint32()
    call("vclhip_ell_from_csr_f64", n, align, _ptr(rp), _ptr(cl), _ptr(vl), C.byref(w), None, None)
    ld = (n + align - 1) // align * align
    coords = np.zeros(max(ld * w.value, 1), np.uint32)
    el = np.zeros(max(ld * w.value, 1))
    call("vclhip_ell_from_csr_f64", n, align, _ptr(rp), _ptr(cl), _ptr(vl), C.byref(w), _ptr(coords), _ptr(el))
    return ld, w.value, coords, el


def hyb_from_csr(row_ptr, col, val, csr_threshold=0.8, align=1):
    """Host CSR -> the hyb_matrix arrays (ld, width, ell coords, ell elements, csr rows, cols, elements)."""
    rp, cl, vl = _u32(row_ptr), _u32(col), np.ascontiguousarray(val, dtype=np.float64)
    n = rp.size - 1
    w, cn = C.c_uint32(), C.c_uint32()
    args = (n, csr_threshold, align, _ptr(rp), _ptr(cl), _ptr(vl), C.byref(w), C.byref(cn))
    call("vclhip_hyb_from_csr_f64", *args, None, None, None, None, None)
    ld = (n + align - 1) // align * align
    ec, ee = np.zeros(max(ld * w.value, 1), np.uint32), np.zeros(max(ld * w.value, 1))
    cr, cc, ce = np.zeros(n + 1, np.uint32), np.zeros(cn.value, np.uint32), np.zeros(cn.value)
    call("vclhip_hyb_from_csr_f64", *args, _ptr(ec), _ptr(ee), _ptr(cr), _ptr(cc), _ptr(ce))
    return ld, w.value, ec, ee, cr, cc, ce


class EllMatrix:
    """ell_matrix on the device (column-major ELL, internal_size1 rows per column)."""

    format = FORMAT_ELL

    def __init__(self, row_ptr, col, val, dtype=np.float64, align=1, stream=None):
        n = len(row_ptr) - 1
        ld, w, coords, el = ell_from_csr(row_ptr, col, val, align)
        self.n, self.dtype, self.ld, self.width = n, np.dtype(dtype), ld, w
        self.host = (coords, el)
        self.d_coords = DeviceArray.from_numpy(coords, stream)
        self.d_el = DeviceArray.from_numpy(el.astype(dtype), stream)
        self.desc = Ell(n, n, ld, w, self.d_coords.ptr, self.d_el.ptr)

    def spmv(self, x, y, alpha=1.0, beta=0.0, xs=0, xi=1, ys=0, yi=1, stream=None):
        fn = "vclhip_ell_spmv_f64" if self.dtype == np.float64 else "vclhip_ell_spmv_f32"
        call(fn, C.byref(self.desc), x.ptr, xs, xi, alpha, y.ptr, ys, yi, beta, stream)

    def free(self):
        for a in (self.d_coords, self.d_el):
            a.free()


class HybMatrix:
    """hyb_matrix on the device: ELL part plus a CSR part for the entries beyond the ELL width.

    plan=True uploads the row-block plan of the CSR part, so that long CSR parts are summed by a
    wavefront; plan=False sums every row in one lane, in the reference's order; plan="auto" (the
    default) takes the plan when some row's CSR part is longer than 32 entries."""

    format = FORMAT_HYB

    def __init__(self, row_ptr, col, val, dtype=np.float64, csr_threshold=0.8, align=1, stream=None, plan="auto"):
        n = len(row_ptr) - 1
        ld, w, ec, ee, cr, cc, ce = hyb_from_csr(row_ptr, col, val, csr_threshold, align)
        self.n, self.dtype, self.ld, self.width, self.csr_nnz = n, np.dtype(dtype), ld, w, cc.size
        self.host = (ec, ee, cr, cc, ce)
        arrays = [ec, ee.astype(dtype), cr, cc, ce.astype(dtype)]
        if plan == "auto":
            plan = bool(n) and int(np.diff(cr.astype(np.int64)).max()) > 32
        self.planned = bool(plan)
        nb = 0
        if plan:
            blocks = plan_blocks(cr)
            nb = blocks.size - 1
            arrays.append(blocks)
        self.d = [DeviceArray.from_numpy(a, stream) for a in arrays]
        self.desc = Hyb(n, n, ld, w, *(a.ptr for a in self.d[:5]), nb, self.d[5].ptr if plan else None)

    def spmv(self, x, y, alpha=1.0, beta=0.0, xs=0, xi=1, ys=0, yi=1, stream=None):
        fn = "vclhip_hyb_spmv_f64" if self.dtype == np.float64 else "vclhip_hyb_spmv_f32"
        call(fn, C.byref(self.desc), x.ptr, xs, xi, alpha, y.ptr, ys, yi, beta, stream)

    def free(self):
        for a in self.d:
            a.free()


def generic(A):
    """The format-generic descriptor (vclhip_matrix) of a device matrix; keep the result alive."""
    fmt = getattr(A, "format", FORMAT_SELL if isinstance(A, SellMatrix) else FORMAT_CSR)
    return Matrix(fmt, C.cast(C.pointer(A.desc), C.c_void_p))


def matrix_diag(A, d, stream=None):
    m = generic(A)
    fn = "vclhip_matrix_diag_f32" if A.dtype == np.float32 else "vclhip_matrix_diag_f64"
    call(fn, C.byref(m), d.ptr, stream)


class DiaOperator:
    """The solvers' DIA detection and copy of a CSR matrix made once (vclhip_dia_create_*), for
    repeated solves of the same matrix: pass it as `dia=` to cg / bicgstab / gmres.  `built` says
    whether the matrix qualified (then the solves run on the copy), `nd` its diagonal count.  The
    matrix's arrays and values must not change while it is used."""

    def __init__(self, A, stream=None):
        self.handle = C.c_void_p()
        fn = "vclhip_dia_create_f64" if A.dtype == np.float64 else "vclhip_dia_create_f32"
        desc = A.desc if hasattr(A, "desc") else A
        call(fn, C.byref(desc), C.byref(self.handle), stream)
        built, nd = C.c_int(), C.c_uint32()
        call("vclhip_dia_info", self.handle, C.byref(built), C.byref(nd))
        self.built, self.nd = bool(built.value), int(nd.value)

    def close(self):
        if self.handle:
            call("vclhip_dia_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solver_params(tol=1e-8, max_iters=300, abs_tol=0.0, max_iters_before_restart=200, inner_tol=1e-2,
                  chunk=16, use_graph=False, mixed_mode=0, variant=0, history=None, monitor=None, inner_format=0,
                  krylov_dim=20, operator_format=0, dia=None):
    p = SolverParams()
    lib().vclhip_solver_params_default(C.byref(p))
    p.variant = variant
    p.tol = tol
    p.abs_tol = abs_tol
    p.max_iters = max_iters
    p.max_iters_before_restart = max_iters_before_restart
    p.inner_tol = inner_tol
    p.chunk = chunk
    p.use_graph = 1 if use_graph else 0
    p.mixed_mode = mixed_mode
    p.inner_format = inner_format
    p.krylov_dim = krylov_dim
    p.operator_format = operator_format
    if dia is not None:
        p.dia = dia.handle
    if history is not None:
        p.history_host = history.ctypes.data_as(C.POINTER(C.c_double))
        p.history_len = history.size
    if monitor is not None:
        p.monitor = monitor
    return p


def _check_vectors(A, *vecs):
    """The C ABI takes plain device pointers, so a right-hand side or solution of another precision or
    shorter than the matrix would be read past its end: refuse it on the host."""
    for v in vecs:
        if v.dtype != A.dtype:
            raise ValueError(f"vector of dtype {v.dtype} for a matrix of dtype {A.dtype}")
        if v.size < A.n:
            raise ValueError(f"vector of {v.size} elements for a matrix of {A.n} rows")


def _solve(kind, A, b, x, precond, params, stream):
    _check_vectors(A, b, x)
    res = SolverResult()
    if isinstance(A, (CsrMatrix, SellMatrix)) and A.dtype == np.float64:
        mat = "csr" if isinstance(A, CsrMatrix) else "sell"
        call(f"vclhip_{kind}_solve_{mat}_f64", C.byref(A.desc), b.ptr, x.ptr, precond, C.byref(params), C.byref(res),
             stream)
    else:
        m = generic(A)
        prec = "f32" if A.dtype == np.float32 else "f64"
        call(f"vclhip_{kind}_solve_{prec}", C.byref(m), b.ptr, x.ptr, precond, C.byref(params), C.byref(res), stream)
    return res


def cg(A, b, x, precond=PRECOND_NONE, stream=None, **kw):
    """solve(A, b, cg_tag [, jacobi_precond]); returns SolverResult (iters, error, converged)."""
    return _solve("cg", A, b, x, precond, solver_params(**kw), stream)


def bicgstab(A, b, x, precond=PRECOND_NONE, stream=None, **kw):
    kw.setdefault("max_iters", 400)
    return _solve("bicgstab", A, b, x, precond, solver_params(**kw), stream)


def mixed_cg(A, b, x, precond=PRECOND_NONE, values32=None, stream=None, **kw):
    """solve(A, b, mixed_precision_cg_tag): FP32 inner CG, FP64 corrections (mixed_mode 0 = the
    reference's restart schedule, 1 = reliable updates).  A holds FP64 values; values32 is an
    optional DeviceArray with their FP32 copy.  result.restarts counts the FP64 corrections."""
    _check_vectors(A, b, x)
    params = solver_params(**kw)
    res = SolverResult()
    mat = "csr" if isinstance(A, CsrMatrix) else "sell"
    call(f"vclhip_mixed_cg_solve_{mat}", C.byref(A.desc), values32.ptr if values32 is not None else None, b.ptr,
         x.ptr, precond, C.byref(params), C.byref(res), stream)
    return res


def gmres(A, b, x, stream=None, **kw):
    """solve(A, b, gmres_tag(tol, max_iters, krylov_dim)) without a preconditioner; returns
    SolverResult (iters = Krylov steps, error = the last cycle's estimate, restarts = cycles run).
    The gmres_tag defaults are tol=1e-10, max_iters=300, krylov_dim=20."""
    kw.setdefault("tol", 1e-10)
    _check_vectors(A, b, x)
    params = solver_params(**kw)
    res = SolverResult()
    prec = "f32" if A.dtype == np.float32 else "f64"
    if isinstance(A, (CsrMatrix, SellMatrix)):
        mat = "csr" if isinstance(A, CsrMatrix) else "sell"
        call(f"vclhip_gmres_solve_{mat}_{prec}", C.byref(A.desc), b.ptr, x.ptr, C.byref(params), C.byref(res), stream)
    else:
        m = generic(A)
        call(f"vclhip_gmres_solve_{prec}", C.byref(m), b.ptr, x.ptr, C.byref(params), C.byref(res), stream)
    return res


class BatchPlan:
    """Launch order and lanes-per-row table of a batch of systems (vclhip_batch_plan)."""

    def __init__(self, sys_ptr, row_ptr=None):
        self.sys_ptr = np.ascontiguousarray(sys_ptr, dtype=np.uint32)
        self.num_systems = self.sys_ptr.size - 1
        rp = None if row_ptr is None else np.ascontiguousarray(row_ptr, dtype=np.uint32)
        h = C.c_void_p()
        call("vclhip_batch_plan_create", self.num_systems, _ptr(self.sys_ptr), _ptr(rp) if rp is not None else None,
             C.byref(h))
        self.handle = h
        self.d_sys_ptr = DeviceArray.from_numpy(self.sys_ptr)

    def autotune(self, A, single=False, stream=None):
        call("vclhip_batch_plan_autotune", self.handle, C.byref(A.desc), self.d_sys_ptr.ptr, int(bool(single)), stream)

    def lanes(self, single=False):
        out = np.zeros(16, np.int32)
        call("vclhip_batch_plan_lanes", self.handle, int(bool(single)), _ptr(out))
        return out

    def solve(self, A, b, x, stream=None, **kw):
        """Batched CG of every system (device arrays b, x); returns per-system iters, errors, converged."""
        ns = self.num_systems
        it = DeviceArray.zeros(ns, np.uint32)
        er = DeviceArray.zeros(ns, np.float64)
        cv = DeviceArray.zeros(ns, np.uint32)
        self.solve_into(A, b, x, it, er, cv, stream, **kw)
        call("vclhip_device_synchronize")
        return it.to_numpy(), er.to_numpy(), cv.to_numpy()

    def solve_into(self, A, b, x, iters, errors, converged, stream=None, **kw):
        params = solver_params(**kw)
        fn = "vclhip_batched_cg_solve_f32" if A.dtype == np.float32 else "vclhip_batched_cg_solve_f64"
        call(fn, self.handle, C.byref(A.desc), self.d_sys_ptr.ptr, b.ptr, x.ptr, C.byref(params), iters.ptr,
             errors.ptr, converged.ptr, stream)

    def close(self):
        if self.handle:
            call("vclhip_batch_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ MatrixMarket files (host)

def read_matrix_market(path, index_base=1):
    """io::read_matrix_market_file: returns (row_ptr, col, val, n_cols, symmetric, lines_read)."""
    h = C.c_void_p()
    nr, nc, nz, sym, lines = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int(), C.c_int64()
    call("vclhip_mm_read", str(path).encode(), index_base, C.byref(h), C.byref(nr), C.byref(nc), C.byref(nz),
         C.byref(sym), C.byref(lines))
    try:
        rp = np.zeros(nr.value + 1, np.uint32)
        col = np.zeros(max(nz.value, 1), np.uint32)
        val = np.zeros(max(nz.value, 1), np.float64)
        call("vclhip_mm_csr", h, _ptr(rp), _ptr(col), _ptr(val))
    finally:
        call("vclhip_mm_free", h)
    return rp, col[: nz.value], val[: nz.value], nc.value, bool(sym.value), lines.value


def write_matrix_market(path, row_ptr, col, val, n_cols=None, index_base=1):
    """io::write_matrix_market_file for CSR host arrays."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    n = rp.size - 1
    call("vclhip_mm_write", str(path).encode(), index_base, n, n if n_cols is None else n_cols, _ptr(rp),
         _ptr(col) if col.size else None, _ptr(val) if val.size else None)


# ------------------------------------------------------------------ generators (host)

def random_uniform(seed, n, a=-1.0, b=1.0):
    out = np.empty(n, np.float64)
    call("vclhip_random_uniform", C.c_uint64(seed), C.c_uint64(n), a, b, _ptr(out))
    return out


def random_normal(seed, n, mean=0.0, sigma=1.0):
    out = np.empty(n, np.float64)
    call("vclhip_random_normal", C.c_uint64(seed), C.c_uint64(n), mean, sigma, _ptr(out))
    return out


def fdm_laplace_2d(px, py):
    n = px * py
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(5 * n, np.uint32)
    val = np.zeros(5 * n, np.float64)
    nnz = C.c_uint32()
    call("vclhip_gen_fdm_laplace_2d", px, py, _ptr(rp), _ptr(col), _ptr(val), C.byref(nnz))
    return rp, col[: nnz.value].copy(), val[: nnz.value].copy()


def laplace_3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(7 * n, np.uint32)
    val = np.zeros(7 * n, np.float64)
    nnz = C.c_uint32()
    call("vclhip_gen_laplace_3d", nx, ny, nz, _ptr(rp), _ptr(col), _ptr(val), C.byref(nnz))
    return rp, col[: nnz.value].copy(), val[: nnz.value].copy()


def diffusion(nx, ny, nz, k):
    k = np.ascontiguousarray(k, dtype=np.float64)
    n = nx * ny * nz
    assert k.size == n
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(7 * n, np.uint32)
    val = np.zeros(7 * n, np.float64)
    nnz = C.c_uint32()
    call("vclhip_gen_diffusion", nx, ny, nz, _ptr(k), _ptr(rp), _ptr(col), _ptr(val), C.byref(nnz))
    return rp, col[: nnz.value].copy(), val[: nnz.value].copy()


def convection_diffusion(nx, ny, nz, beta=(0.5, 0.3, 0.2)):
    """Nonsymmetric test matrix: the 7-point Laplacian of laplace_3d with first-order upwind
    convection of velocity beta (grid units).  Row i: (6 + sum beta) on the diagonal, -1 - beta_d
    for the upstream neighbour along axis d and -1 for the downstream one; diagonally dominant."""
    rp, col, val = laplace_3d(nx, ny, nz)
    row = np.repeat(np.arange(rp.size - 1, dtype=np.int64), np.diff(rp.astype(np.int64)))
    d = col.astype(np.int64) - row
    val = val.copy()
    for step, bd in ((1, beta[0]), (nx, beta[1]), (nx * ny, beta[2])):
        val[d == -step] -= bd
    val[d == 0] += float(sum(beta))
    return rp, col, val


def powerlaw(n, max_len=512, alpha=2.0, diag_factor=1.1, seed=1):
    nnz = C.c_uint32()
    call("vclhip_gen_powerlaw", n, max_len, alpha, diag_factor, C.c_uint64(seed), None, None, None, C.byref(nnz))
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(nnz.value, np.uint32)
    val = np.zeros(nnz.value, np.float64)
    call("vclhip_gen_powerlaw", n, max_len, alpha, diag_factor, C.c_uint64(seed), _ptr(rp), _ptr(col), _ptr(val),
         C.byref(nnz))
    return rp, col, val


def graph_laplacian_batch(num_systems, n_min=64, n_max=4096, avg_degree=8.0, w_lo=0.1, w_hi=1.0, shift=0.01, seed=1):
    rows = C.c_uint32()
    nnz = C.c_uint32()
    args = (num_systems, n_min, n_max, avg_degree, w_lo, w_hi, shift, C.c_uint64(seed))
    call("vclhip_gen_graph_laplacian_batch", *args, None, None, None, None, C.byref(rows), C.byref(nnz))
    sys_ptr = np.zeros(num_systems + 1, np.uint32)
    rp = np.zeros(rows.value + 1, np.uint32)
    col = np.zeros(nnz.value, np.uint32)
    val = np.zeros(nnz.value, np.float64)
    call("vclhip_gen_graph_laplacian_batch", *args, _ptr(sys_ptr), _ptr(rp), _ptr(col), _ptr(val), C.byref(rows),
         C.byref(nnz))
    return sys_ptr, rp, col, val
