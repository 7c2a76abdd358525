"""ctypes binding of the CPU oracle (liboracle.so) and of the reference build (_ref/libvclref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, never by the product.  See vcl_oracle.h for what each function restates.
"""
import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libvclref.so"

u32, u64, f32, f64, vp, sz, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_void_p, C.c_size_t, C.c_int
P = C.POINTER


class _Csr(C.Structure):
    _fields_ = [("n_rows", u32), ("row_ptr", vp), ("col", vp), ("val", vp)]


class _Sell(C.Structure):
    _fields_ = [("n_rows", u32), ("C", u32), ("cpb", vp), ("ci", vp), ("bs", vp), ("el", vp)]


class _Res(C.Structure):
    _fields_ = [("iters", u32), ("error", f64)]


class _Fmt(C.Structure):
    _fields_ = [("format", i32), ("n_rows", u32), ("nnz", u32), ("ld", u32), ("width", u32), ("coords", vp),
                ("elements", vp), ("csr_rows", vp), ("csr_cols", vp), ("csr_elements", vp)]


_ORC_SIGS = {
    "orc_set_num_threads": (None, [i32]),
    "orc_get_num_threads": (i32, []),
    "orc_set_reduction_blocks": (None, [i32]),
    "orc_csr_spmv_f64": (None, [u32, vp, vp, vp, vp, sz, sz, f64, vp, sz, sz, f64]),
    "orc_csr_spmv_f32": (None, [u32, vp, vp, vp, vp, sz, sz, f32, vp, sz, sz, f32]),
    "orc_sell_spmv_f64": (None, [u32, u32, vp, vp, vp, vp, vp, sz, sz, f64, vp, sz, sz, f64]),
    "orc_sell_spmv_f32": (None, [u32, u32, vp, vp, vp, vp, vp, sz, sz, f32, vp, sz, sz, f32]),
    "orc_sell_sizes": (None, [u32, u32, vp, P(u32), P(u64)]),
    "orc_sell_from_csr_f64": (None, [u32, u32, vp, vp, vp, vp, vp, vp, vp]),
    "orc_row_blocks": (u32, [u32, vp, vp]),
    "orc_csr_row_info_f64": (None, [u32, vp, vp, vp, vp, i32]),
    "orc_dot_f64": (f64, [sz, vp, vp]),
    "orc_dot_f32": (f32, [sz, vp, vp]),
    "orc_norm2_f64": (f64, [sz, vp]),
    "orc_norm2_f32": (f32, [sz, vp]),
    "orc_pipelined_prod_csr_f64": (None, [u32, vp, vp, vp, vp, vp, vp, vp, sz, sz]),
    "orc_pipelined_prod_sell_f64": (None, [u32, u32, vp, vp, vp, vp, vp, vp, vp, vp, sz, sz]),
    "orc_pipelined_cg_vector_update_f64": (None, [sz, vp, f64, vp, vp, vp, f64, vp]),
    "orc_pipelined_bicgstab_update_s_f64": (None, [sz, vp, vp, vp, vp, sz, sz]),
    "orc_pipelined_bicgstab_vector_update_f64": (None, [sz, vp, f64, vp, f64, vp, vp, vp, f64, vp, vp, vp]),
    "orc_cg_pipelined_csr": (_Res, [P(_Csr), vp, vp, f64, f64, u32, vp]),
    "orc_cg_pipelined_sell": (_Res, [P(_Sell), vp, vp, f64, f64, u32, vp]),
    "orc_cg_classic_csr": (_Res, [P(_Csr), vp, vp, vp, f64, f64, u32, vp]),
    "orc_bicgstab_pipelined_csr": (_Res, [P(_Csr), vp, vp, f64, f64, u32, vp]),
    "orc_bicgstab_pipelined_sell": (_Res, [P(_Sell), vp, vp, f64, f64, u32, vp]),
    "orc_bicgstab_classic_csr": (_Res, [P(_Csr), vp, vp, vp, f64, f64, u32, u32, vp]),
    "orc_mixed_cg_csr": (_Res, [P(_Csr), vp, vp, f64, u32, f32, P(u32)]),
    "orc_mixed_cg2_csr": (_Res, [P(_Csr), vp, vp, vp, f64, u32, f32, i32, i32, P(u32), vp]),
    "orc_gmres_pipelined_csr": (_Res, [P(_Csr), vp, vp, f64, f64, u32, u32, i32, vp]),
    "orc_fmt_spmv_f64": (None, [P(_Fmt), vp, f64, vp, f64]),
    "orc_fmt_pipelined_prod_f64": (None, [P(_Fmt), vp, vp, vp, vp, sz, sz]),
    "orc_cg_pipelined_fmt": (_Res, [P(_Fmt), vp, vp, f64, f64, u32, vp]),
    "orc_bicgstab_pipelined_fmt": (_Res, [P(_Fmt), vp, vp, f64, f64, u32, vp]),
    "orc_gmres_normalize_vk": (None, [sz, vp, vp, vp, sz, vp, vp, sz, sz]),
    "orc_gmres_stage1": (None, [vp, sz, sz, sz, vp, sz, i32]),
    "orc_gmres_stage2": (None, [vp, sz, sz, sz, vp, vp, sz, vp, sz]),
    "orc_gmres_update_result": (None, [vp, vp, vp, sz, sz, vp, sz]),
    "orc_batched_cg_csr": (None, [u32, vp, P(_Csr), vp, vp, f64, u32, vp, vp]),
    "orc_batched_cg2_csr": (None, [u32, vp, P(_Csr), vp, vp, f64, u32, i32, vp, vp]),
    "orc_adjoint_grad_csr": (None, [u32, vp, vp, vp, vp, vp]),
}

_REF_SIGS = {
    "ref_set_num_threads": (None, [i32]),
    "ref_fdm_laplace": (u32, [u32, u32, vp, vp, vp]),
    "ref_csr_spmv_f64": (None, [u32, vp, vp, vp, vp, f64, vp, f64]),
    "ref_sell_spmv_f64": (None, [u32, u32, vp, vp, vp, vp, f64, vp, f64]),
    "ref_sell_arrays": (None, [u32, u32, vp, vp, vp, vp, vp, vp, vp, u64]),
    "ref_row_blocks": (u32, [u32, vp, vp, vp, vp]),
    "ref_row_info_diag": (None, [u32, vp, vp, vp, vp]),
    "ref_dot_f64": (f64, [u32, vp, vp]),
    "ref_norm2_f64": (f64, [u32, vp]),
    "ref_pipelined_cg_vector_update": (None, [u32, vp, f64, vp, vp, vp, f64, vp]),
    "ref_pipelined_cg_prod_csr": (None, [u32, vp, vp, vp, vp, vp, vp]),
    "ref_pipelined_bicgstab_prod_csr": (None, [u32, vp, vp, vp, vp, vp, vp, vp, u32]),
    "ref_pipelined_bicgstab_update_s": (None, [u32, vp, vp, vp, vp, u32]),
    "ref_pipelined_bicgstab_vector_update": (None, [u32, vp, f64, vp, f64, vp, vp, vp, f64, vp, vp, vp]),
    "ref_cg_csr": (None, [u32, vp, vp, vp, vp, vp, f64, f64, u32, i32, P(u32), P(f64)]),
    "ref_cg_sell": (None, [u32, u32, vp, vp, vp, vp, vp, f64, f64, u32, P(u32), P(f64)]),
    "ref_bicgstab_csr": (None, [u32, vp, vp, vp, vp, vp, f64, f64, u32, u32, i32, P(u32), P(f64)]),
    "ref_bicgstab_sell": (None, [u32, u32, vp, vp, vp, vp, vp, f64, f64, u32, P(u32), P(f64)]),
    "ref_mixed_cg_csr": (None, [u32, vp, vp, vp, vp, vp, f64, u32, f32, P(u32), P(f64)]),
    "ref_gmres_csr": (None, [u32, vp, vp, vp, vp, vp, f64, f64, u32, u32, P(u32), P(f64)]),
    "ref_coo_arrays": (None, [u32, vp, vp, vp, vp, vp]),
    "ref_ell_arrays": (None, [u32, vp, vp, vp, P(u32), P(u32), vp, vp]),
    "ref_hyb_arrays": (None, [u32, vp, vp, vp, P(u32), P(u32), P(u32), vp, vp, vp, vp, vp]),
    "ref_fmt_spmv_f64": (None, [i32, u32, vp, vp, vp, vp, f64, vp, f64]),
    "ref_fmt_pipelined_prod": (None, [i32, u32, vp, vp, vp, vp, vp, vp, vp, u32]),
    "ref_fmt_solve": (None, [i32, i32, u32, vp, vp, vp, vp, vp, f64, u32, u32, P(u32), P(f64)]),
    "ref_pipelined_gmres_normalize_vk": (None, [u32, vp, vp, vp, u32, u32, vp, vp, u32, u32, u32]),
    "ref_pipelined_gmres_stage1": (None, [vp, u32, u32, u32, u32, vp, u32]),
    "ref_pipelined_gmres_stage2": (None, [vp, u32, u32, u32, u32, vp, vp, vp, u32]),
    "ref_pipelined_gmres_update_result": (None, [vp, vp, vp, u32, u32, u32, vp, u32]),
    "ref_bench_csr_spmv": (f64, [u32, vp, vp, vp, vp, vp, u32]),
    "ref_bench_solve": (f64, [u32, vp, vp, vp, vp, vp, f64, u32, i32, i32, P(u32), P(f64)]),
    "ref_max_threads": (i32, []),
    "ref_mm_read": (C.c_long, [C.c_char_p, C.c_long, P(u32), P(u32)]),
    "ref_mm_csr": (None, [vp, vp, vp]),
}

_orc = None
_ref = None


def _load(path, sigs):
    L = C.CDLL(str(path))
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def orc():
    global _orc
    if _orc is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"oracle not built: {ORACLE_LIB}")
        _orc = _load(ORACLE_LIB, _ORC_SIGS)
    return _orc


def ref_available():
    return REF_LIB.exists()


def ref():
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError(f"reference build not available: {REF_LIB}")
        _ref = _load(REF_LIB, _REF_SIGS)
    return _ref


def p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def cu32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class CsrHost:
    """Host CSR arrays kept alive together with the oracle descriptor."""

    def __init__(self, row_ptr, col, val):
        self.row_ptr, self.col, self.val = cu32(row_ptr), cu32(col), c64(val)
        self.n = self.row_ptr.size - 1
        self.desc = _Csr(self.n, p(self.row_ptr), p(self.col), p(self.val))


class SellHost:
    def __init__(self, row_ptr, col, val, C_):
        row_ptr, col, val = cu32(row_ptr), cu32(col), c64(val)
        n = row_ptr.size - 1
        ns, stored = u32(), u64()
        orc().orc_sell_sizes(n, C_, p(row_ptr), C.byref(ns), C.byref(stored))
        self.cpb = np.zeros(ns.value, np.uint32)
        self.bs = np.zeros(ns.value, np.uint32)
        self.ci = np.zeros(max(stored.value, 1), np.uint32)
        self.el = np.zeros(max(stored.value, 1), np.float64)
        orc().orc_sell_from_csr_f64(n, C_, p(row_ptr), p(col), p(val), p(self.cpb), p(self.ci), p(self.bs), p(self.el))
        self.ci, self.el = self.ci[: stored.value], self.el[: stored.value]
        self.n, self.C = n, C_
        self.desc = _Sell(n, C_, p(self.cpb), p(self.ci), p(self.bs), p(self.el))


FORMAT_COO, FORMAT_ELL, FORMAT_HYB = 2, 3, 4


class FmtHost:
    """coordinate_matrix / ell_matrix / hyb_matrix arrays kept alive with the oracle descriptor.

    Built by the reference build's copy() (ref_*_arrays) when `arrays` is None, else from the given
    dict of arrays (a fixture): coords/elements (COO), ld/width/coords/elements (ELL), and for HYB
    also csr_rows/csr_cols/csr_elements."""

    def __init__(self, fmt, row_ptr, col, val, arrays=None):
        rp, cl, vl = cu32(row_ptr), cu32(col), c64(val)
        n = rp.size - 1
        self.format, self.n = fmt, n
        if arrays is None:
            arrays = ref_format_arrays(fmt, rp, cl, vl)
        self.arrays = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) and v.ndim else v)
                       for k, v in arrays.items()}
        a = self.arrays
        ld = int(a.get("ld", 0))
        width = int(a.get("width", 0))
        nnz = int(rp[-1]) if fmt == FORMAT_COO else 0
        self.desc = _Fmt(fmt, n, nnz, ld, width, p(a["coords"]), p(a["elements"]),
                         p(a["csr_rows"]) if fmt == FORMAT_HYB else None,
                         p(a["csr_cols"]) if fmt == FORMAT_HYB else None,
                         p(a["csr_elements"]) if fmt == FORMAT_HYB else None)


def ref_format_arrays(fmt, rp, col, val):
    """The arrays the reference's copy() writes for the format."""
    rp, col, val = cu32(rp), cu32(col), c64(val)
    n = rp.size - 1
    R = ref()
    if fmt == FORMAT_COO:
        nnz = int(rp[-1])
        coords = np.zeros(2 * max(nnz, 1), np.uint32)
        el = np.zeros(max(nnz, 1))
        R.ref_coo_arrays(n, p(rp), p(col), p(val), p(coords), p(el))
        return {"coords": coords[:2 * nnz], "elements": el[:nnz]}
    ld, w = u32(), u32()
    if fmt == FORMAT_ELL:
        R.ref_ell_arrays(n, p(rp), p(col), p(val), C.byref(ld), C.byref(w), None, None)
        coords = np.zeros(max(ld.value * w.value, 1), np.uint32)
        el = np.zeros(max(ld.value * w.value, 1))
        R.ref_ell_arrays(n, p(rp), p(col), p(val), C.byref(ld), C.byref(w), p(coords), p(el))
        return {"ld": ld.value, "width": w.value, "coords": coords, "elements": el}
    cn = u32()
    R.ref_hyb_arrays(n, p(rp), p(col), p(val), C.byref(ld), C.byref(w), C.byref(cn), None, None, None, None, None)
    m = max(ld.value * w.value, 1)
    ec, ee = np.zeros(m, np.uint32), np.zeros(m)
    cr, cc, ce = np.zeros(n + 1, np.uint32), np.zeros(cn.value, np.uint32), np.zeros(cn.value)
    R.ref_hyb_arrays(n, p(rp), p(col), p(val), C.byref(ld), C.byref(w), C.byref(cn), p(ec), p(ee), p(cr), p(cc),
                     p(ce))
    return {"ld": ld.value, "width": w.value, "coords": ec, "elements": ee, "csr_rows": cr, "csr_cols": cc,
            "csr_elements": ce}


def fmt_spmv(A, x, alpha=1.0, y=None, beta=0.0):
    x = c64(x)
    y = np.zeros(A.n) if y is None else c64(y).copy()
    orc().orc_fmt_spmv_f64(C.byref(A.desc), p(x), alpha, p(y), beta)
    return y


def fmt_pipelined_prod(A, pv, r0=None, chunk=256, offset=0):
    pv = c64(pv)
    Ap = np.zeros(A.n)
    buf = np.zeros(6 * chunk)
    orc().orc_fmt_pipelined_prod_f64(C.byref(A.desc), p(pv), p(Ap), p(c64(r0)) if r0 is not None else None, p(buf),
                                     chunk, offset)
    return Ap, buf


def ref_fmt_solve(fmt, kind, rp, col, val, b, tol, max_iters, krylov_dim=20):
    """solve(A, b, cg_tag / bicgstab_tag / gmres_tag) (kind 0 / 1 / 2) with A in the format."""
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_fmt_solve(fmt, kind, rp.size - 1, p(rp), p(col), p(val), p(b), p(x), tol, max_iters, krylov_dim,
                        C.byref(it), C.byref(err))
    return x, it.value, err.value


# ------------------------------------------------------------------ oracle wrappers

def set_threads(n):
    orc().orc_set_num_threads(n)
    if ref_available():
        ref().ref_set_num_threads(n)


class reduction_blocks:
    """Context manager: run the oracle's inner products as k blocks (k = 0: reference order)."""

    def __init__(self, k):
        self.k = k

    def __enter__(self):
        orc().orc_set_reduction_blocks(self.k)
        return self

    def __exit__(self, *exc):
        orc().orc_set_reduction_blocks(0)


ENVELOPE_BLOCKS = (0, 2, 3, 4, 6, 8, 16, 64, 256)


def iteration_envelope(run, blocks=ENVELOPE_BLOCKS):
    """[min, max] of the iteration count returned by run() (a tuple whose second entry is the count)
    over the reference order and the k-block reduction orders."""
    its = []
    for k in blocks:
        with reduction_blocks(k):
            its.append(int(run()[1]))
    return min(its), max(its)


def csr_spmv(rp, col, val, x, alpha=1.0, y=None, beta=0.0, xs=0, xi=1, ys=0, yi=1, dtype=np.float64):
    rp, col = cu32(rp), cu32(col)
    val = np.ascontiguousarray(val, dtype=dtype)
    x = np.ascontiguousarray(x, dtype=dtype)
    n = rp.size - 1
    y = np.zeros(ys + (n - 1) * yi + 1 if n else 0, dtype) if y is None else np.array(y, dtype=dtype)
    fn = orc().orc_csr_spmv_f64 if dtype == np.float64 else orc().orc_csr_spmv_f32
    fn(n, p(rp), p(col), p(val), p(x), xs, xi, alpha, p(y), ys, yi, beta)
    return y


def sell_spmv(S, x, alpha=1.0, y=None, beta=0.0, dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=dtype)
    y = np.zeros(S.n, dtype) if y is None else np.array(y, dtype=dtype)
    if dtype == np.float64:
        orc().orc_sell_spmv_f64(S.n, S.C, p(S.cpb), p(S.ci), p(S.bs), p(S.el), p(x), 0, 1, alpha, p(y), 0, 1, beta)
    else:
        el = S.el.astype(np.float32)
        orc().orc_sell_spmv_f32(S.n, S.C, p(S.cpb), p(S.ci), p(S.bs), p(el), p(x), 0, 1, alpha, p(y), 0, 1, beta)
    return y


def row_info(rp, col, val, info):
    rp, col, val = cu32(rp), cu32(col), c64(val)
    out = np.zeros(rp.size - 1)
    orc().orc_csr_row_info_f64(rp.size - 1, p(rp), p(col), p(val), p(out), info)
    return out


def dot(x, y):
    x, y = c64(x), c64(y)
    return orc().orc_dot_f64(x.size, p(x), p(y))


def norm2(x):
    x = c64(x)
    return orc().orc_norm2_f64(x.size, p(x))


def cg_pipelined(A, b, tol, max_iters, abs_tol=0.0, history=False):
    b = c64(b)
    x = np.zeros_like(b)
    h = np.zeros(max(max_iters, 1)) if history else None
    if isinstance(A, SellHost):
        r = orc().orc_cg_pipelined_sell(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, p(h))
    else:
        r = orc().orc_cg_pipelined_csr(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, p(h))
    return x, r.iters, r.error, (h[: r.iters] if history else None)


def cg_classic(A, b, tol, max_iters, diag=None, abs_tol=0.0, history=False):
    b = c64(b)
    x = np.zeros_like(b)
    d = c64(diag) if diag is not None else None
    h = np.zeros(max(max_iters, 1)) if history else None
    r = orc().orc_cg_classic_csr(C.byref(A.desc), p(b), p(x), p(d), tol, abs_tol, max_iters, p(h))
    return x, r.iters, r.error, (h[: r.iters] if history else None)


def bicgstab_pipelined(A, b, tol, max_iters, abs_tol=0.0, history=False):
    b = c64(b)
    x = np.zeros_like(b)
    h = np.zeros(max(max_iters, 1)) if history else None
    if isinstance(A, SellHost):
        r = orc().orc_bicgstab_pipelined_sell(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, p(h))
    else:
        r = orc().orc_bicgstab_pipelined_csr(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, p(h))
    return x, r.iters, r.error, (h[: r.iters] if history else None)


def bicgstab_classic(A, b, tol, max_iters, diag=None, max_before_restart=200, abs_tol=0.0, history=False):
    b = c64(b)
    x = np.zeros_like(b)
    d = c64(diag) if diag is not None else None
    h = np.zeros(max(max_iters, 1)) if history else None
    r = orc().orc_bicgstab_classic_csr(C.byref(A.desc), p(b), p(x), p(d), tol, abs_tol, max_iters,
                                       max_before_restart, p(h))
    return x, r.iters, r.error, (h[: r.iters] if history else None)


def mixed_cg(A, b, tol, max_iters, inner_tol=1e-2):
    b = c64(b)
    x = np.zeros_like(b)
    rest = u32()
    r = orc().orc_mixed_cg_csr(C.byref(A.desc), p(b), p(x), tol, max_iters, inner_tol, C.byref(rest))
    return x, r.iters, r.error, rest.value


def mixed_cg2(A, b, tol, max_iters, inner_tol=1e-2, mode=1, acc64=True, diag=None, history=False):
    """Mixed-precision CG of the HIP path (mode 0: reference schedule, 1: reliable updates).
    Returns x, inner iterations, final FP64 relative residual, FP64 corrections, history."""
    b = c64(b)
    x = np.zeros_like(b)
    d = c64(diag) if diag is not None else None
    h = np.zeros(max(max_iters, 1)) if history else None
    upd = u32()
    r = orc().orc_mixed_cg2_csr(C.byref(A.desc), p(b), p(x), p(d), tol, max_iters, inner_tol, int(mode),
                                int(bool(acc64)), C.byref(upd), p(h))
    return x, r.iters, r.error, upd.value, (h[: r.iters] if history else None)


def gmres(A, b, tol, max_iters, krylov_dim=20, gs_skip_tail=False, abs_tol=0.0, history=False):
    """Pipelined GMRES (gmres.hpp:182-370); gs_skip_tail reproduces the reference's Gram-Schmidt range."""
    b = c64(b)
    x = np.zeros_like(b)
    h = np.zeros(max(max_iters, 1)) if history else None
    r = orc().orc_gmres_pipelined_csr(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, krylov_dim,
                                      int(bool(gs_skip_tail)), p(h))
    return x, r.iters, r.error, (h[: r.iters] if history else None)


def gmres_normalize_vk(vk, residual, R, offset_in_R, buf, r_dot_vk, chunk, chunk_offset):
    """host_based pipelined_gmres_normalize_vk (:734-779); updates vk, R, r_dot_vk in place."""
    orc().orc_gmres_normalize_vk(vk.size, p(vk), p(residual), p(R), offset_in_R, p(buf), p(r_dot_vk), chunk,
                                 chunk_offset)


def gmres_stage1(basis, n, ld, k, vi_in_vk, chunk, skip_tail=False):
    """host_based pipelined_gmres_gram_schmidt_stage1 (:787-842); skip_tail: the single-thread range."""
    orc().orc_gmres_stage1(p(basis), n, ld, k, p(vi_in_vk), chunk, int(bool(skip_tail)))


def gmres_stage2(basis, n, ld, k, vi_in_vk, R, krylov_dim, buf, chunk):
    """host_based pipelined_gmres_gram_schmidt_stage2 (:851-893); updates basis, R, buf in place."""
    orc().orc_gmres_stage2(p(basis), n, ld, k, p(vi_in_vk), p(R), krylov_dim, p(buf), chunk)


def gmres_update_result(x, residual, basis, n, ld, coef, k):
    """host_based pipelined_gmres_update_result (:898-926); updates x in place."""
    orc().orc_gmres_update_result(p(x), p(residual), p(basis), n, ld, p(coef), k)


def cg_pipelined_fmt(A, b, tol, max_iters, abs_tol=0.0):
    b = c64(b)
    x = np.zeros_like(b)
    r = orc().orc_cg_pipelined_fmt(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, None)
    return x, r.iters, r.error


def bicgstab_pipelined_fmt(A, b, tol, max_iters, abs_tol=0.0):
    b = c64(b)
    x = np.zeros_like(b)
    r = orc().orc_bicgstab_pipelined_fmt(C.byref(A.desc), p(b), p(x), tol, abs_tol, max_iters, None)
    return x, r.iters, r.error


def batched_cg(sys_ptr, A, b, tol, max_iters, single=False):
    """Per-system pipelined CG of a block-diagonal CSR (single: FP32 vectors, FP64 accumulation)."""
    sys_ptr = cu32(sys_ptr)
    b = c64(b)
    x = np.zeros_like(b)
    ns = sys_ptr.size - 1
    iters = np.zeros(ns, np.uint32)
    errs = np.zeros(ns)
    orc().orc_batched_cg2_csr(ns, p(sys_ptr), C.byref(A.desc), p(b), p(x), tol, max_iters, int(bool(single)),
                              p(iters), p(errs))
    return x, iters, errs


def adjoint_grad(rp, col, lam, x):
    rp, col, lam, x = cu32(rp), cu32(col), c64(lam), c64(x)
    out = np.zeros(col.size)
    orc().orc_adjoint_grad_csr(rp.size - 1, p(rp), p(col), p(lam), p(x), p(out))
    return out


# ------------------------------------------------------------------ reference-build wrappers

def ref_cg(rp, col, val, b, tol, max_iters, jacobi=False, abs_tol=0.0):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_cg_csr(rp.size - 1, p(rp), p(col), p(val), p(b), p(x), tol, abs_tol, max_iters, int(jacobi),
                     C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_cg_sell(rp, col, val, C_, b, tol, max_iters, abs_tol=0.0):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_cg_sell(rp.size - 1, C_, p(rp), p(col), p(val), p(b), p(x), tol, abs_tol, max_iters,
                      C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_bicgstab(rp, col, val, b, tol, max_iters, jacobi=False, max_before_restart=200, abs_tol=0.0):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_bicgstab_csr(rp.size - 1, p(rp), p(col), p(val), p(b), p(x), tol, abs_tol, max_iters,
                           max_before_restart, int(jacobi), C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_bicgstab_sell(rp, col, val, C_, b, tol, max_iters, abs_tol=0.0):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_bicgstab_sell(rp.size - 1, C_, p(rp), p(col), p(val), p(b), p(x), tol, abs_tol, max_iters,
                            C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_gmres(rp, col, val, b, tol, max_iters, krylov_dim=20, abs_tol=0.0):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_gmres_csr(rp.size - 1, p(rp), p(col), p(val), p(b), p(x), tol, abs_tol, max_iters, krylov_dim,
                        C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_mixed_cg(rp, col, val, b, tol, max_iters, inner_tol=1e-2):
    rp, col, val, b = cu32(rp), cu32(col), c64(val), c64(b)
    x = np.zeros_like(b)
    it, err = u32(), f64()
    ref().ref_mixed_cg_csr(rp.size - 1, p(rp), p(col), p(val), p(b), p(x), tol, max_iters, inner_tol,
                           C.byref(it), C.byref(err))
    return x, it.value, err.value


def ref_read_matrix_market(path, index_base=1):
    """The reference reader (io/matrix_market.hpp:102-326, into std::vector<std::map>): returns
    (result, row_ptr, col, val); result is the reader's own return value (0 on a parse error, 1 when
    the file cannot be opened, else the number of lines read)."""
    nr, nz = u32(), u32()
    r = ref().ref_mm_read(str(path).encode(), index_base, C.byref(nr), C.byref(nz))
    rp = np.zeros(nr.value + 1, np.uint32)
    col = np.zeros(nz.value, np.uint32)
    val = np.zeros(nz.value, np.float64)
    ref().ref_mm_csr(p(rp), p(col), p(val))
    return int(r), rp, col, val
