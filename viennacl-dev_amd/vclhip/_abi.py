"""ctypes declarations of the C ABI in include/vclhip.h.

The shared library is built in-tree (viennacl-dev_amd/lib/libvclhip.so).  Loading fails loudly
when it is missing: there is no CPU fallback for any compute entry point.
"""
import ctypes as C
import os
import pathlib

PKG_ROOT = pathlib.Path(__file__).resolve().parent.parent          # viennacl-dev_amd/
REPO_ROOT = PKG_ROOT.parent
LIB_PATH = PKG_ROOT / "lib" / "libvclhip.so"
HEADER_PATH = REPO_ROOT / "include" / "vclhip.h"

u32 = C.c_uint32
u64 = C.c_uint64
i32 = C.c_int
f32 = C.c_float
f64 = C.c_double
vp = C.c_void_p
sz = C.c_size_t
P = C.POINTER

SUCCESS = 0
ERROR_INVALID_VALUE = -1
ERROR_ZERO_DIAGONAL = -2
ERROR_NOT_SUPPORTED = -3
ERROR_NO_DEVICE = -4
ERROR_TOO_LARGE = -5

PRECOND_NONE = 0
PRECOND_JACOBI = 1

OPERATOR_AUTO = 0
OPERATOR_AS_GIVEN = 1
OPERATOR_DIA_UNCODED = 2
USED_AS_GIVEN = 0
USED_DIA = 1
USED_DIA_CODED = 2
INNER_AUTO, INNER_CSR, INNER_SELL64, INNER_SELL64_D16, INNER_DIA = 0, 1, 2, 3, 4
INNER_DIA_CODED = 5  # reported only


class Csr(C.Structure):
    _fields_ = [("n_rows", u32), ("n_cols", u32), ("nnz", u32), ("num_blocks", u32),
                ("row_ptr", vp), ("col_idx", vp), ("values", vp), ("blocks", vp)]


class Sell(C.Structure):
    _fields_ = [("n_rows", u32), ("n_cols", u32), ("C", u32), ("num_slices", u32),
                ("columns_per_block", vp), ("column_indices", vp), ("block_start", vp),
                ("elements", vp)]


class Coo(C.Structure):
    _fields_ = [("n_rows", u32), ("n_cols", u32), ("nnz", u32), ("num_blocks", u32),
                ("coords", vp), ("elements", vp), ("row_ptr", vp), ("blocks", vp)]


class Ell(C.Structure):
    _fields_ = [("n_rows", u32), ("n_cols", u32), ("internal_size1", u32), ("internal_maxnnz", u32),
                ("coords", vp), ("elements", vp)]


class Hyb(C.Structure):
    _fields_ = [("n_rows", u32), ("n_cols", u32), ("internal_size1", u32), ("internal_ellnnz", u32),
                ("ell_coords", vp), ("ell_elements", vp), ("csr_rows", vp), ("csr_cols", vp),
                ("csr_elements", vp), ("num_blocks", u32), ("blocks", vp)]


FORMAT_CSR, FORMAT_SELL, FORMAT_COO, FORMAT_ELL, FORMAT_HYB = 0, 1, 2, 3, 4


class Matrix(C.Structure):
    _fields_ = [("format", i32), ("desc", vp)]


MONITOR = C.CFUNCTYPE(C.c_int, vp, f64, vp)


class SolverParams(C.Structure):
    _fields_ = [("tol", f64), ("abs_tol", f64), ("max_iters", u32),
                ("max_iters_before_restart", u32), ("inner_tol", f32), ("chunk", u32),
                ("use_graph", i32), ("mixed_mode", i32), ("variant", i32), ("history_host", P(f64)),
                ("history_len", u32), ("monitor", MONITOR), ("monitor_user", vp), ("inner_format", i32),
                ("krylov_dim", u32), ("operator_format", i32), ("dia", vp)]


class SolverResult(C.Structure):
    _fields_ = [("iters", u32), ("error", f64), ("restarts", u32), ("converged", i32), ("operator_used", i32),
                ("kernels_per_iter", u32)]


# name -> (restype, argtypes); every function of include/vclhip.h
_SIGS = {
    "vclhip_version": (C.c_char_p, []),
    "vclhip_last_error": (C.c_char_p, []),
    "vclhip_status_string": (C.c_char_p, [i32]),
    "vclhip_device_count": (i32, [P(i32)]),
    "vclhip_set_device": (i32, [i32]),
    "vclhip_get_device": (i32, [P(i32)]),
    "vclhip_device_info": (i32, [i32, C.c_char_p, sz, P(sz), P(i32), P(i32), P(i32)]),
    "vclhip_malloc": (i32, [P(vp), sz]),
    "vclhip_free": (i32, [vp]),
    "vclhip_memcpy_h2d": (i32, [vp, vp, sz, vp]),
    "vclhip_memcpy_d2h": (i32, [vp, vp, sz, vp]),
    "vclhip_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "vclhip_fill_zero": (i32, [vp, sz, vp]),
    "vclhip_stream_read_f64": (i32, [u64, vp, vp, u32, vp]),
    "vclhip_gather_read_f64": (i32, [u32, vp, vp, u32, vp, u32, vp]),
    "vclhip_empty_launch": (i32, [u32, vp]),
    "vclhip_grid_sync_probe": (i32, [u32, u32, vp, vp]),
    "vclhip_stream_create": (i32, [P(vp)]),
    "vclhip_stream_destroy": (i32, [vp]),
    "vclhip_stream_synchronize": (i32, [vp]),
    "vclhip_device_synchronize": (i32, []),
    "vclhip_time_next_launch": (i32, [vp, vp]),
    "vclhip_event_create": (i32, [P(vp)]),
    "vclhip_event_destroy": (i32, [vp]),
    "vclhip_event_record": (i32, [vp, vp]),
    "vclhip_event_elapsed_ms": (i32, [vp, vp, P(f32)]),
    "vclhip_csr_plan_blocks": (i32, [vp, u32, vp, P(u32)]),
    "vclhip_csr_reference_row_blocks": (i32, [vp, u32, vp, P(u32)]),
    "vclhip_sell_from_csr_f64": (i32, [u32, u32, vp, vp, vp, P(u32), P(u64), vp, vp, vp, vp]),
    "vclhip_csr_spmv_f64": (i32, [P(Csr), vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_csr_spmv_f32": (i32, [P(Csr), vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_sell_spmv_f64": (i32, [P(Sell), vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_sell_spmv_f32": (i32, [P(Sell), vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_csr_row_info_f64": (i32, [P(Csr), vp, i32, vp]),
    "vclhip_csr_row_info_f32": (i32, [P(Csr), vp, i32, vp]),
    "vclhip_csr_adjoint_grad_f64": (i32, [P(Csr), vp, vp, vp, vp]),
    "vclhip_csr_adjoint_grad_f32": (i32, [P(Csr), vp, vp, vp, vp]),
    "vclhip_dot_f64": (i32, [u32, vp, u32, u32, vp, u32, u32, P(f64), vp]),
    "vclhip_dot_f32": (i32, [u32, vp, u32, u32, vp, u32, u32, P(f32), vp]),
    "vclhip_dot_f32_acc64": (i32, [u32, vp, vp, P(f64), vp]),
    "vclhip_norm2_f64": (i32, [u32, vp, u32, u32, P(f64), vp]),
    "vclhip_norm2_f32": (i32, [u32, vp, u32, u32, P(f32), vp]),
    "vclhip_axpby_f64": (i32, [u32, f64, vp, f64, vp, vp, vp]),
    "vclhip_axpby_f32": (i32, [u32, f32, vp, f32, vp, vp, vp]),
    "vclhip_avbv_f64": (i32, [u32, vp, u32, u32, vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_avbv_f32": (i32, [u32, vp, u32, u32, vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_fill_f64": (i32, [u32, vp, u32, u32, f64, vp]),
    "vclhip_fill_f32": (i32, [u32, vp, u32, u32, f32, vp]),
    "vclhip_element_div_f64": (i32, [u32, vp, vp, vp, vp]),
    "vclhip_element_div_f32": (i32, [u32, vp, vp, vp, vp]),
    "vclhip_convert_f64_to_f32": (i32, [u32, vp, vp, vp]),
    "vclhip_convert_f32_to_f64": (i32, [u32, vp, vp, vp]),
    "vclhip_asum_f64": (i32, [u32, vp, u32, u32, P(f64), vp]),
    "vclhip_asum_f32": (i32, [u32, vp, u32, u32, P(f32), vp]),
    "vclhip_iamax_f64": (i32, [u32, vp, u32, u32, P(u32), vp]),
    "vclhip_iamax_f32": (i32, [u32, vp, u32, u32, P(u32), vp]),
    "vclhip_scal_f64": (i32, [u32, vp, u32, u32, f64, vp]),
    "vclhip_scal_f32": (i32, [u32, vp, u32, u32, f32, vp]),
    "vclhip_copy_f64": (i32, [u32, vp, u32, u32, vp, u32, u32, vp]),
    "vclhip_copy_f32": (i32, [u32, vp, u32, u32, vp, u32, u32, vp]),
    "vclhip_swap_f64": (i32, [u32, vp, u32, u32, vp, u32, u32, vp]),
    "vclhip_swap_f32": (i32, [u32, vp, u32, u32, vp, u32, u32, vp]),
    "vclhip_rot_f64": (i32, [u32, vp, u32, u32, vp, u32, u32, f64, f64, vp]),
    "vclhip_rot_f32": (i32, [u32, vp, u32, u32, vp, u32, u32, f32, f32, vp]),
    "vclhip_pipelined_cg_vector_update_f64": (i32, [u32, vp, f64, vp, vp, vp, f64, vp, u32, vp]),
    "vclhip_pipelined_cg_prod_csr_f64": (i32, [P(Csr), vp, vp, vp, u32, vp]),
    "vclhip_pipelined_cg_prod_sell_f64": (i32, [P(Sell), vp, vp, vp, u32, vp]),
    "vclhip_pipelined_bicgstab_update_s_f64": (i32, [u32, vp, vp, vp, vp, u32, u32, vp]),
    "vclhip_pipelined_bicgstab_vector_update_f64":
        (i32, [u32, vp, f64, vp, f64, vp, vp, vp, f64, vp, vp, vp, u32, vp]),
    "vclhip_pipelined_bicgstab_prod_csr_f64": (i32, [P(Csr), vp, vp, vp, vp, u32, u32, vp]),
    "vclhip_pipelined_bicgstab_prod_sell_f64": (i32, [P(Sell), vp, vp, vp, vp, u32, u32, vp]),
    "vclhip_pipelined_gmres_prod_csr_f64": (i32, [P(Csr), vp, vp, vp, u32, vp]),
    "vclhip_pipelined_gmres_prod_sell_f64": (i32, [P(Sell), vp, vp, vp, u32, vp]),
    "vclhip_pipelined_gmres_normalize_vk_f64": (i32, [u32, vp, vp, vp, C.c_size_t, vp, vp, u32, C.c_size_t, vp]),
    "vclhip_pipelined_gmres_gram_schmidt_stage1_f64": (i32, [vp, u32, C.c_size_t, u32, vp, u32, vp]),
    "vclhip_pipelined_gmres_gram_schmidt_stage2_f64": (i32, [vp, u32, C.c_size_t, u32, vp, vp, u32, vp, u32, vp]),
    "vclhip_pipelined_gmres_update_result_f64": (i32, [vp, vp, vp, u32, C.c_size_t, vp, u32, vp]),
    "vclhip_coo_plan": (i32, [vp, u32, u32, vp, vp, P(u32)]),
    "vclhip_ell_from_csr_f64": (i32, [u32, u32, vp, vp, vp, P(u32), vp, vp]),
    "vclhip_hyb_from_csr_f64": (i32, [u32, f64, u32, vp, vp, vp, P(u32), P(u32), vp, vp, vp, vp, vp]),
    "vclhip_coo_spmv_f64": (i32, [P(Coo), vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_coo_spmv_f32": (i32, [P(Coo), vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_ell_spmv_f64": (i32, [P(Ell), vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_ell_spmv_f32": (i32, [P(Ell), vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_hyb_spmv_f64": (i32, [P(Hyb), vp, u32, u32, f64, vp, u32, u32, f64, vp]),
    "vclhip_hyb_spmv_f32": (i32, [P(Hyb), vp, u32, u32, f32, vp, u32, u32, f32, vp]),
    "vclhip_matrix_diag_f64": (i32, [P(Matrix), vp, vp]),
    "vclhip_matrix_diag_f32": (i32, [P(Matrix), vp, vp]),
    "vclhip_pipelined_cg_prod_f64": (i32, [P(Matrix), vp, vp, vp, u32, vp]),
    "vclhip_pipelined_bicgstab_prod_f64": (i32, [P(Matrix), vp, vp, vp, vp, u32, u32, vp]),
    "vclhip_cg_solve_f64": (i32, [P(Matrix), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_cg_solve_f32": (i32, [P(Matrix), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_f64": (i32, [P(Matrix), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_f32": (i32, [P(Matrix), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_f64": (i32, [P(Matrix), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_f32": (i32, [P(Matrix), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_solver_params_default": (None, [P(SolverParams)]),
    "vclhip_dia_create_f64": (i32, [P(Csr), P(vp), vp]),
    "vclhip_dia_create_f32": (i32, [P(Csr), P(vp), vp]),
    "vclhip_dia_info": (i32, [vp, P(i32), P(u32)]),
    "vclhip_dia_destroy": (i32, [vp]),
    "vclhip_cg_solve_csr_f64": (i32, [P(Csr), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_cg_solve_csr_f64_noalloc": (i32, [P(Csr), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_cg_solve_sell_f64": (i32, [P(Sell), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_csr_f64": (i32, [P(Csr), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_sell_f64": (i32, [P(Sell), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_cg_solve_csr_f32": (i32, [P(Csr), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_cg_solve_sell_f32": (i32, [P(Sell), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_csr_f32": (i32, [P(Csr), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_bicgstab_solve_sell_f32": (i32, [P(Sell), vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_mixed_cg_solve_csr": (i32, [P(Csr), vp, vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_csr_f64": (i32, [P(Csr), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_sell_f64": (i32, [P(Sell), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_csr_f32": (i32, [P(Csr), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_gmres_solve_sell_f32": (i32, [P(Sell), vp, vp, P(SolverParams), P(SolverResult), vp]),
    "vclhip_mixed_cg_solve_sell": (i32, [P(Sell), vp, vp, vp, i32, P(SolverParams), P(SolverResult), vp]),
    "vclhip_batch_plan_create": (i32, [u32, vp, vp, P(vp)]),
    "vclhip_batch_plan_destroy": (i32, [vp]),
    "vclhip_batch_plan_autotune": (i32, [vp, P(Csr), vp, i32, vp]),
    "vclhip_batch_plan_lanes": (i32, [vp, i32, vp]),
    "vclhip_batched_cg_solve_f64": (i32, [vp, P(Csr), vp, vp, vp, P(SolverParams), vp, vp, vp, vp]),
    "vclhip_batched_cg_solve_f32": (i32, [vp, P(Csr), vp, vp, vp, P(SolverParams), vp, vp, vp, vp]),
    "vclhip_mm_read": (i32, [C.c_char_p, C.c_int64, P(vp), P(u32), P(u32), P(u32), P(i32), P(C.c_int64)]),
    "vclhip_mm_csr": (i32, [vp, vp, vp, vp]),
    "vclhip_mm_free": (i32, [vp]),
    "vclhip_mm_write": (i32, [C.c_char_p, C.c_int64, u32, u32, vp, vp, vp]),
    "vclhip_random_uniform": (i32, [u64, u64, f64, f64, vp]),
    "vclhip_random_normal": (i32, [u64, u64, f64, f64, vp]),
    "vclhip_gen_fdm_laplace_2d": (i32, [u32, u32, vp, vp, vp, P(u32)]),
    "vclhip_gen_laplace_3d": (i32, [u32, u32, u32, vp, vp, vp, P(u32)]),
    "vclhip_gen_diffusion": (i32, [u32, u32, u32, vp, vp, vp, vp, P(u32)]),
    "vclhip_gen_powerlaw": (i32, [u32, u32, f64, f64, u64, vp, vp, vp, P(u32)]),
    "vclhip_gen_graph_laplacian_batch":
        (i32, [u32, u32, u32, f64, f64, f64, f64, u64, vp, vp, vp, vp, P(u32), P(u32)]),
}

_lib = None


class VclHipError(RuntimeError):
    def __init__(self, fn, status, text):
        super().__init__(f"{fn} failed with status {status}: {text}")
        self.status = status


def lib():
    """Loads libvclhip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libvclhip is not built: {LIB_PATH} missing "
                               f"(run __graft_entry__.build() or make -C viennacl-dev_amd)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Calls an int-returning entry point and raises VclHipError on a nonzero status."""
    L = lib()
    st = getattr(L, name)(*args)
    if st != 0:
        raise VclHipError(name, st, L.vclhip_last_error().decode(errors="replace"))
    return st


def declared_functions():
    """Names of every function declared in include/vclhip.h (parsed from the header text)."""
    import re
    text = HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(vclhip_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_t")))
