"""ctypes binding of the C ABI in include/fadesim.h (libfadesim.so).

The library is built in-tree by ``fadesim_amd.build``; nothing here falls
back to another implementation: if the library is missing or a call fails,
a RuntimeError says so.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libfadesim.so")

FADESIM_OK = 0
ST_CONVERGED, ST_MAXIT, ST_NONFINITE, ST_ABORTED = 0, 1, 2, 3
F_FIXED_SWEEPS, F_NO_REFINE, F_DWR, F_TIMING, F_ZERO_CORRECTION, F_TEST_FIRST, F_NO_FASTPATH = 1, 2, 4, 8, 16, 32, 64
MAX_IR2S_ITERS = 64

_vp = ctypes.c_void_p


class _SizedDesc(ctypes.Structure):
    """A descriptor whose leading ``struct_size`` is filled in at creation."""

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.struct_size = ctypes.sizeof(self)


class BatchDesc(_SizedDesc):
    """Mirror of ``fadesim_batch_desc``."""
    _fields_ = [
        ("struct_size", ctypes.c_size_t),
        ("batch", ctypes.c_int), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
        ("d_g", _vp), ("d_gidx", _vp), ("d_r_wire", _vp), ("d_vin", _vp), ("d_vout", _vp), ("d_warm", _vp),
        ("omega", ctypes.c_double), ("tol", ctypes.c_double), ("safety", ctypes.c_double),
        ("switch_tol", ctypes.c_double), ("max_sweeps", ctypes.c_int),
        ("flags", ctypes.c_uint),
        ("d_iout", _vp), ("d_sweeps", _vp), ("d_status", _vp), ("d_est", _vp), ("d_field", _vp),
        ("ir2s_iters", ctypes.c_int), ("corr_iters", ctypes.c_int), ("ir2s_omega", ctypes.c_double),
        ("d_warm_coarse", _vp), ("d_warm_scale", _vp), ("warm_coarse_ch", ctypes.c_int), ("warm_level", ctypes.c_int),
        ("d_guess_rejected", _vp),
    ]


class Gs64Desc(_SizedDesc):
    """Mirror of ``fadesim_gs64_desc``."""
    _fields_ = [
        ("struct_size", ctypes.c_size_t),
        ("rows", ctypes.c_int), ("cols", ctypes.c_int),
        ("d_g", _vp), ("d_ru", _vp), ("ru_ld", ctypes.c_int), ("d_rd", _vp), ("rd_ld", ctypes.c_int),
        ("d_vin", _vp), ("d_vout", _vp),
        ("iter_number", ctypes.c_int), ("omega", ctypes.c_double),
        ("enable_break", ctypes.c_int), ("tol", ctypes.c_double),
        ("d_vup", _vp), ("d_vdown", _vp), ("d_iout", _vp), ("d_record", _vp), ("d_sweeps", _vp),
        ("general", ctypes.c_int), ("d_on_top", _vp), ("d_on_left", _vp), ("d_on_bottom", _vp),
        ("d_on_right", _vp), ("d_v_top", _vp), ("d_v_right", _vp),
        ("order", ctypes.c_int), ("d_init_up", _vp), ("d_init_down", _vp),
    ]


GS64_LEXICOGRAPHIC, GS64_REDBLACK = 0, 1


class HookDesc(ctypes.Structure):
    """Mirror of ``fadesim_hook_desc``."""
    _fields_ = [
        ("dtype_bytes", ctypes.c_int),
        ("d_input", _vp), ("d_G", _vp),
        ("batch_size", ctypes.c_int), ("row_size", ctypes.c_int), ("col_size", ctypes.c_int),
        ("numX", ctypes.c_int), ("numY", ctypes.c_int), ("iter_times", ctypes.c_int),
        ("g_wire", ctypes.c_double), ("beta", ctypes.c_double),
        ("enable_break", ctypes.c_int), ("break_threshold", ctypes.c_double),
        ("nonlinear", ctypes.c_int), ("nonlinear_params", ctypes.c_double * 6),
        ("ret_v", ctypes.c_int), ("d_out", _vp), ("d_sweeps", _vp),
        ("g_load_left", ctypes.c_double), ("g_load_down", ctypes.c_double),
        ("ou_rows", ctypes.c_int), ("ou_cols", ctypes.c_int),
    ]


class Ir2sDesc(ctypes.Structure):
    """Mirror of ``fadesim_ir2s_desc``."""
    _fields_ = [
        ("dtype_bytes", ctypes.c_int), ("layout", ctypes.c_int),
        ("d_vin", _vp), ("d_G", _vp), ("d_vout", _vp),
        ("batch_size", ctypes.c_int), ("row_size", ctypes.c_int), ("col_size", ctypes.c_int),
        ("numX", ctypes.c_int), ("numY", ctypes.c_int), ("iter_times", ctypes.c_int),
        ("r_wire", ctypes.c_double), ("r_source", ctypes.c_double), ("r_sink", ctypes.c_double),
        ("divide_by_g", ctypes.c_int), ("carry_residue", ctypes.c_int),
        ("enable_break", ctypes.c_int), ("break_threshold", ctypes.c_double),
        ("nonlinear", ctypes.c_int), ("nonlinear_params", ctypes.c_double * 6),
        ("ret_v", ctypes.c_int), ("d_out", _vp), ("d_vup", _vp), ("d_vdown", _vp), ("d_iters", _vp),
        ("ou_rows", ctypes.c_int), ("ou_cols", ctypes.c_int),
    ]


IR2S_HOOK, IR2S_PLAIN = 0, 1
IR2S_BREAK_BATCH, IR2S_BREAK_EACH = 1, 2

# Every symbol include/fadesim.h declares.
EXPORTED = [
    "fadesim_last_error", "fadesim_version", "fadesim_solve_batch", "fadesim_solve_batch_noalloc",
    "fadesim_default_omega",
    "fadesim_last_kernel_ms", "fadesim_last_guess_ms", "fadesim_cell_conductance",
    "fadesim_ragged_plan_create", "fadesim_ragged_plan_destroy", "fadesim_ragged_plan_info",
    "fadesim_ragged_solve", "fadesim_ragged_aborted", "fadesim_ragged_last_ms",
    "fadesim_gs64_solve", "fadesim_hook_ir_drop_gs", "fadesim_ir2s",
    "fadesim_nn_features", "fadesim_nn_features_pooled", "fadesim_nn_coarse_forward", "fadesim_nn_combine", "fadesim_nn_combine_up", "fadesim_nn_conv3x3_fwd", "fadesim_nn_flip_weights",
    "fadesim_nn_conv3x3_wgrad", "fadesim_nn_pool2", "fadesim_nn_pool2_bwd", "fadesim_nn_up2_add",
    "fadesim_nn_up2_bwd", "fadesim_nn_relu_mask", "fadesim_nn_add", "fadesim_nn_mse", "fadesim_nn_adam",
    "fadesim_nn_cast", "fadesim_nn_solve_batch",
]

_LIB = None


def lib():
    """Load libfadesim.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m fadesim_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        L.fadesim_last_error.restype = ctypes.c_char_p
        L.fadesim_version.restype = ctypes.c_int
        L.fadesim_solve_batch.argtypes = [ctypes.POINTER(BatchDesc), _vp]
        L.fadesim_solve_batch.restype = ctypes.c_int
        if hasattr(L, "fadesim_solve_batch_noalloc"):
            L.fadesim_solve_batch_noalloc.argtypes = [ctypes.POINTER(BatchDesc), _vp]
            L.fadesim_solve_batch_noalloc.restype = ctypes.c_int
        if hasattr(L, "fadesim_nn_solve_batch"):
            L.fadesim_nn_solve_batch.argtypes = [ctypes.POINTER(BatchDesc), _vp, _vp, _vp]
            L.fadesim_nn_solve_batch.restype = ctypes.c_int
        L.fadesim_last_kernel_ms.restype = ctypes.c_double
        if hasattr(L, "fadesim_last_guess_ms"):
            L.fadesim_last_guess_ms.restype = ctypes.c_double
        L.fadesim_default_omega.argtypes = [ctypes.c_int, ctypes.c_int]
        L.fadesim_default_omega.restype = ctypes.c_double
        if hasattr(L, "fadesim_gs64_solve"):
            L.fadesim_gs64_solve.argtypes = [ctypes.POINTER(Gs64Desc), _vp]
            L.fadesim_gs64_solve.restype = ctypes.c_int
        if hasattr(L, "fadesim_hook_ir_drop_gs"):
            L.fadesim_hook_ir_drop_gs.argtypes = [ctypes.POINTER(HookDesc), _vp]
            L.fadesim_hook_ir_drop_gs.restype = ctypes.c_int
        if hasattr(L, "fadesim_ir2s"):
            L.fadesim_ir2s.argtypes = [ctypes.POINTER(Ir2sDesc), _vp]
            L.fadesim_ir2s.restype = ctypes.c_int
        _declare_nn(L)
        _declare_ragged(L)
        _LIB = L
    return _LIB


def _declare_nn(L):
    i, f, sz = ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    sigs = {
        "fadesim_nn_features": [i, i, i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "fadesim_nn_features_pooled": [i, i, i, i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "fadesim_nn_coarse_forward": [i, _vp, _vp, _vp, _vp, _vp],
        "fadesim_nn_combine": [i, i, i, _vp, _vp, _vp, i, _vp, _vp],
        "fadesim_nn_combine_up": [i, i, i, _vp, _vp, _vp, i, i, _vp, _vp],
        "fadesim_nn_conv3x3_fwd": [i, i, i, i, i, i, _vp, _vp, _vp, i, _vp],
        "fadesim_nn_flip_weights": [i, i, i, i, _vp, _vp, _vp],
        "fadesim_nn_conv3x3_wgrad": [i, i, i, i, i, i, _vp, _vp, _vp, _vp],
        "fadesim_nn_pool2": [i, i, i, i, _vp, _vp, _vp],
        "fadesim_nn_pool2_bwd": [i, i, i, i, _vp, i, _vp, _vp],
        "fadesim_nn_up2_add": [i, i, i, i, _vp, _vp, _vp, _vp],
        "fadesim_nn_up2_bwd": [i, i, i, i, _vp, _vp, _vp],
        "fadesim_nn_relu_mask": [sz, _vp, _vp, _vp, _vp],
        "fadesim_nn_add": [sz, _vp, _vp, _vp, _vp],
        "fadesim_nn_mse": [i, i, i, _vp, i, _vp, _vp, _vp, _vp],
        "fadesim_nn_adam": [sz, _vp, _vp, _vp, _vp, _vp, f, f, f, f, i, f, f, _vp],
        "fadesim_nn_cast": [sz, _vp, i, _vp, _vp],
        "fadesim_cell_conductance": [i, i, i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    }
    for name, args in sigs.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int


class RaggedClassOpts(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double), ("check_first", ctypes.c_int), ("check_min", ctypes.c_int),
                ("check_max", ctypes.c_int), ("ir2s_iters", ctypes.c_int), ("corr_iters", ctypes.c_int),
                ("ir2s_omega", ctypes.c_double)]


class RaggedPlanStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int), ("cells", ctypes.c_longlong), ("rows_total", ctypes.c_longlong),
                ("cols_total", ctypes.c_longlong), ("grid", ctypes.c_int), ("teams", ctypes.c_int),
                ("big_tasks", ctypes.c_int), ("small_items", ctypes.c_int), ("omega", ctypes.c_double * 4),
                ("check_first", ctypes.c_int * 4), ("check_min", ctypes.c_int * 4), ("check_max", ctypes.c_int * 4),
                ("ir2s_iters", ctypes.c_int * 4), ("corr_iters", ctypes.c_int * 4), ("ir2s_omega", ctypes.c_double * 4)]


class RaggedDesc(ctypes.Structure):
    _fields_ = [("d_g", _vp), ("d_r_wire", _vp), ("d_vin", _vp), ("d_vout", _vp), ("d_warm", _vp),
                ("tol", ctypes.c_double), ("safety", ctypes.c_double), ("switch_tol", ctypes.c_double),
                ("max_sweeps", ctypes.c_int), ("flags", ctypes.c_uint), ("d_iout", _vp), ("d_sweeps", _vp),
                ("d_status", _vp), ("d_est", _vp), ("d_field", _vp)]


def _declare_ragged(L):
    if not hasattr(L, "fadesim_ragged_plan_create"):
        return
    L.fadesim_ragged_plan_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, _vp, _vp, _vp]
    L.fadesim_ragged_plan_create.restype = ctypes.c_int
    L.fadesim_ragged_plan_destroy.argtypes = [_vp]
    L.fadesim_ragged_plan_destroy.restype = None
    L.fadesim_ragged_plan_info.argtypes = [_vp, ctypes.POINTER(RaggedPlanStats)]
    L.fadesim_ragged_plan_info.restype = ctypes.c_int
    L.fadesim_ragged_solve.argtypes = [_vp, ctypes.POINTER(RaggedDesc), _vp]
    L.fadesim_ragged_solve.restype = ctypes.c_int
    L.fadesim_ragged_aborted.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.fadesim_ragged_aborted.restype = ctypes.c_int
    L.fadesim_ragged_last_ms.argtypes = [_vp]
    L.fadesim_ragged_last_ms.restype = ctypes.c_double


def check(rc: int, what: str) -> None:
    if rc != FADESIM_OK:
        msg = lib().fadesim_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def stream_handle(device=None) -> int:
    """hipStream_t of torch's current stream on `device`."""
    import torch
    return torch.cuda.current_stream(device).cuda_stream
