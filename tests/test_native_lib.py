"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import glob
import os
import re
import subprocess

import pytest

from conftest import REPO


def _declared():
    names = []
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(fadesim_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_header_symbols():
    from fadesim_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libfadesim.so not built (python -m fadesim_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b[TW] (\w+)", out))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing
    lib = _native.lib()
    assert lib.fadesim_version() >= 100
    for n in _declared():
        assert hasattr(lib, n)


def test_default_omega_rule():
    from fadesim_amd import _native
    w = _native.lib().fadesim_default_omega(128, 128)
    assert 1.97 < w < 1.99
    assert _native.lib().fadesim_default_omega(64, 64) < w


def test_shape_validation_without_gpu():
    import torch
    from fadesim_amd.solver import solve_batch
    with pytest.raises(ValueError):
        solve_batch(torch.ones((2, 48, 48)), 1.0, torch.zeros((2, 48)))
    with pytest.raises(ValueError):
        solve_batch(torch.ones((2, 64, 64)), 1.0, torch.zeros((2, 64)))  # not a GPU tensor


# positional parameters of the extension's six entry points, in the order
# pybind11 binds them (split_arr_matmul.cu:204-220, 419-433, 565-581, 690-706,
# 953-969, 1197-1213; module table :1431-1436)
_HOOK_SIGNATURES = {
    "ir_drop_gs": ["input", "G", "batch_size", "row_size", "col_size", "numX", "numY", "iter_times", "g_wire",
                   "beta", "enable_break", "break_threshould", "nonlinear_cell_enable", "nonlinear_params", "ret_V"],
    "ir2s": ["input", "G", "batch_size", "row_size", "col_size", "numX", "numY", "iter_times", "r_wire",
             "enable_break", "break_threshould", "nonlinear_cell_enable", "nonlinear_params", "ret_V"],
    "ir2s_ou_local": ["input", "G", "batch_size", "row_size", "col_size", "numX", "numY", "iter_times", "r_wire",
                      "r_load_left", "r_load_down", "enable_break", "break_threshould", "nonlinear_cell_enable",
                      "nonlinear_params"],
    "ir_gs_ou_local": ["input", "G", "batch_size", "row_size", "col_size", "numX", "numY", "iter_times", "g_wire",
                       "g_load_left", "g_load_down", "beta", "enable_break", "break_threshould",
                       "nonlinear_cell_enable", "nonlinear_params"],
    "ir2s_ou": ["input", "G", "batch_size", "row_size", "col_size", "ou_row_size", "ou_col_size", "numX", "numY",
                "iter_times", "r_wire", "enable_break", "break_threshould", "nonlinear_cell_enable",
                "nonlinear_params"],
    "ir_gs_ou": ["input", "G", "batch_size", "row_size", "col_size", "ou_row_size", "ou_col_size", "numX", "numY",
                 "iter_times", "g_wire", "beta", "enable_break", "break_threshould", "nonlinear_cell_enable",
                 "nonlinear_params"],
}


def test_hook_module_has_the_extension_entry_points():
    import inspect
    from fadesim_amd import hook
    for name, params in _HOOK_SIGNATURES.items():
        for holder in (hook, hook._HookModule):
            f = getattr(holder, name)
            assert list(inspect.signature(f).parameters) == params, name


def test_hook_refuses_cpu_tensors():
    import torch
    from fadesim_amd import hook
    inp, G = torch.zeros((1, 1, 4)), torch.ones((1, 4, 4))
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        hook.ir_gs_ou(inp, G, 1, 4, 4, 2, 2, 1, 1, 1, 1.0, 1.0, False, 0.0, False, [])
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        hook.ir2s_ou_local(inp, G, 1, 4, 4, 1, 1, 1, 1.0, 1.0, 1.0, False, 0.0, False, [])


def test_c_abi_rejects_bad_descriptors_without_a_gpu():
    """Argument validation of the C entry points happens before any device
    work: each call returns FADESIM_E_INVALID / _UNSUPPORTED with a message."""
    import ctypes

    import torch
    from fadesim_amd import _native as nat
    if torch.cuda.is_available():
        # the descriptors below hold fake device addresses: a validation
        # regression must never reach a real device
        pytest.skip("fake-pointer validation runs only where no GPU is visible")
    L = nat.lib()
    E_INVALID = -1

    def err(rc):
        assert rc != 0
        return L.fadesim_last_error().decode()

    assert "null descriptor" in err(L.fadesim_solve_batch(None, None))
    assert "null descriptor" in err(L.fadesim_hook_ir_drop_gs(None, None))
    h = nat.HookDesc()
    h.row_size, h.col_size, h.batch_size, h.numX, h.numY = 16, 12, 1, 1, 1
    h.d_input = h.d_G = h.d_out = 8  # never dereferenced: validation fails first
    h.ou_rows, h.ou_cols = 5, 4
    assert "OU sizes" in err(L.fadesim_hook_ir_drop_gs(ctypes.byref(h), None))
    h.row_size = 1
    assert "row_size" in err(L.fadesim_hook_ir_drop_gs(ctypes.byref(h), None))
    q = nat.Ir2sDesc()
    q.dtype_bytes, q.layout, q.batch_size, q.row_size, q.col_size, q.numX, q.numY = 8, nat.IR2S_HOOK, 1, 8, 8, 1, 1
    q.r_wire = 0.0
    assert "r_wire" in err(L.fadesim_ir2s(ctypes.byref(q), None))
    kp = (ctypes.c_int * 7)(96, 320, 320, 608, 608, 320, 320)
    w = (ctypes.c_void_p * 7)(*([8] * 7))
    assert L.fadesim_nn_coarse_forward(-1, ctypes.c_void_p(8), w, kp, ctypes.c_void_p(8), None) == E_INVALID
    kp[3] = 600
    assert "Kp" in err(L.fadesim_nn_coarse_forward(1, ctypes.c_void_p(8), w, kp, ctypes.c_void_p(8), None))
    p = [ctypes.c_void_p(8)] * 6
    assert "128x128" in err(L.fadesim_nn_features_pooled(1, 64, 64, 2, *p, None))
    # warm start combined at load: needs the scale, an even channel count and
    # a level that divides the shape (the base field is optional: the solve
    # forms it when d_warm is NULL)
    d = nat.BatchDesc()
    d.batch, d.rows, d.cols = 4, 128, 128
    d.d_g = d.d_r_wire = d.d_vin = d.d_iout = 8
    d.tol = 1e-6
    d.d_warm_coarse, d.d_warm_scale, d.warm_coarse_ch, d.warm_level = 8, None, 16, 2
    assert "d_warm_scale" in err(L.fadesim_solve_batch(ctypes.byref(d), None))
    d.d_warm_scale = 8
    d.warm_coarse_ch = 3
    assert "d_warm_coarse" in err(L.fadesim_solve_batch(ctypes.byref(d), None))
    d.warm_coarse_ch, d.warm_level = 16, 7
    assert "d_warm_coarse" in err(L.fadesim_solve_batch(ctypes.byref(d), None))
    # the in-kernel network start: coarse16f weights, 128x128, no given
    # start or vout, the first test before any sweep
    n = nat.BatchDesc()
    n.batch, n.rows, n.cols, n.tol, n.flags = 4, 128, 128, 1e-6, nat.F_TEST_FIRST
    n.d_g = n.d_r_wire = n.d_vin = n.d_iout = 8
    kp16 = (ctypes.c_int * 7)(96, 160, 160, 320, 320, 160, 160)
    kp32 = (ctypes.c_int * 7)(96, 320, 320, 608, 608, 320, 320)
    assert "coarse16f" in err(L.fadesim_nn_solve_batch(ctypes.byref(n), w, kp32, None))
    n.rows = n.cols = 64
    assert "128 x 128" in err(L.fadesim_nn_solve_batch(ctypes.byref(n), w, kp16, None))
    n.rows = n.cols = 128
    n.d_vout = 8
    assert "d_vout" in err(L.fadesim_nn_solve_batch(ctypes.byref(n), w, kp16, None))
    n.d_vout = None
    n.flags = 0
    assert "TEST_FIRST" in err(L.fadesim_nn_solve_batch(ctypes.byref(n), w, kp16, None))


def test_c_abi_refuses_a_descriptor_of_another_size():
    """struct_size guards the descriptor layout: a caller built against an
    older or newer header is refused before any field past the size is read."""
    import ctypes

    import torch
    from fadesim_amd import _native as nat
    if torch.cuda.is_available():
        pytest.skip("fake-pointer validation runs only where no GPU is visible")
    L = nat.lib()
    d = nat.BatchDesc()
    assert d.struct_size == ctypes.sizeof(nat.BatchDesc)
    d.batch, d.rows, d.cols, d.tol = 4, 128, 128, 1e-6
    d.d_g = d.d_r_wire = d.d_vin = d.d_iout = 8
    d.struct_size = ctypes.sizeof(nat.BatchDesc) - 24  # the descriptor without its last four fields
    assert L.fadesim_solve_batch(ctypes.byref(d), None) == -1
    assert "struct_size" in L.fadesim_last_error().decode()
    q = nat.Gs64Desc()
    assert q.struct_size == ctypes.sizeof(nat.Gs64Desc)
    q.rows = q.cols = 4
    q.struct_size = ctypes.sizeof(nat.Gs64Desc) - 24
    assert L.fadesim_gs64_solve(ctypes.byref(q), None) == -1
    assert "struct_size" in L.fadesim_last_error().decode()
    q.struct_size = ctypes.sizeof(nat.Gs64Desc)
    q.order = 7
    assert L.fadesim_gs64_solve(ctypes.byref(q), None) == -1
    assert "order" in L.fadesim_last_error().decode()
