"""Drop-in for the PimTorch array-simulation hook (``ir_drop_gs`` and ``ir2s``).

PimTorch builds ``phyArrParams.cpp_ir_drop_accelerate`` at import time
(pimtorch/pimfixedpoint/fixedPoint/nn/commonConst.py:273-280) and calls
``ir_drop_gs`` positionally from splitArrayArithmetic.py:488-511.  This module
exposes the same function (same positional signature, argument meaning,
return shapes and error messages as
pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/split_arr_matmul.cu:204-333),
backed by the HIP library:

* ``mode = "reference"`` (default): the reference's own line-GS iteration,
  fixed ``iter_times`` sweeps or the batch-wide break -- the same iterates,
  for callers that depend on its (often unconverged) results;
* ``mode = "converged"``: every crossbar solved to ``tol`` relative
  bitline-current error by the batched red-black solver
  (``fadesim_solve_batch``); ``iter_times`` and the break settings are then
  ignored.  Linear cells and shapes the batched kernel supports only.

The extension's other entry points are here too: ``ir2s`` (the wire-solve
iteration), ``ir_gs_ou_local`` / ``ir2s_ou_local`` (separate source and sink
loads) and the OU-mode ``ir_gs_ou`` / ``ir2s_ou`` (every OU block of a tile
solved as its own crossbar).

``install(phyArrParams)`` assigns this module where PimTorch looks for it.
"""
from __future__ import annotations

import ctypes
import sys
from typing import Sequence

import torch

from . import _native as nat

mode = "reference"
tol = 1e-6


def _check_input(x: torch.Tensor, name: str) -> None:
    if not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def ir_drop_gs(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int,
               col_size: int, numX: int, numY: int, iter_times: int, g_wire: float,
               beta: float, enable_break: bool, break_threshould: float,
               nonlinear_cell_enable: bool, nonlinear_params: Sequence[float],
               ret_V: bool = False) -> torch.Tensor:
    """IR-drop solve of ``batch_size x numX x numY`` crossbars of
    ``row_size x col_size`` cells.

    input [batch, numX, row] wordline voltages; G [numX, row, numY*col]
    conductances shared by the batch.  Returns the bitline currents
    [batch, numX, numY*col], or with ``ret_V`` the cell voltages
    [batch, numX, numY, row, col].  dtype follows G.
    """
    _check_input(input, "input")
    _check_input(G, "G")
    if row_size < 2:
        raise RuntimeError("row_sizeshould be >=2")
    if col_size < 2:
        raise RuntimeError("col_sizeshould be >=2")
    if G.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"ir_drop_gs not implemented for '{G.dtype}'")
    if input.dtype != G.dtype:
        raise RuntimeError("expected input and G to have the same dtype")
    if input.numel() != batch_size * numX * row_size or G.numel() != numX * row_size * numY * col_size:
        raise RuntimeError("input/G sizes do not match batch_size, numX, numY, row_size, col_size")
    nl = list(nonlinear_params) if nonlinear_cell_enable else []
    if nonlinear_cell_enable and len(nl) < 6:
        raise RuntimeError("nonlinear_params needs [Ron, Roff, a_on, b_on, a_off, b_off]")
    if mode == "converged" and not nonlinear_cell_enable:
        return _converged(input, G, batch_size, row_size, col_size, numX, numY, g_wire, ret_V)
    dev = G.device
    if ret_V:
        out = torch.empty((batch_size, numX, numY, row_size, col_size), dtype=G.dtype, device=dev)
    else:
        out = torch.empty((batch_size, numX, numY * col_size), dtype=G.dtype, device=dev)
    d = nat.HookDesc()
    d.dtype_bytes = G.element_size()
    d.d_input, d.d_G = input.data_ptr(), G.data_ptr()
    d.batch_size, d.row_size, d.col_size = int(batch_size), int(row_size), int(col_size)
    d.numX, d.numY, d.iter_times = int(numX), int(numY), int(iter_times)
    d.g_wire, d.beta = float(g_wire), float(beta)
    d.g_load_left = d.g_load_down = float(g_wire)
    d.enable_break, d.break_threshold = int(bool(enable_break)), float(break_threshould)
    d.nonlinear = int(bool(nonlinear_cell_enable))
    for k in range(6):
        d.nonlinear_params[k] = float(nl[k]) if k < len(nl) else 0.0
    d.ret_v = int(bool(ret_V))
    d.d_out = out.data_ptr()
    d.d_sweeps = None
    with torch.cuda.device(dev):
        rc = nat.lib().fadesim_hook_ir_drop_gs(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(dev)))
    nat.check(rc, "ir_drop_gs")
    return out


def _converged(input, G, batch_size, row, col, numX, numY, g_wire, ret_V):
    """All crossbars to `tol` through fadesim_solve_batch (linear cells).

    The hook's network has the source and sink conductance equal to g_wire
    and grounded bitlines (split_arr_matmul.cu:274-275, 290-291): uniform
    wire resistance 1/g_wire, vout = 0."""
    from .solver import solve_batch
    B = batch_size * numX * numY
    # conductance slot x*numY + y holds G[x, :, y*col:(y+1)*col], shared by the batch
    slots = G.view(numX, row, numY, col).permute(0, 2, 1, 3).reshape(numX * numY, row, col)
    slots = slots.to(torch.float32).contiguous()
    gidx = torch.arange(numX * numY, dtype=torch.int32, device=G.device).repeat(batch_size)
    vin = input.view(batch_size, numX, 1, row).expand(batch_size, numX, numY, row).reshape(B, row)
    res = solve_batch(slots, 1.0 / float(g_wire), vin, gidx=gidx, tol=tol, want_field=bool(ret_V))
    if ret_V:
        # u - d = (vin - p) - q
        f = res.field.to(torch.float64)
        V = vin.to(torch.float64)[:, :, None] - f[:, 0] - f[:, 1]
        return V.view(batch_size, numX, numY, row, col).to(G.dtype)
    return res.iout.view(batch_size, numX, numY * col).to(G.dtype)


def ir2s(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int, col_size: int, numX: int,
         numY: int, iter_times: int, r_wire: float, enable_break: bool, break_threshould: float,
         nonlinear_cell_enable: bool, nonlinear_params: Sequence[float], ret_V: bool = False) -> torch.Tensor:
    """IR2S solve of ``batch_size x numX x numY`` crossbars (the hook's
    ``ir2s``, split_arr_matmul.cu:446-545): ``iter_times`` iterations of
    exact wordline and bitline solves for the present cell currents, both
    layers from the same currents, source and sink segments ``r_wire``,
    grounded bitlines; with ``enable_break`` the loop ends after the first
    iteration in which no node of the whole batch moved by more than
    ``break_threshould``.  Same shapes, dtype rule and errors as ir_drop_gs;
    the arithmetic is the reference's, operation for operation
    (``fadesim_ir2s`` with the hook's settings)."""
    _check_input(input, "input")
    _check_input(G, "G")
    if row_size < 2:
        raise RuntimeError("row_sizeshould be >=2")
    if col_size < 2:
        raise RuntimeError("col_sizeshould be >=2")
    if G.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"ir2s not implemented for '{G.dtype}'")
    if input.dtype != G.dtype:
        raise RuntimeError("expected input and G to have the same dtype")
    if input.numel() != batch_size * numX * row_size or G.numel() != numX * row_size * numY * col_size:
        raise RuntimeError("input/G sizes do not match batch_size, numX, numY, row_size, col_size")
    nl = list(nonlinear_params) if nonlinear_cell_enable else []
    if nonlinear_cell_enable and len(nl) < 6:
        raise RuntimeError("nonlinear_params needs [Ron, Roff, a_on, b_on, a_off, b_off]")
    dev = G.device
    if ret_V:
        out = torch.empty((batch_size, numX, numY, row_size, col_size), dtype=G.dtype, device=dev)
    else:
        out = torch.empty((batch_size, numX, numY * col_size), dtype=G.dtype, device=dev)
    d = nat.Ir2sDesc()
    d.dtype_bytes, d.layout = G.element_size(), nat.IR2S_HOOK
    d.d_vin, d.d_G, d.d_vout = input.data_ptr(), G.data_ptr(), None
    d.batch_size, d.row_size, d.col_size = int(batch_size), int(row_size), int(col_size)
    d.numX, d.numY, d.iter_times = int(numX), int(numY), int(iter_times)
    d.r_wire, d.r_source, d.r_sink = float(r_wire), float(r_wire), float(r_wire)
    d.divide_by_g, d.carry_residue = 0, 1
    d.enable_break = nat.IR2S_BREAK_BATCH if enable_break else 0
    d.break_threshold = float(break_threshould)
    d.nonlinear = int(bool(nonlinear_cell_enable))
    for k in range(6):
        d.nonlinear_params[k] = float(nl[k]) if k < len(nl) else 0.0
    d.ret_v = int(bool(ret_V))
    d.d_out = out.data_ptr()
    d.d_vup = d.d_vdown = d.d_iters = None
    with torch.cuda.device(dev):
        rc = nat.lib().fadesim_ir2s(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(dev)))
    nat.check(rc, "ir2s")
    return out


def _check_common(fn, input, G, batch_size, row_size, col_size, numX, numY, nonlinear_cell_enable,
                  nonlinear_params):
    _check_input(input, "input")
    _check_input(G, "G")
    if row_size < 2:
        raise RuntimeError("row_sizeshould be >=2")
    if col_size < 2:
        raise RuntimeError("col_sizeshould be >=2")
    if G.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"{fn} not implemented for '{G.dtype}'")
    if input.dtype != G.dtype:
        raise RuntimeError("expected input and G to have the same dtype")
    if input.numel() != batch_size * numX * row_size or G.numel() != numX * row_size * numY * col_size:
        raise RuntimeError("input/G sizes do not match batch_size, numX, numY, row_size, col_size")
    nl = list(nonlinear_params) if nonlinear_cell_enable else []
    if nonlinear_cell_enable and len(nl) < 6:
        raise RuntimeError("nonlinear_params needs [Ron, Roff, a_on, b_on, a_off, b_off]")
    return nl + [0.0] * (6 - len(nl))


def _check_ou(row_size, col_size, ou_row_size, ou_col_size):
    if ou_row_size < 2 or ou_col_size < 2 or row_size % ou_row_size or col_size % ou_col_size:
        raise RuntimeError("OU sizes must be >= 2 and divide row_size and col_size")


def _gs_call(fn, input, G, batch_size, row_size, col_size, numX, numY, iter_times, g_wire, g_load_left,
             g_load_down, beta, enable_break, break_threshould, nonlinear_cell_enable, nonlinear_params,
             ou_rows, ou_cols):
    """One fadesim_hook_ir_drop_gs call; returns [batch, numX, num_ou_X, numY*col]."""
    nl = _check_common(fn, input, G, batch_size, row_size, col_size, numX, numY, nonlinear_cell_enable,
                       nonlinear_params)
    if ou_rows or ou_cols:
        _check_ou(row_size, col_size, ou_rows, ou_cols)
    nOuX = row_size // ou_rows if ou_rows else 1
    out = torch.empty((batch_size, numX, nOuX, numY * col_size), dtype=G.dtype, device=G.device)
    d = nat.HookDesc()
    d.dtype_bytes = G.element_size()
    d.d_input, d.d_G = input.data_ptr(), G.data_ptr()
    d.batch_size, d.row_size, d.col_size = int(batch_size), int(row_size), int(col_size)
    d.numX, d.numY, d.iter_times = int(numX), int(numY), int(iter_times)
    d.g_wire, d.beta = float(g_wire), float(beta)
    d.g_load_left, d.g_load_down = float(g_load_left), float(g_load_down)
    d.ou_rows, d.ou_cols = int(ou_rows), int(ou_cols)
    d.enable_break, d.break_threshold = int(bool(enable_break)), float(break_threshould)
    d.nonlinear = int(bool(nonlinear_cell_enable))
    for k in range(6):
        d.nonlinear_params[k] = float(nl[k])
    d.ret_v = 0
    d.d_out = out.data_ptr()
    d.d_sweeps = None
    with torch.cuda.device(G.device):
        rc = nat.lib().fadesim_hook_ir_drop_gs(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(G.device)))
    nat.check(rc, fn)
    return out


def _ir2s_call(fn, input, G, batch_size, row_size, col_size, numX, numY, iter_times, r_wire, r_source, r_sink,
               enable_break, break_threshould, nonlinear_cell_enable, nonlinear_params, ou_rows, ou_cols):
    """One fadesim_ir2s call (hook layout); returns [batch, numX, num_ou_X, numY*col]."""
    nl = _check_common(fn, input, G, batch_size, row_size, col_size, numX, numY, nonlinear_cell_enable,
                       nonlinear_params)
    if ou_rows or ou_cols:
        _check_ou(row_size, col_size, ou_rows, ou_cols)
    nOuX = row_size // ou_rows if ou_rows else 1
    out = torch.empty((batch_size, numX, nOuX, numY * col_size), dtype=G.dtype, device=G.device)
    d = nat.Ir2sDesc()
    d.dtype_bytes, d.layout = G.element_size(), nat.IR2S_HOOK
    d.d_vin, d.d_G, d.d_vout = input.data_ptr(), G.data_ptr(), None
    d.batch_size, d.row_size, d.col_size = int(batch_size), int(row_size), int(col_size)
    d.numX, d.numY, d.iter_times = int(numX), int(numY), int(iter_times)
    d.r_wire, d.r_source, d.r_sink = float(r_wire), float(r_source), float(r_sink)
    d.ou_rows, d.ou_cols = int(ou_rows), int(ou_cols)
    d.divide_by_g, d.carry_residue = 0, 1
    d.enable_break = nat.IR2S_BREAK_BATCH if enable_break else 0
    d.break_threshold = float(break_threshould)
    d.nonlinear = int(bool(nonlinear_cell_enable))
    for k in range(6):
        d.nonlinear_params[k] = float(nl[k])
    d.ret_v = 0
    d.d_out = out.data_ptr()
    d.d_vup = d.d_vdown = d.d_iters = None
    with torch.cuda.device(G.device):
        rc = nat.lib().fadesim_ir2s(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(G.device)))
    nat.check(rc, fn)
    return out


def ir_gs_ou_local(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int, col_size: int,
                   numX: int, numY: int, iter_times: int, g_wire: float, g_load_left: float, g_load_down: float,
                   beta: float, enable_break: bool, break_threshould: float, nonlinear_cell_enable: bool,
                   nonlinear_params: Sequence[float]) -> torch.Tensor:
    """The hook's line-GS iteration with its own source and sink load
    conductances (``ir_gs_ou_local``, split_arr_matmul.cu:690-815): the
    wordline's first segment conducts ``g_load_left``, the bitline's last
    ``g_load_down``, inner segments ``g_wire``.  Returns the bitline currents
    [batch, numX, numY*col].  Loads must be > 0 (an open line end is not
    supported)."""
    if not (g_load_left > 0 and g_load_down > 0):
        raise RuntimeError("g_load_left and g_load_down must be > 0")
    return _gs_call("ir_gs_ou_local", input, G, batch_size, row_size, col_size, numX, numY, iter_times, g_wire,
                    g_load_left, g_load_down, beta, enable_break, break_threshould, nonlinear_cell_enable,
                    nonlinear_params, 0, 0).view(batch_size, numX, numY * col_size)


def ir_gs_ou(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int, col_size: int,
             ou_row_size: int, ou_col_size: int, numX: int, numY: int, iter_times: int, g_wire: float,
             beta: float, enable_break: bool, break_threshould: float, nonlinear_cell_enable: bool,
             nonlinear_params: Sequence[float]) -> torch.Tensor:
    """OU-mode line-GS (``ir_gs_ou``, split_arr_matmul.cu:1086-1327, called
    from splitArrayArithmetic.py:187-188): every ``ou_row_size x ou_col_size``
    block of a tile is activated alone, so it is solved as its own crossbar
    whose wordline is driven through the c0+1 wire segments left of it
    (load g_wire/(c0+1)) and whose bitline drains through the row_size - r1
    segments below its last row r1 (load g_wire/(row_size - r1)).  Returns
    the per-OU bitline currents [batch, numX*num_ou_X, numY*col]."""
    out = _gs_call("ir_gs_ou", input, G, batch_size, row_size, col_size, numX, numY, iter_times, g_wire,
                   g_wire, g_wire, beta, enable_break, break_threshould, nonlinear_cell_enable, nonlinear_params,
                   max(int(ou_row_size), 1), max(int(ou_col_size), 1))
    return out.view(batch_size, -1, numY * col_size)


def ir2s_ou_local(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int, col_size: int,
                  numX: int, numY: int, iter_times: int, r_wire: float, r_load_left: float, r_load_down: float,
                  enable_break: bool, break_threshould: float, nonlinear_cell_enable: bool,
                  nonlinear_params: Sequence[float]) -> torch.Tensor:
    """The hook's IR2S iteration with its own source and sink resistances
    (``ir2s_ou_local``, split_arr_matmul.cu:546-688): first wordline segment
    ``r_load_left``, last bitline segment ``r_load_down``.  Returns the
    bitline currents [batch, numX, numY*col]."""
    if not (r_load_left > 0 and r_load_down > 0):
        raise RuntimeError("r_load_left and r_load_down must be > 0")
    return _ir2s_call("ir2s_ou_local", input, G, batch_size, row_size, col_size, numX, numY, iter_times, r_wire,
                      r_load_left, r_load_down, enable_break, break_threshould, nonlinear_cell_enable,
                      nonlinear_params, 0, 0).view(batch_size, numX, numY * col_size)


def ir2s_ou(input: torch.Tensor, G: torch.Tensor, batch_size: int, row_size: int, col_size: int,
            ou_row_size: int, ou_col_size: int, numX: int, numY: int, iter_times: int, r_wire: float,
            enable_break: bool, break_threshould: float, nonlinear_cell_enable: bool,
            nonlinear_params: Sequence[float]) -> torch.Tensor:
    """OU-mode IR2S (``ir2s_ou``, split_arr_matmul.cu:816-1085, called from
    splitArrayArithmetic.py:171-172): every OU block its own crossbar with
    source resistance r_wire*(c0+1) and sink resistance r_wire*(row_size - r1).
    Returns the per-OU bitline currents [batch, numX*num_ou_X, numY*col]."""
    out = _ir2s_call("ir2s_ou", input, G, batch_size, row_size, col_size, numX, numY, iter_times, r_wire, r_wire,
                     r_wire, enable_break, break_threshould, nonlinear_cell_enable, nonlinear_params,
                     max(int(ou_row_size), 1), max(int(ou_col_size), 1))
    return out.view(batch_size, -1, numY * col_size)


class _HookModule:
    """Module-like object with the reference extension's function names."""
    ir_drop_gs = staticmethod(ir_drop_gs)
    ir2s = staticmethod(ir2s)
    ir_gs_ou_local = staticmethod(ir_gs_ou_local)
    ir_gs_ou = staticmethod(ir_gs_ou)
    ir2s_ou_local = staticmethod(ir2s_ou_local)
    ir2s_ou = staticmethod(ir2s_ou)


def install(phyArrParams) -> None:
    """Make PimTorch call this implementation:
    ``phyArrParams.cpp_ir_drop_accelerate = <this module>``."""
    phyArrParams.cpp_ir_drop_accelerate = sys.modules[__name__]
