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
    d.r_wire, d.r_source, d.r_sink = float(r_wire), 0.0, 0.0
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


class _HookModule:
    """Module-like object with the reference extension's function names."""
    ir_drop_gs = staticmethod(ir_drop_gs)
    ir2s = staticmethod(ir2s)


def install(phyArrParams) -> None:
    """Make PimTorch call this implementation:
    ``phyArrParams.cpp_ir_drop_accelerate = <this module>``."""
    phyArrParams.cpp_ir_drop_accelerate = sys.modules[__name__]
