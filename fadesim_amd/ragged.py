"""Mixed-size batched IR-drop solve (64x64 .. 512x512 crossbars in one launch).

Host layer over the plan API of include/fadesim.h (``fadesim_ragged_*``):
``RaggedPlan(rows, cols)`` fixes the sizes and owns the device memory of the
solve; ``plan.solve(...)`` launches one persistent kernel for the whole mixed
batch; ``plan.capture(...)`` records that launch into a HIP graph (through
torch's graph API) for replay with no host work.  Per-size relaxation factors
and test intervals come from the autotuner's table (fadesim_amd.autotune)
when one is present for the device.

Packing (instance k in input order): conductances / fields at the running sum
of rows*cols (fields: two planes p, q per instance), wordline drives at the
running sum of rows, bitline currents / sink voltages at the running sum of
cols.  ``pack`` / ``unpack`` convert between lists of per-instance arrays and
the packed tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as nat

SIZES = (64, 128, 256, 512)


@dataclass
class RaggedResult:
    iout: torch.Tensor            # packed [sum cols] float64
    sweeps: torch.Tensor          # [K] int32
    status: torch.Tensor          # [K] int32
    est: torch.Tensor             # [K] float64
    field: Optional[torch.Tensor]  # packed [2 * sum rows*cols] float32 or None


class RaggedPlan:
    def __init__(self, rows: Sequence[int], cols: Optional[Sequence[int]] = None, device=None,
                 tuning: Optional[dict] = None):
        self.rows = np.asarray(list(rows), dtype=np.int32)
        self.cols = self.rows.copy() if cols is None else np.asarray(list(cols), dtype=np.int32)
        if self.rows.shape != self.cols.shape:
            raise ValueError("rows and cols differ in length")
        self.count = int(self.rows.size)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        cells = self.rows.astype(np.int64) * self.cols
        self.g_off = np.concatenate([[0], np.cumsum(cells)[:-1]]).astype(np.int64) if self.count else np.zeros(0, np.int64)
        self.vin_off = np.concatenate([[0], np.cumsum(self.rows)[:-1]]).astype(np.int64) if self.count else np.zeros(0, np.int64)
        self.out_off = np.concatenate([[0], np.cumsum(self.cols)[:-1]]).astype(np.int64) if self.count else np.zeros(0, np.int64)
        self.cells = int(cells.sum())
        self.rows_total = int(self.rows.sum())
        self.cols_total = int(self.cols.sum())
        if tuning is None:
            from . import autotune
            tuning = autotune.load_table(self.device)
        self.tuning = tuning
        opts = (nat.RaggedClassOpts * 4)()
        if tuning:
            for c, n in enumerate(SIZES):
                t = tuning.get(str(n))
                if t:
                    opts[c].omega = float(t.get("omega", 0.0))
                    opts[c].check_first = int(t.get("check_first", 0))
                    opts[c].check_min = int(t.get("check_min", 0))
                    opts[c].check_max = int(t.get("check_max", 0))
                    opts[c].ir2s_iters = int(t.get("ir2s_iters", 0))
                    opts[c].corr_iters = int(t.get("corr_iters", 0))
                    opts[c].ir2s_omega = float(t.get("ir2s_omega", 0.0))
        h = ctypes.c_void_p()
        L = nat.lib()
        with torch.cuda.device(self.device):
            rc = L.fadesim_ragged_plan_create(ctypes.byref(h), self.count,
                                              self.rows.ctypes.data_as(ctypes.c_void_p),
                                              self.cols.ctypes.data_as(ctypes.c_void_p), ctypes.cast(opts, ctypes.c_void_p))
        nat.check(rc, "fadesim_ragged_plan_create")
        self._h = h
        st = nat.RaggedPlanStats()
        nat.check(L.fadesim_ragged_plan_info(self._h, ctypes.byref(st)), "fadesim_ragged_plan_info")
        self.stats = {"grid": st.grid, "teams": st.teams, "big_tasks": st.big_tasks, "small_items": st.small_items,
                      "omega": list(st.omega), "check_first": list(st.check_first),
                      "check_min": list(st.check_min), "check_max": list(st.check_max),
                      "ir2s_iters": list(st.ir2s_iters), "corr_iters": list(st.corr_iters),
                      "ir2s_omega": list(st.ir2s_omega)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                nat.lib().fadesim_ragged_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------ packing
    def pack(self, arrays: List[np.ndarray], what: str, dtype=None) -> np.ndarray:
        """Concatenate per-instance arrays in the plan's packed layout
        (what = 'g' [rows, cols], 'field' [2, rows, cols], 'vin' [rows],
        'cols' [cols])."""
        if len(arrays) != self.count:
            raise ValueError("one array per instance")
        for k, a in enumerate(arrays):
            a = np.asarray(a)
            exp = {"g": (self.rows[k], self.cols[k]), "field": (2, self.rows[k], self.cols[k]),
                   "vin": (self.rows[k],), "cols": (self.cols[k],)}[what]
            if a.shape != tuple(int(x) for x in exp):
                raise ValueError(f"instance {k}: shape {a.shape}, expected {exp}")
        flat = [np.asarray(a).reshape(-1) for a in arrays]
        out = np.concatenate(flat) if flat else np.zeros(0)
        return out.astype(dtype) if dtype is not None else out

    def unpack(self, packed, what: str) -> List[np.ndarray]:
        x = packed.detach().cpu().numpy() if torch.is_tensor(packed) else np.asarray(packed)
        res = []
        for k in range(self.count):
            r, c = int(self.rows[k]), int(self.cols[k])
            if what == "cols":
                res.append(x[self.out_off[k]:self.out_off[k] + c])
            elif what == "vin":
                res.append(x[self.vin_off[k]:self.vin_off[k] + r])
            elif what == "g":
                res.append(x[self.g_off[k]:self.g_off[k] + r * c].reshape(r, c))
            else:
                res.append(x[2 * self.g_off[k]:2 * self.g_off[k] + 2 * r * c].reshape(2, r, c))
        return res

    # ------------------------------------------------------------ solve
    def _desc(self, g, r_wire, vin, vout, warm, tol, safety, switch_tol, max_sweeps, flags, out):
        d = nat.RaggedDesc()
        d.d_g, d.d_r_wire, d.d_vin = g.data_ptr(), r_wire.data_ptr(), vin.data_ptr()
        d.d_vout = None if vout is None else vout.data_ptr()
        d.d_warm = None if warm is None else warm.data_ptr()
        d.tol, d.safety, d.switch_tol = float(tol), float(safety), float(switch_tol)
        d.max_sweeps, d.flags = int(max_sweeps), int(flags)
        d.d_iout, d.d_sweeps, d.d_status, d.d_est = (out.iout.data_ptr(), out.sweeps.data_ptr(),
                                                     out.status.data_ptr(), out.est.data_ptr())
        d.d_field = None if out.field is None else out.field.data_ptr()
        return d

    def _check_inputs(self, g, r_wire, vin, vout, warm):
        dev = self.device
        def chk(t, n, dt, name):
            if t is None:
                return None
            if not torch.is_tensor(t) or t.device != dev or t.dtype != dt or not t.is_contiguous() or t.numel() != n:
                raise ValueError(f"{name} must be a contiguous {dt} tensor of {n} elements on {dev}")
            return t
        chk(g, self.cells, torch.float32, "g")
        chk(r_wire, self.count, torch.float64, "r_wire")
        chk(vin, self.rows_total, torch.float64, "vin")
        chk(vout, self.cols_total, torch.float64, "vout")
        chk(warm, 2 * self.cells, torch.float32, "warm")

    def alloc_result(self, want_field=False) -> RaggedResult:
        dev = self.device
        return RaggedResult(torch.empty(self.cols_total, dtype=torch.float64, device=dev),
                            torch.empty(self.count, dtype=torch.int32, device=dev),
                            torch.empty(self.count, dtype=torch.int32, device=dev),
                            torch.empty(self.count, dtype=torch.float64, device=dev),
                            torch.empty(2 * self.cells, dtype=torch.float32, device=dev) if want_field else None)

    def solve(self, g: torch.Tensor, r_wire: torch.Tensor, vin: torch.Tensor, vout: Optional[torch.Tensor] = None,
              warm: Optional[torch.Tensor] = None, *, tol: float = 1e-6, safety: float = 2.0,
              switch_tol: float = 0.0, max_sweeps: int = 100000, fixed_sweeps: Optional[int] = None,
              refine: bool = True, want_field: bool = False, timing: bool = False,
              zero_correction: bool = False, test_first: bool = True,
              out: Optional[RaggedResult] = None, check_abort: bool = True) -> RaggedResult:
        """Solve every crossbar of the plan; packed float32 g / warm, float64
        r_wire [K], vin, vout (None = 0 V).  ``test_first``: size classes whose
        controls apply IR2S iterations to the initial guess test it before
        the first sweep."""
        self._check_inputs(g, r_wire, vin, vout, warm)
        res = out if out is not None else self.alloc_result(want_field)
        flags = 0
        if fixed_sweeps is not None:
            flags |= nat.F_FIXED_SWEEPS
            max_sweeps = int(fixed_sweeps)
        if not refine:
            flags |= nat.F_NO_REFINE
        if timing:
            flags |= nat.F_TIMING
        if zero_correction:
            flags |= nat.F_ZERO_CORRECTION
        if test_first:
            flags |= nat.F_TEST_FIRST
        d = self._desc(g, r_wire, vin, vout, warm, tol, safety, switch_tol, max_sweeps, flags, res)
        with torch.cuda.device(self.device):
            rc = nat.lib().fadesim_ragged_solve(self._h, ctypes.byref(d),
                                                ctypes.c_void_p(nat.stream_handle(self.device)))
        nat.check(rc, "fadesim_ragged_solve")
        if check_abort and self.aborted():
            raise RuntimeError("mixed-size solve aborted: a team wait timed out")
        return res

    def aborted(self) -> bool:
        a = ctypes.c_int(0)
        with torch.cuda.device(self.device):
            nat.check(nat.lib().fadesim_ragged_aborted(self._h, ctypes.byref(a)), "fadesim_ragged_aborted")
        return bool(a.value)

    def last_ms(self) -> float:
        return float(nat.lib().fadesim_ragged_last_ms(self._h))

    def capture(self, g, r_wire, vin, vout=None, warm=None, *, tol=1e-6, safety=2.0, switch_tol=0.0,
                max_sweeps=100000, want_field=False, test_first=True):
        """Record the solve (fill kernel + persistent kernel) into a graph.
        Returns (graph, result): ``graph.replay()`` re-solves from the current
        contents of the same input tensors into ``result``."""
        self._check_inputs(g, r_wire, vin, vout, warm)
        res = self.alloc_result(want_field)
        # one eager call first: nothing may be allocated during capture
        self.solve(g, r_wire, vin, vout, warm, tol=tol, safety=safety, switch_tol=switch_tol,
                   max_sweeps=max_sweeps, out=res, test_first=test_first)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        d = self._desc(g, r_wire, vin, vout, warm, tol, safety, switch_tol, max_sweeps,
                       nat.F_TEST_FIRST if test_first else 0, res)
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(graph, stream=s):
                rc = nat.lib().fadesim_ragged_solve(self._h, ctypes.byref(d),
                                                    ctypes.c_void_p(nat.stream_handle(self.device)))
        nat.check(rc, "fadesim_ragged_solve (capture)")
        torch.cuda.current_stream(self.device).wait_stream(s)
        return graph, res
