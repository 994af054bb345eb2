"""CPU oracle for the FADESim IR-drop solver path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, and only as the checker.  The product
(``fadesim_amd``) never imports it.

numpy wrappers over ``liboracle.so`` (oracle/fadesim_oracle.c):

* :func:`nodegs`       lexicographic SOR restating
  ``NodebasedGSMethod::nodeBaseGS`` (crossbarHspice/V2/concrete_command.h:1136-1361)
  and ``array2D::nodeBasedGS`` (crossbarHspice/2dArray.cc:1377-1687).
* :func:`exact`        exact FP64 banded-Cholesky solve of the same network.
* :func:`exact_batch`  the same, batched over instances with uniform wire r.
* :func:`rb_f64`       FP64 red-black replay of the GPU drop-in kernel.
* :func:`rb_f32`       FP32 red-black replay of the GPU batched kernel.
* :func:`hook_gs`      the PimTorch hook's line-GS ordering
  (pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/split_arr_matmul.cu:27-333).
* :func:`errisum`      the partial-KCL "error I sum" (concrete_command.h:1293-1319).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        i, d = ctypes.c_int, ctypes.c_double
        L.oracle_nodegs.argtypes = [i, i, _dp, i, _dp, i, _dp, _dp, _dp, i, d, i, d,
                                    _dp, _dp, _dp, _ip]
        L.oracle_nodegs.restype = i
        L.oracle_nodegs_init.argtypes = [i, i, _dp, i, _dp, i, _dp, _dp, _dp, _dp, _dp, i, d, i, d,
                                         _dp, _dp, _dp, ctypes.POINTER(i)]
        L.oracle_nodegs_init.restype = i
        L.oracle_scichina_guess.argtypes = [i, i, d, _dp, _dp, _dp, _dp]
        L.oracle_scichina_guess.restype = None
        L.oracle_exact_nonlinear_batch.argtypes = [i, i, i, _fp, _dp, _dp, _dp, d, i, _dp, _ip]
        L.oracle_exact_nonlinear_batch.restype = None
        L.oracle_errisum.argtypes = [i, i, d, _dp, _dp, _dp]
        L.oracle_errisum.restype = d
        L.oracle_exact.argtypes = [i, i, _dp, i, _dp, i, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_exact.restype = i
        L.oracle_exact_batch_f32g.argtypes = [i, i, i, _fp, _dp, _dp, _dp, _dp]
        L.oracle_exact_batch_f32g.restype = i
        L.oracle_nodegs_batch_f32g.argtypes = [i, i, i, _fp, _dp, _dp, i, d, d, _dp, _ip]
        L.oracle_nodegs_batch_f32g.restype = i
        L.oracle_rb_f64.argtypes = [i, i, _dp, i, _dp, i, _dp, _dp, _dp, i, d, i, d,
                                    _dp, _dp, _dp, _ip]
        L.oracle_rb_f64.restype = i
        L.oracle_rb_f32.argtypes = [i, i, _fp, d, _dp, _dp, d, i, _fp, _fp]
        L.oracle_rb_f32.restype = None
        L.oracle_exact_general.argtypes = [i, i, _dp, i, _dp, i, _dp] + [_ip] * 4 + [_dp] * 4 + [_dp, _dp, _dp]
        L.oracle_exact_general.restype = i
        L.oracle_nodegs_general.argtypes = ([i, i, _dp, i, _dp, i, _dp] + [_ip] * 4 + [_dp] * 4 +
                                            [i, d, i, d, _dp, _dp, _dp, _ip])
        L.oracle_nodegs_general.restype = i
        L.oracle_rb_f64_general.argtypes = ([i, i, _dp, i, _dp, i, _dp] + [_ip] * 4 + [_dp] * 4 +
                                            [i, d, i, d, _dp, _dp, _dp, _ip])
        L.oracle_rb_f64_general.restype = i
        L.oracle_errisum_general.argtypes = [i, i, d, _dp, _dp, _dp]
        L.oracle_errisum_general.restype = d
        L.oracle_ir2s.argtypes = [i, i, _dp, _dp, _dp, d, d, i, i, d, _dp, _dp, _dp]
        L.oracle_ir2s.restype = i
        for name, fp in (("oracle_hook_ir2s_f64", _dp), ("oracle_hook_ir2s_f32", _fp)):
            f = getattr(L, name)
            f.argtypes = [fp, fp, i, i, i, i, i, i, d, i, d, i, _dp, i, fp]
            f.restype = i
        for name, fp in (("oracle_hook_gs_f64", _dp), ("oracle_hook_gs_f32", _fp)):
            f = getattr(L, name)
            f.argtypes = [fp, fp, i, i, i, i, i, i, d, d, i, d, i, _dp, i, fp]
            f.restype = i
        for name, fp in (("oracle_hook_gs_ou_f64", _dp), ("oracle_hook_gs_ou_f32", _fp)):
            f = getattr(L, name)
            f.argtypes = [fp, fp, i, i, i, i, i, i, d, d, d, d, i, d, i, _dp, i, i, i, fp]
            f.restype = i
        for name, fp in (("oracle_hook_ir2s_ou_f64", _dp), ("oracle_hook_ir2s_ou_f32", _fp)):
            f = getattr(L, name)
            f.argtypes = [fp, fp, i, i, i, i, i, i, d, d, d, i, d, i, _dp, i, i, i, fp]
            f.restype = i
        _LIB = L
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pf(a):
    return a.ctypes.data_as(_fp)


def segments_uniform(n: int, m: int, r: float):
    """RU [n, m+1] and RD [m, n+1] filled with one wire resistance."""
    return np.full((n, m + 1), float(r)), np.full((m, n + 1), float(r))


def nodegs(g, vin, vout, RU, RD, iters, w, enable_break=False, tol=0.0, init=None):
    """Lexicographic SOR (from the initial state `init` = (vup0, vdown0) when
    given).  Returns (vup, vdown, iout, record)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    vup = np.empty((n, m))
    vdn = np.empty((n, m))
    iout = np.empty(m)
    rec = ctypes.c_int(0)
    iu, idn = (None, None) if init is None else (_d(init[0]), _d(init[1]))
    rc = lib().oracle_nodegs_init(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                                  _pd(_d(vin)), _pd(_d(vout)), None if iu is None else _pd(iu),
                                  None if idn is None else _pd(idn), int(iters), float(w),
                                  int(bool(enable_break)), float(tol), _pd(vup), _pd(vdn),
                                  _pd(iout), ctypes.byref(rec))
    if rc != 0:
        raise ValueError("m and n should greater than 2")
    return vup, vdn, iout, rec.value


def exact_nonlinear_batch(glin, r, vin, params, tol=1e-14, max_iter=60):
    """Sinh-nonlinear cells converged: Picard iteration of the hook's cell
    model around exact FP64 solves (uniform wire r per instance, grounded
    bitlines).  glin [B, n, m] float32, r [B] or scalar, vin [B, n].
    Returns (iout [B, m], solves [B])."""
    glin = np.ascontiguousarray(glin, dtype=np.float32)
    B, n, m = glin.shape
    r = np.ascontiguousarray(np.broadcast_to(np.asarray(r, dtype=np.float64), (B,)))
    vin = _d(np.asarray(vin).reshape(B, n))
    prm = _d(np.asarray(params, dtype=np.float64).reshape(6))
    iout = np.empty((B, m))
    its = np.empty(B, dtype=np.int32)
    lib().oracle_exact_nonlinear_batch(B, n, m, glin.ctypes.data_as(_fp), _pd(r), _pd(vin), _pd(prm), float(tol),
                                       int(max_iter), _pd(iout), its.ctypes.data_as(_ip))
    if (its < 0).any():
        raise RuntimeError(f"oracle_exact_nonlinear_batch failed: {its.tolist()}")
    return iout, its


def scichina_guess(g, vin, r_seg):
    """V1 mode bit 0's initial state (2dArray.cc:1425-1535).  Returns (vup0, vdown0)."""
    g = _d(g)
    n, m = g.shape
    vu, vd = np.empty((n, m)), np.empty((n, m))
    lib().oracle_scichina_guess(n, m, float(r_seg), _pd(g), _pd(_d(vin)), _pd(vu), _pd(vd))
    return vu, vd


def nodegs_batch(g32, r, vin, iters, w, tol):
    """Lexicographic SOR (reference rule, break at max|dv| < tol) over a batch
    [B, n, m] of FP32 conductances with one wire resistance each and grounded
    bitlines, OpenMP over instances.  Returns (iout [B, m], sweeps [B], threads)."""
    g32 = _f(g32)
    B, n, m = g32.shape
    r = _d(np.broadcast_to(np.asarray(r, dtype=np.float64), (B,)))
    vin = _d(vin).reshape(B, n)
    iout = np.empty((B, m))
    sweeps = np.empty(B, dtype=np.int32)
    nt = lib().oracle_nodegs_batch_f32g(B, n, m, _pf(g32), _pd(r), _pd(vin), int(iters), float(w),
                                        float(tol), _pd(iout), sweeps.ctypes.data_as(_ip))
    return iout, sweeps, nt


def errisum(g, vup, vdown, g_wire):
    g = _d(g)
    n, m = g.shape
    return lib().oracle_errisum(n, m, float(g_wire), _pd(g), _pd(_d(vup)), _pd(_d(vdown)))


def exact(g, vin, vout, RU, RD):
    """Exact FP64 solve.  Returns (u, d, iout)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    u = np.empty((n, m))
    d = np.empty((n, m))
    iout = np.empty(m)
    rc = lib().oracle_exact(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                            _pd(_d(vin)), _pd(_d(vout)), _pd(u), _pd(d), _pd(iout))
    if rc != 0:
        raise RuntimeError(f"oracle_exact failed ({rc})")
    return u, d, iout


def exact_batch(g32, r, vin, vout=None):
    """Exact FP64 readout for a batch [B, n, m] of FP32 conductances with one
    wire resistance per instance.  Returns iout [B, m] (FP64)."""
    g32 = _f(g32)
    B, n, m = g32.shape
    r = _d(np.broadcast_to(np.asarray(r, dtype=np.float64), (B,)))
    vin = _d(vin).reshape(B, n)
    vo = None if vout is None else _d(vout).reshape(B, m)
    iout = np.empty((B, m))
    rc = lib().oracle_exact_batch_f32g(B, n, m, _pf(g32), _pd(r), _pd(vin),
                                       None if vo is None else _pd(vo), _pd(iout))
    if rc != 0:
        raise RuntimeError("oracle_exact_batch_f32g failed")
    return iout


def exact_slots(g32, r, slot, vin):
    """Exact FP64 readout of K drives over S FP32 conductance slots [S, n, m]
    (uniform wire resistance r[S] or scalar, vout = 0): instance k uses slot
    slot[k] and wordline drive vin[k].  Returns iout [K, m]."""
    g32 = _f(g32)
    S, n, m = g32.shape
    r = _d(np.broadcast_to(np.asarray(r, dtype=np.float64), (S,)))
    slot = np.ascontiguousarray(slot, dtype=np.int32).reshape(-1)
    K = slot.size
    if K and (slot.min() < 0 or slot.max() >= S):
        raise ValueError("slot out of range")
    vin = _d(vin).reshape(K, n)
    iout = np.empty((K, m))
    rc = lib().oracle_exact_slots_f32g(S, n, m, _pf(g32), _pd(r), K,
                                       slot.ctypes.data_as(ctypes.c_void_p), _pd(vin), _pd(iout))
    if rc != 0:
        raise RuntimeError("oracle_exact_slots_f32g failed")
    return iout


def rb_f64(g, vin, vout, RU, RD, iters, w, enable_break=False, tol=0.0):
    """FP64 red-black SOR.  Returns (u, d, iout, record, sweeps)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    u = np.empty((n, m))
    d = np.empty((n, m))
    iout = np.empty(m)
    rec = ctypes.c_int(0)
    sweeps = lib().oracle_rb_f64(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                                 _pd(_d(vin)), _pd(_d(vout)), int(iters), float(w),
                                 int(bool(enable_break)), float(tol), _pd(u), _pd(d),
                                 _pd(iout), ctypes.byref(rec))
    return u, d, iout, rec.value, sweeps


def rb_f32(g32, r, vin, vout, w, sweeps, p=None, q=None):
    """FP32 red-black SOR on drop variables for exactly `sweeps` sweeps.
    Returns (p, q) (FP32)."""
    g32 = _f(g32)
    n, m = g32.shape
    p = np.zeros((n, m), np.float32) if p is None else _f(p).copy()
    q = np.zeros((n, m), np.float32) if q is None else _f(q).copy()
    vo = None if vout is None else _d(vout)
    lib().oracle_rb_f32(n, m, _pf(g32), float(r), _pd(_d(vin)),
                        None if vo is None else _pd(vo), float(w), int(sweeps), _pf(p), _pf(q))
    return p, q


def hook_gs(inp, G, batch_size, row_size, col_size, numX, numY, iter_times, g_wire, beta,
            enable_break, break_threshold, nonlinear_cell_enable, nonlinear_params,
            ret_V=False):
    """Line-GS of the PimTorch hook.  dtype follows G (float32 or float64).
    Returns (out, sweeps) with out [bsize, numX, numY*col] or
    [bsize, numX, numY, row, col] when ret_V."""
    f64 = np.asarray(G).dtype == np.float64
    cv = _d if f64 else _f
    pp = _pd if f64 else _pf
    inp, G = cv(inp), cv(G)
    nl = _d(list(nonlinear_params) + [0.0] * (6 - len(nonlinear_params)))
    if ret_V:
        out = np.empty((batch_size, numX, numY, row_size, col_size), G.dtype)
    else:
        out = np.empty((batch_size, numX, numY * col_size), G.dtype)
    f = lib().oracle_hook_gs_f64 if f64 else lib().oracle_hook_gs_f32
    sweeps = f(pp(inp), pp(G), batch_size, row_size, col_size, numX, numY, int(iter_times),
               float(g_wire), float(beta), int(bool(enable_break)), float(break_threshold),
               int(bool(nonlinear_cell_enable)), _pd(nl), int(bool(ret_V)), pp(out))
    return out, sweeps


def ir2s(g, vin, vout, g_wire, r_load, times, enable_break=False, break_th=1e-9, vup=None, vdown=None):
    """IR2S wire-solve iteration, voltage form (ir_drop_fast_solver.h:24-163).
    Starts from vup = vin (broadcast), vdown = 0 unless given.  Returns
    (vup, vdown, iout, count) with the reference's iteration count."""
    g = _d(g)
    n, m = g.shape
    vin, vout = _d(vin).reshape(n), _d(vout).reshape(m)
    vu = np.repeat(vin[:, None], m, 1).copy() if vup is None else _d(vup).copy()
    vd = np.zeros((n, m)) if vdown is None else _d(vdown).copy()
    out = np.empty(m)
    cnt = lib().oracle_ir2s(n, m, _pd(g), _pd(vin), _pd(vout), float(g_wire), float(r_load), int(times),
                            int(bool(enable_break)), float(break_th), _pd(vu), _pd(vd), _pd(out))
    if cnt < 0:
        raise MemoryError("oracle_ir2s")
    return vu, vd, out, cnt


def hook_ir2s(inp, G, batch_size, row_size, col_size, numX, numY, iter_times, r_wire, enable_break,
              break_threshould, nonlinear_cell_enable, nonlinear_params, ret_V=False):
    """The hook's ir2s (split_arr_matmul.cu:336-545) on the CPU, in the dtype
    of G (float32 or float64).  Returns (out, iterations)."""
    f32 = np.asarray(G).dtype == np.float32
    cv = _f if f32 else _d
    inp, G = cv(inp), cv(G)
    nl = _d(list(nonlinear_params) + [0.0] * (6 - len(nonlinear_params)) if nonlinear_cell_enable else np.zeros(6))
    if ret_V:
        out = np.empty((batch_size, numX, numY, row_size, col_size), dtype=G.dtype)
    else:
        out = np.empty((batch_size, numX, numY * col_size), dtype=G.dtype)
    fn = lib().oracle_hook_ir2s_f32 if f32 else lib().oracle_hook_ir2s_f64
    pp = _pf if f32 else _pd
    it = fn(pp(inp), pp(G), int(batch_size), int(row_size), int(col_size), int(numX), int(numY), int(iter_times),
            float(r_wire), int(bool(enable_break)), float(break_threshould), int(bool(nonlinear_cell_enable)),
            _pd(nl), int(bool(ret_V)), pp(out))
    return out, it


def hook_gs_ou(inp, G, batch_size, row_size, col_size, numX, numY, iter_times, g_wire, beta, enable_break,
               break_threshold, nonlinear_cell_enable, nonlinear_params, g_load_left=None, g_load_down=None,
               ou_rows=0, ou_cols=0):
    """Line-GS with separate source / sink loads (ir_gs_ou_local,
    split_arr_matmul.cu:690-815) or in OU mode (ir_gs_ou, :1086-1327; loads
    from the block's position).  Returns (out, sweeps), out
    [bsize, numX, num_ou_X, numY*col] (num_ou_X = 1 without OUs)."""
    f64 = np.asarray(G).dtype == np.float64
    cv, pp = (_d, _pd) if f64 else (_f, _pf)
    inp, G = cv(inp), cv(G)
    nl = _d(list(nonlinear_params) + [0.0] * (6 - len(nonlinear_params)))
    nOuX = row_size // ou_rows if ou_rows > 0 else 1
    out = np.empty((batch_size, numX, nOuX, numY * col_size), G.dtype)
    gl = g_wire if g_load_left is None else g_load_left
    gd = g_wire if g_load_down is None else g_load_down
    f = lib().oracle_hook_gs_ou_f64 if f64 else lib().oracle_hook_gs_ou_f32
    sweeps = f(pp(inp), pp(G), batch_size, row_size, col_size, numX, numY, int(iter_times), float(g_wire),
               float(gl), float(gd), float(beta), int(bool(enable_break)), float(break_threshold),
               int(bool(nonlinear_cell_enable)), _pd(nl), 0, int(ou_rows), int(ou_cols), pp(out))
    return out, sweeps


def hook_ir2s_ou(inp, G, batch_size, row_size, col_size, numX, numY, iter_times, r_wire, enable_break,
                 break_threshould, nonlinear_cell_enable, nonlinear_params, r_source=None, r_sink=None,
                 ou_rows=0, ou_cols=0):
    """The hook's IR2S with separate source / sink resistances (ir2s_ou_local,
    split_arr_matmul.cu:546-688) or in OU mode (ir2s_ou, :816-1085).
    Returns (out, iterations), out [bsize, numX, num_ou_X, numY*col]."""
    f32 = np.asarray(G).dtype == np.float32
    cv, pp = (_f, _pf) if f32 else (_d, _pd)
    inp, G = cv(inp), cv(G)
    nl = _d(list(nonlinear_params) + [0.0] * (6 - len(nonlinear_params)) if nonlinear_cell_enable else np.zeros(6))
    nOuX = row_size // ou_rows if ou_rows > 0 else 1
    out = np.empty((batch_size, numX, nOuX, numY * col_size), G.dtype)
    rs = r_wire if r_source is None else r_source
    rk = r_wire if r_sink is None else r_sink
    fn = lib().oracle_hook_ir2s_ou_f32 if f32 else lib().oracle_hook_ir2s_ou_f64
    it = fn(pp(inp), pp(G), int(batch_size), int(row_size), int(col_size), int(numX), int(numY), int(iter_times),
            float(r_wire), float(rs), float(rk), int(bool(enable_break)), float(break_threshould),
            int(bool(nonlinear_cell_enable)), _pd(nl), 0, int(ou_rows), int(ou_cols), pp(out))
    return out, it


def _edges(n, m, on, v):
    """on / v: four arrays (top [m], left [n], bottom [m], right [n])."""
    sizes = (m, n, m, n)
    on_c = [np.ascontiguousarray(np.asarray(on[e], dtype=np.int32).reshape(sizes[e])) for e in range(4)]
    v_c = [_d(np.asarray(v[e], dtype=np.float64).reshape(sizes[e])) for e in range(4)]
    return on_c, v_c


def exact_general(g, RU, RD, on, v):
    """Exact solve of the general-boundary network.  Returns (u, d, iout) with
    iout the sink wire currents of driven bottoms (2dArray.cc:1353-1356)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    on_c, v_c = _edges(n, m, on, v)
    u, dn, io = np.empty((n, m)), np.empty((n, m)), np.empty(m)
    rc = lib().oracle_exact_general(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                                    *[a.ctypes.data_as(_ip) for a in on_c], *[_pd(a) for a in v_c],
                                    _pd(u), _pd(dn), _pd(io))
    if rc == -3:
        raise np.linalg.LinAlgError("singular network (a floating line)")
    if rc:
        raise MemoryError("oracle_exact_general")
    return u, dn, io


def nodegs_general(g, RU, RD, on, v, iters, w, enable_break=False, tol=0.0):
    """nodeBasedGeneralGS restated.  Returns (vup, vdown, iout, record)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    on_c, v_c = _edges(n, m, on, v)
    vu, vd, io = np.empty((n, m)), np.empty((n, m)), np.empty(m)
    rec = ctypes.c_int(0)
    lib().oracle_nodegs_general(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                                *[a.ctypes.data_as(_ip) for a in on_c], *[_pd(a) for a in v_c],
                                int(iters), float(w), int(bool(enable_break)), float(tol), _pd(vu), _pd(vd),
                                _pd(io), ctypes.byref(rec))
    return vu, vd, io, rec.value


def errisum_general(g, vup, vdown, g_wire):
    g = _d(g)
    n, m = g.shape
    return lib().oracle_errisum_general(n, m, float(g_wire), _pd(g), _pd(_d(vup)), _pd(_d(vdown)))


def rb_f64_general(g, RU, RD, on, v, iters, w, enable_break=False, tol=0.0):
    """FP64 red-black replay of the GPU general-boundary solve.  Returns
    (vup, vdown, iout, record, sweeps)."""
    g = _d(g)
    n, m = g.shape
    RU, RD = _d(RU), _d(RD)
    on_c, v_c = _edges(n, m, on, v)
    vu, vd, io = np.empty((n, m)), np.empty((n, m)), np.empty(m)
    rec = ctypes.c_int(0)
    sw = lib().oracle_rb_f64_general(n, m, _pd(RU), RU.shape[1], _pd(RD), RD.shape[1], _pd(g),
                                     *[a.ctypes.data_as(_ip) for a in on_c], *[_pd(a) for a in v_c],
                                     int(iters), float(w), int(bool(enable_break)), float(tol), _pd(vu), _pd(vd),
                                     _pd(io), ctypes.byref(rec))
    return vu, vd, io, rec.value, sw
