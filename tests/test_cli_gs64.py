"""The `sim` CLI drop-in and its FP64 solvers (fadesim_gs64_solve, fadesim_ir2s).

GPU: the FP64 red-black kernel equals its CPU replay (oracle_rb_f64) bit for
bit, sweep count included; the CLI run on the reference's own configs agrees
with the reference's golden output to the precision both reach, and writes
the reference's file formats; `fastsolve` (IR2S) reproduces the reference
binary's fastmode.out digits and iteration count.  CPU: command parsing and
error behaviour (no GPU call is reached).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import REPO, STATES4, load_golden

SIM = os.path.join(REPO, "fadesim_amd", "lib", "sim")


def _write_cfg(path, n, m, states, cell_R, r, vin, vout, gs_args, v1_names=False, extra=(),
               cmd="nodebasedGSMethod"):
    L = [f"arraySize {n} {m}", f"line_resistance {r!r}", "setUseLine left -1", "setUseLine down -1",
         "senseBitlineI down -1", "// a comment", "% another", ""]
    L += list(extra)
    for i, v in enumerate(vin):
        L.append(f"setLineV left {i} {float(v)!r}")
    for j, v in enumerate(vout):
        L.append(f"setLineV down {j} {float(v)!r}")
    kw = "cellRstates" if v1_names else "cellRStates"
    L.append(f"{kw} {len(cell_R)} " + " ".join(repr(float(x)) for x in cell_R))
    for i in range(n):
        for j in range(m):
            L.append(f"setCellR {i} {j} {int(states[i, j])}")
    L.append(cmd + " " + " ".join(str(a) for a in gs_args))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def _parse_out(path, n, m):
    lines = open(path).read().split("\n")
    err = float(lines[0].split("=")[1])
    k = lines.index("vup = ")
    vup = np.array([[float(x) for x in lines[k + 1 + i].split()] for i in range(n)])
    k = lines.index("vdown = ")
    vdn = np.array([[float(x) for x in lines[k + 1 + i].split()] for i in range(n)])
    k = lines.index("iout = ")
    iout = np.array([float(x) for x in lines[k + 1].split()])
    return err, vup, vdn, iout


def _run(cfg, cwd):
    r = subprocess.run([SIM, cfg], cwd=cwd, capture_output=True, text=True, timeout=300)
    return r.stdout


# ------------------------------------------------------------------ CPU side
def test_cli_unknown_command_and_bad_args(tmp_path):
    if not os.path.exists(SIM):
        pytest.fail("sim not built")
    cfg = tmp_path / "c1"
    cfg.write_text("arraySize 2 2\nfrobnicate 1 2\n")
    out = _run(str(cfg), tmp_path)
    assert "Error(s) occured in line 2: Unknown command: frobnicate" in out
    cfg.write_text("arraySize 2 2\ncellRStates 1 1e6\nnodebasedGSMethod 10 1.5 0 yes\n")
    assert "args number not matched (!=5)" in _run(str(cfg), tmp_path)
    cfg.write_text("arraySize 2 2\ncellRStates 1 1e6\nnodebasedGSMethod 10 1.5\n")
    assert "args number not matched (<3)" in _run(str(cfg), tmp_path)
    cfg.write_text("arraySize 1 4\ncellRStates 1 1e6\nline_resistance 1\nnodebasedGSMethod 10 1.5 0\n")
    assert "m and n should greater than 2" in _run(str(cfg), tmp_path)
    cfg.write_text("arraySize 2 2\ncellRStates 1 1e6\nline_resistance 1\nnodebasedGSMethod 10 1.5 1\n")
    assert "mode bit 0 is not supported now" in _run(str(cfg), tmp_path)
    cfg.write_text("arraySize 2 2\nsetCellR 5 0 1\n")
    assert "out of range" in _run(str(cfg), tmp_path)
    env = dict(os.environ, FADESIM_STRICT_EXIT="1")
    r = subprocess.run([SIM, str(cfg)], cwd=tmp_path, capture_output=True, text=True, env=env)
    assert r.returncode == 1


def test_cli_mode2_reports_initial_guess(tmp_path):
    """mode bit 1 skips the sweeps: the output is the initial state (no GPU call)."""
    n = m = 3
    st = np.zeros((n, m), int)
    _write_cfg(tmp_path / "cfg", n, m, st, [1e6], 2.0, [1.0, 0.5, 0.25], [0, 0, 0], [100, 1.5, 2])
    out = _run(str(tmp_path / "cfg"), tmp_path)
    assert "Run end && Succeed" in out, out
    err, vup, vdn, iout = _parse_out(tmp_path / "fastmode_nodebasedGS.out", n, m)
    np.testing.assert_allclose(vup, np.repeat([[1.0], [0.5], [0.25]], 3, 1))
    np.testing.assert_allclose(iout, [1.75e-6] * 3, rtol=1e-12)
    log = open(tmp_path / "default.log").read()
    assert "--------Operation Log---------" in log and "iter times = 100" in log


# ------------------------------------------------------------------ GPU side
@pytest.mark.gpu
@pytest.mark.parametrize("n,m,iters,w,brk,tol", [(2, 2, 300, 1.9, True, 1e-13), (3, 5, 17, 1.5, False, 0.0),
                                                  (8, 4, 1000, 1.8, True, 1e-12), (33, 17, 400, 1.9, True, 1e-11),
                                                  (128, 128, 5000, 1.97, True, 1e-10)])
def test_gs64_bitwise_vs_replay(cuda, n, m, iters, w, brk, tol):
    import torch
    from fadesim_amd import _native as nat
    rng = np.random.default_rng(n * 7 + m)
    g = 1.0 / STATES4[rng.integers(0, 4, (n, m))]
    RU = rng.uniform(1, 5, (n, m + 1))
    RD = rng.uniform(1, 5, (m, n + 1))
    vin = rng.uniform(0, 1, n)
    vout = rng.uniform(0, 0.05, m)
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)
    dg, dru, drd, dvin, dvout = T(g), T(RU), T(RD), T(vin), T(vout)
    vup = torch.empty((n, m), dtype=torch.float64, device=cuda)
    vdn = torch.empty_like(vup)
    iout = torch.empty(m, dtype=torch.float64, device=cuda)
    ints = torch.zeros(2, dtype=torch.int32, device=cuda)
    d = nat.Gs64Desc()
    d.rows, d.cols = n, m
    d.d_g, d.d_ru, d.ru_ld, d.d_rd, d.rd_ld = dg.data_ptr(), dru.data_ptr(), m + 1, drd.data_ptr(), n + 1
    d.d_vin, d.d_vout = dvin.data_ptr(), dvout.data_ptr()
    d.iter_number, d.omega, d.enable_break, d.tol = iters, w, int(brk), tol
    d.d_vup, d.d_vdown, d.d_iout = vup.data_ptr(), vdn.data_ptr(), iout.data_ptr()
    d.d_record, d.d_sweeps = ints.data_ptr(), ints.data_ptr() + 4
    nat.check(nat.lib().fadesim_gs64_solve(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(cuda))), "gs64")
    torch.cuda.synchronize()
    u, dd, io, rec, sweeps = oracle.rb_f64(g, vin, vout, RU, RD, iters, w, brk, tol)
    assert ints.cpu().tolist() == [rec, sweeps]
    assert np.array_equal(vup.cpu().numpy(), u)
    assert np.array_equal(vdn.cpu().numpy(), dd)
    assert np.array_equal(iout.cpu().numpy(), io)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_v2_tol1e-12", "c2_v2_0", "c2_v2_3", "size64_v2", "small0_v2_2x2",
                                  "small4_v2_16x16", "small0_v1_3x5", "small1_v1_8x4", "small3_v1_2x7"])
def test_cli_matches_reference_golden(cuda, tmp_path, name):
    """Converged runs: the GPU red-black CLI at tol 1e-13 against the reference
    output at its (tight) tolerance -- both within ~1e-8 of the fixed point."""
    z = load_golden(name)
    n, m = int(z["n"]), int(z["m"])
    L = [f"arraySize {n} {m}"]
    # per-segment resistances exactly as the fixture's RU/RD
    L.append(f"line_resistance {float(z['RD'][0, 0])!r}")
    RU = z["RU"]  # V2 fixtures are square with one wire resistance: RU and RD agree
    for i in range(n):
        for j in range(m + 1):
            L.append(f"setLineR up {i} {j} {float(RU[i, j])!r}")
    for j in range(m):
        for i in range(n + 1):
            L.append(f"setLineR down {j} {i} {float(z['RD'][j, i])!r}")
    L += [f"setLineV left {i} {float(v)!r}" for i, v in enumerate(z["vin"])]
    L += [f"setLineV down {j} {float(v)!r}" for j, v in enumerate(z["vout"])]
    L.append(f"cellRStates {len(z['cell_R'])} " + " ".join(repr(float(x)) for x in z["cell_R"]))
    L += [f"setCellR {i} {j} {int(z['states'][i, j])}" for i in range(n) for j in range(m)]
    L.append("nodebasedGSMethod 200000 1.9 0 yes 1e-13")
    (tmp_path / "cfg").write_text("\n".join(L) + "\n")
    out = _run(str(tmp_path / "cfg"), tmp_path)
    assert "Run end && Succeed" in out, out[-2000:]
    err, vup, vdn, iout = _parse_out(tmp_path / "fastmode_nodebasedGS.out", n, m)
    g = 1.0 / z["cell_R"][z["states"]]
    _, _, iex = oracle.exact(g, z["vin"], z["vout"], RU, z["RD"])
    assert np.max(np.abs(iout - iex) / np.abs(iex)) < 1e-9
    if int(z["record"]) >= 0 and float(z["tol"]) <= 1e-10:
        # the reference stopped on max|dv| < tol, a few 1e-8 short of the fixed
        # point: the two answers differ by that iteration error and no more
        ref_err = np.max(np.abs(z["iout"] - iex) / np.abs(iex))
        assert ref_err < 1e-7
        assert np.max(np.abs(iout - z["iout"]) / np.abs(z["iout"])) <= ref_err + 2e-9
    log = open(tmp_path / "default.log").read()
    assert re.search(r"threshould break, iter times =  \d+", log)
    assert "Duration time = " in log


IR2S_CASES = sorted(os.path.basename(p)[:-4] for p in
                    __import__("glob").glob(os.path.join(REPO, "tests", "golden", "ir2s*.npz")))


@pytest.mark.gpu
@pytest.mark.parametrize("name", IR2S_CASES)
def test_cli_fastsolve_matches_reference(tmp_path, name):
    z = load_golden(name)
    n, m = int(z["n"]), int(z["m"])
    r = float(z["RD"][0, 0])
    rl = float(z["rload_left"])
    extra = () if np.isnan(rl) else (f"rload left {rl!r}",)
    args = [int(z["iters"]), 0, int(z["enable_break"]), repr(float(z["tol"]))]
    cfg = tmp_path / "cfg"
    _write_cfg(str(cfg), n, m, z["states"], z["cell_R"], r, z["vin"], z["vout"], args, v1_names=True,
               extra=extra, cmd="fastsolve")
    out = _run(str(cfg), tmp_path)
    assert "Error" not in out, out[:400]
    cnt = int(re.findall(r"ir2s fast iter times = (\d+)", out)[-1])
    assert cnt == int(z["record"])
    err, vup, vdn, iout = _parse_out(str(tmp_path / "fastmode.out"), n, m)
    np.testing.assert_allclose(iout, z["iout"], rtol=1e-12)
    assert err == pytest.approx(float(z["errsum"]), rel=1e-5, abs=1e-17)
    if "vup" in z:
        np.testing.assert_allclose(vup, z["vup"], rtol=1e-12)
        np.testing.assert_allclose(vdn, z["vdown"], rtol=1e-12, atol=1e-300)


def test_cli_fastsolve_rejects_table_mode(tmp_path):
    if not os.path.exists(SIM):
        pytest.fail("sim not built")
    cfg = tmp_path / "c"
    cfg.write_text("arraySize 2 2\ncellRstates 1 1e6\nline_resistance 1\nfastsolve 3 1\n")
    assert "initial voltages from a table are not supported" in _run(str(cfg), tmp_path)
    cfg.write_text("arraySize 2 2\ncellRstates 1 1e6\nline_resistance 1\nfastsolve\n")
    assert "fastsolve needs iter_times" in _run(str(cfg), tmp_path)


# ------------------------------------------------------------------ general boundary
GEN_CASES = sorted(os.path.basename(p)[:-4] for p in
                   __import__("glob").glob(os.path.join(REPO, "tests", "golden", "gen*.npz")))
DIRS = ("up", "left", "down", "right")


def _gen(z):
    n, m = int(z["n"]), int(z["m"])
    sizes = (m, n, m, n)
    on = [z["on"][e][:sizes[e]].astype(np.int32) for e in range(4)]
    v = [z["v"][e][:sizes[e]].astype(np.float64) for e in range(4)]
    g = 1.0 / z["cell_R"][z["states"]]
    r = float(z["r"])
    return n, m, g, r, np.full((n, m + 1), r), np.full((m, n + 1), r), on, v


@pytest.mark.gpu
@pytest.mark.parametrize("name", GEN_CASES)
@pytest.mark.parametrize("iters,brk,tol", [(37, False, 0.0), (100000, True, 1e-12)])
def test_gs64_general_bitwise_vs_replay(cuda, name, iters, brk, tol):
    import torch
    from fadesim_amd import _native as nat
    n, m, g, r, RU, RD, on, v = _gen(load_golden(name))
    T = lambda a, dt=torch.float64: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=cuda)
    dg, dru, drd = T(g), T(RU), T(RD)
    dv = [T(x) for x in v]
    don = [T(x, torch.int32) for x in on]
    vup = torch.empty((n, m), dtype=torch.float64, device=cuda)
    vdn = torch.empty_like(vup)
    iout = torch.empty(m, dtype=torch.float64, device=cuda)
    ints = torch.zeros(2, dtype=torch.int32, device=cuda)
    d = nat.Gs64Desc()
    d.rows, d.cols = n, m
    d.d_g, d.d_ru, d.ru_ld, d.d_rd, d.rd_ld = dg.data_ptr(), dru.data_ptr(), m + 1, drd.data_ptr(), n + 1
    d.d_vin, d.d_vout = dv[1].data_ptr(), dv[2].data_ptr()
    d.iter_number, d.omega, d.enable_break, d.tol = iters, 1.9, int(brk), tol
    d.d_vup, d.d_vdown, d.d_iout = vup.data_ptr(), vdn.data_ptr(), iout.data_ptr()
    d.d_record, d.d_sweeps = ints.data_ptr(), ints.data_ptr() + 4
    d.general = 1
    d.d_on_top, d.d_on_left, d.d_on_bottom, d.d_on_right = (x.data_ptr() for x in don)
    d.d_v_top, d.d_v_right = dv[0].data_ptr(), dv[3].data_ptr()
    nat.check(nat.lib().fadesim_gs64_solve(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(cuda))), "gs64")
    torch.cuda.synchronize()
    u, dd, io, rec, sweeps = oracle.rb_f64_general(g, RU, RD, on, v, iters, 1.9, brk, tol)
    assert ints.cpu().tolist() == [rec, sweeps]
    assert np.array_equal(vup.cpu().numpy(), u)
    assert np.array_equal(vdn.cpu().numpy(), dd)
    assert np.array_equal(iout.cpu().numpy(), io)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GEN_CASES)
def test_cli_general_boundary(cuda, tmp_path, name):
    """V1's dispatcher: a general boundary also runs the general solver and
    writes fastmode_nodebasedGeneralGS.out.  Where the reference converged
    (break taken) the CLI agrees with its printed digits; otherwise the CLI's
    red-black iterates are those of the CPU replay."""
    z = load_golden(name)
    n, m, g, r, RU, RD, on, v = _gen(z)
    L = [f"arraySize {n} {m}", f"line_resistance {r!r}"]
    for e, dname in enumerate(DIRS):
        L.append(f"setNotUseLine {dname} -1")
        idx = [k for k in range(len(on[e])) if on[e][k]]
        if idx:
            L.append(f"setUseLine {dname} " + " ".join(str(k) for k in idx))
        for k, x in enumerate(v[e]):
            L.append(f"setLineV {dname} {k} {float(x)!r}")
    L.append(f"cellRstates {len(z['cell_R'])} " + " ".join(repr(float(x)) for x in z["cell_R"]))
    for i in range(n):
        for j in range(m):
            L.append(f"setCellR {i} {j} {int(z['states'][i, j])}")
    args = [int(z["iters"]), z["omega_text"], 0] + (["yes", repr(float(z["tol"]))] if int(z["enable_break"]) else [])
    L.append("nodebasedGSMethod " + " ".join(str(a) for a in args))
    (tmp_path / "cfg").write_text("\n".join(L) + "\n")
    out = _run(str(tmp_path / "cfg"), tmp_path)
    assert "using general GS-method" in out, out[-400:]
    err, vup, vdn, iout = _parse_out(str(tmp_path / "fastmode_nodebasedGeneralGS.out"), n, m)
    w = float(np.float32(float(z["omega_text"])))
    u, dd, io, rec, sw = oracle.rb_f64_general(g, RU, RD, on, v, int(z["iters"]), w, bool(z["enable_break"]),
                                               float(z["tol"]))
    np.testing.assert_allclose(iout, io, rtol=1e-6, atol=1e-15)   # printed with 7 digits
    np.testing.assert_allclose(vup, u, rtol=1e-6, atol=1e-12)
    if int(z["record"]) != -1:  # the reference converged: same answer to its printed digits
        np.testing.assert_allclose(iout, z["iout"], rtol=2e-6, atol=1e-15)
        np.testing.assert_allclose(vup, z["vup"], rtol=2e-6, atol=1e-12)
