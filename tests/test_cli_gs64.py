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
def _gs64(cuda, g, RU, RD, vin, vout, iters, w, brk, tol, order, gen=None, init=None):
    """One fadesim_gs64_solve call through the C ABI.  gen = (on, v) for the
    general boundary.  Returns (vup, vdown, iout, record, sweeps)."""
    import torch
    from fadesim_amd import _native as nat
    n, m = g.shape
    T = lambda a, dt=torch.float64: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=cuda)
    keep = [T(g), T(RU), T(RD), T(vin), T(vout)]
    vup = torch.empty((n, m), dtype=torch.float64, device=cuda)
    vdn = torch.empty_like(vup)
    iout = torch.empty(m, dtype=torch.float64, device=cuda)
    ints = torch.zeros(2, dtype=torch.int32, device=cuda)
    d = nat.Gs64Desc()
    d.rows, d.cols = n, m
    d.d_g, d.d_ru, d.ru_ld = keep[0].data_ptr(), keep[1].data_ptr(), RU.shape[1]
    d.d_rd, d.rd_ld = keep[2].data_ptr(), RD.shape[1]
    d.d_vin, d.d_vout = keep[3].data_ptr(), keep[4].data_ptr()
    d.iter_number, d.omega, d.enable_break, d.tol = iters, w, int(brk), tol
    d.d_vup, d.d_vdown, d.d_iout = vup.data_ptr(), vdn.data_ptr(), iout.data_ptr()
    d.d_record, d.d_sweeps = ints.data_ptr(), ints.data_ptr() + 4
    d.order = order
    if gen is not None:
        on, v = gen
        don = [T(x, torch.int32) for x in on]
        dv = [T(x) for x in v]
        keep += don + dv
        d.general = 1
        d.d_on_top, d.d_on_left, d.d_on_bottom, d.d_on_right = (x.data_ptr() for x in don)
        d.d_v_top, d.d_v_right = dv[0].data_ptr(), dv[3].data_ptr()
        d.d_vin, d.d_vout = dv[1].data_ptr(), dv[2].data_ptr()
    if init is not None:
        iu, idn = T(init[0]), T(init[1])
        keep += [iu, idn]
        d.d_init_up, d.d_init_down = iu.data_ptr(), idn.data_ptr()
    nat.check(nat.lib().fadesim_gs64_solve(ctypes.byref(d), ctypes.c_void_p(nat.stream_handle(cuda))), "gs64")
    torch.cuda.synchronize()
    rec, sw = ints.cpu().tolist()
    return vup.cpu().numpy(), vdn.cpu().numpy(), iout.cpu().numpy(), rec, sw


def _rand_case(n, m, seed):
    rng = np.random.default_rng(seed)
    g = 1.0 / STATES4[rng.integers(0, 4, (n, m))]
    RU = rng.uniform(1, 5, (n, m + 1))
    RD = rng.uniform(1, 5, (m, n + 1))
    return g, RU, RD, rng.uniform(0, 1, n), rng.uniform(0, 0.05, m)


@pytest.mark.gpu
@pytest.mark.parametrize("n,m,iters,w,brk,tol", [(2, 2, 300, 1.9, True, 1e-13), (3, 5, 17, 1.5, False, 0.0),
                                                  (8, 4, 1000, 1.8, True, 1e-12), (33, 17, 400, 1.9, True, 1e-11),
                                                  (128, 128, 5000, 1.97, True, 1e-10)])
def test_gs64_redblack_bitwise_vs_replay(cuda, n, m, iters, w, brk, tol):
    from fadesim_amd import _native as nat
    g, RU, RD, vin, vout = _rand_case(n, m, n * 7 + m)
    got = _gs64(cuda, g, RU, RD, vin, vout, iters, w, brk, tol, nat.GS64_REDBLACK)
    u, dd, io, rec, sweeps = oracle.rb_f64(g, vin, vout, RU, RD, iters, w, brk, tol)
    assert got[3:] == (rec, sweeps)
    assert np.array_equal(got[0], u) and np.array_equal(got[1], dd) and np.array_equal(got[2], io)


# Shapes cover 1, 2, 4 and 8 rows per lane of the wavefront, non-square
# arrays, fixed counts and breaks.
LEX_CASES = [(2, 2, 300, 1.95, True, 1e-13), (3, 5, 17, 1.5, False, 0.0), (8, 4, 1000, 1.8, True, 1e-12),
             (33, 17, 400, 1.9, True, 1e-11), (64, 64, 60, 1.95, False, 0.0), (65, 3, 50, 1.7, False, 0.0),
             (128, 128, 3000, 1.97, True, 1e-10), (130, 40, 30, 1.9, False, 0.0), (200, 21, 25, 1.9, False, 0.0),
             (300, 8, 12, 1.93, False, 0.0), (512, 16, 5, 1.9, False, 0.0), (7, 300, 9, 1.8, False, 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("n,m,iters,w,brk,tol", LEX_CASES)
def test_gs64_lexicographic_bitwise_vs_oracle(cuda, n, m, iters, w, brk, tol):
    """The reference's own sweep order: every field value, the readout, the
    break sweep and the sweep count equal the CPU restatement's bit for bit."""
    from fadesim_amd import _native as nat
    g, RU, RD, vin, vout = _rand_case(n, m, n * 13 + m)
    print(f"lexicographic {n}x{m} iters={iters} w={w}", flush=True)
    vup, vdn, iout, rec, sw = _gs64(cuda, g, RU, RD, vin, vout, iters, w, brk, tol, nat.GS64_LEXICOGRAPHIC)
    u, dd, io, rec_o = oracle.nodegs(g, vin, vout, RU, RD, iters, w, brk, tol)
    assert rec == rec_o
    assert sw == (rec_o + 1 if rec_o >= 0 else iters)
    assert np.array_equal(vup, u)
    assert np.array_equal(vdn, dd)
    assert np.array_equal(iout, io)


@pytest.mark.gpu
def test_gs64_lexicographic_initial_state_and_ieee_division(cuda):
    """A given initial state (V1 mode bit 0) and a divisor whose reciprocal is
    subnormal, which switches the kernel to the IEEE division: still bit for bit."""
    from fadesim_amd import _native as nat
    n, m = 70, 33
    g, RU, RD, vin, vout = _rand_case(n, m, 4242)
    init = oracle.scichina_guess(g, vin, RD[0, 0])
    got = _gs64(cuda, g, RU, RD, vin, vout, 40, 1.9, False, 0.0, nat.GS64_LEXICOGRAPHIC, init=init)
    ref = oracle.nodegs(g, vin, vout, RU, RD, 40, 1.9, False, 0.0, init=init)
    assert all(np.array_equal(a, b) for a, b in zip(got[:3], ref[:3]))
    RU[5, 7] = 1e308  # 1/RU is subnormal
    got = _gs64(cuda, g, RU, RD, vin, vout, 40, 1.9, False, 0.0, nat.GS64_LEXICOGRAPHIC)
    ref = oracle.nodegs(g, vin, vout, RU, RD, 40, 1.9, False, 0.0)
    assert all(np.array_equal(a, b) for a, b in zip(got[:3], ref[:3]))


def _fixture_command(z):
    args = [int(z["iters"]), str(z["omega_text"]), int(z["mode"]) if "mode" in z else 0]
    if int(z["enable_break"]):
        args += ["yes", repr(float(z["tol"]))]
    return "nodebasedGSMethod " + " ".join(str(a) for a in args)


GS_FIXTURES = sorted(os.path.basename(p)[:-4] for p in
                     __import__("glob").glob(os.path.join(REPO, "tests", "golden", "*.npz"))
                     if not os.path.basename(p).startswith(("ir2s", "gen")))


@pytest.mark.gpu
@pytest.mark.parametrize("name", GS_FIXTURES)
def test_cli_reproduces_reference_output(cuda, tmp_path, name):
    """Each reference run re-run through the GPU CLI with its own command
    (config 1 is `nodebasedGSMethod 100000 1.95 0 yes 1e-8`; V1 fixtures with
    -v1): the printed currents are the reference's digits, the printed fields
    and "error I sum" equal its output, and the break sweep is the same."""
    z = dict(load_golden(name))
    n, m = int(z["n"]), int(z["m"])
    L = [f"arraySize {n} {m}", f"line_resistance {float(z['RD'][0, 0])!r}"]
    RU = z["RU"]
    for i in range(n):
        for j in range(m + 1):
            if RU[i, j] != z["RD"][0, 0]:
                L.append(f"setLineR up {i} {j} {float(RU[i, j])!r}")
    for j in range(m):
        for i in range(n + 1):
            if z["RD"][j, i] != z["RD"][0, 0]:
                L.append(f"setLineR down {j} {i} {float(z['RD'][j, i])!r}")
    L += ["setUseLine left -1", "setUseLine down -1"]
    L += [f"setLineV left {i} {float(v)!r}" for i, v in enumerate(z["vin"])]
    L += [f"setLineV down {j} {float(v)!r}" for j, v in enumerate(z["vout"])]
    L.append(f"cellRStates {len(z['cell_R'])} " + " ".join(repr(float(x)) for x in z["cell_R"]))
    L += [f"setCellR {i} {j} {int(z['states'][i, j])}" for i in range(n) for j in range(m)]
    L.append(_fixture_command(z))
    (tmp_path / "cfg").write_text("\n".join(L) + "\n")
    argv = [SIM] + (["-v1"] if int(z["version"]) == 1 else []) + [str(tmp_path / "cfg")]
    print(name, " ".join(argv[1:-1]), L[-1], flush=True)
    r = subprocess.run(argv, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert "Run end && Succeed" in r.stdout, r.stdout[-2000:]
    text = open(tmp_path / "fastmode_nodebasedGS.out").read().split("\n")
    k = text.index("iout = ")
    assert text[k + 1].split() == list(z["iout_txt"])
    err, vup, vdn, iout = _parse_out(tmp_path / "fastmode_nodebasedGS.out", n, m)
    assert err == float(z["errsum"])
    if "vup" in z:
        assert np.array_equal(vup, z["vup"]) and np.array_equal(vdn, z["vdown"])
    else:
        assert np.array_equal(vup[z["rows"]], z["vup_rows"]) and np.array_equal(vdn[z["rows"]], z["vdown_rows"])
        assert np.array_equal(vup.sum(1), z["vup_rowsum"]) and np.array_equal(vdn.sum(1), z["vdown_rowsum"])
    log = open(tmp_path / "default.log").read()
    if int(z["record"]) >= 0:
        assert f"threshould break, iter times =  {int(z['record'])}\n" in log
    else:
        assert "threshould break" not in log
    if name.startswith("c1_"):
        t = float(re.search(r"Duration time = ([0-9.eE+-]+) s", log).group(1))
        print(f"{name}: GPU solve {t:.4f} s, reference CPU solve {float(z['solve_seconds'][0]):.4f} s")


@pytest.mark.gpu
def test_cli_redblack_opt_in(cuda, tmp_path):
    """FADESIM_GS64_REDBLACK=1: red-black order, converged to the same fixed point."""
    z = load_golden("c2_v2_3")
    n, m = int(z["n"]), int(z["m"])
    L = [f"arraySize {n} {m}", f"line_resistance {float(z['RD'][0, 0])!r}", "setUseLine left -1",
         "setUseLine down -1"]
    L += [f"setLineV left {i} {float(v)!r}" for i, v in enumerate(z["vin"])]
    L.append(f"cellRStates {len(z['cell_R'])} " + " ".join(repr(float(x)) for x in z["cell_R"]))
    L += [f"setCellR {i} {j} {int(z['states'][i, j])}" for i in range(n) for j in range(m)]
    L.append("nodebasedGSMethod 200000 1.97 0 yes 1e-13")
    (tmp_path / "cfg").write_text("\n".join(L) + "\n")
    env = dict(os.environ, FADESIM_GS64_REDBLACK="1")
    r = subprocess.run([SIM, str(tmp_path / "cfg")], cwd=tmp_path, capture_output=True, text=True, env=env,
                       timeout=300)
    assert "(GPU, red-black)" in r.stdout, r.stdout[-1000:]
    err, vup, vdn, iout = _parse_out(tmp_path / "fastmode_nodebasedGS.out", n, m)
    g = 1.0 / z["cell_R"][z["states"]]
    _, _, iex = oracle.exact(g, z["vin"], z["vout"], z["RD"], z["RD"])
    assert np.max(np.abs(iout - iex) / np.abs(iex)) < 1e-9


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
def test_gs64_general_bitwise(cuda, name, iters, brk, tol):
    """General boundary, both orders: lexicographic equal to the restatement of
    nodeBasedGeneralGS, red-black equal to its CPU replay, bit for bit."""
    from fadesim_amd import _native as nat
    n, m, g, r, RU, RD, on, v = _gen(load_golden(name))
    got = _gs64(cuda, g, RU, RD, v[1], v[2], iters, 1.9, brk, tol, nat.GS64_LEXICOGRAPHIC, gen=(on, v))
    u, dd, io, rec = oracle.nodegs_general(g, RU, RD, on, v, iters, 1.9, brk, tol)
    assert got[3] == rec and got[4] == (rec + 1 if rec >= 0 else iters)
    assert np.array_equal(got[0], u) and np.array_equal(got[1], dd) and np.array_equal(got[2], io)
    got = _gs64(cuda, g, RU, RD, v[1], v[2], iters, 1.9, brk, tol, nat.GS64_REDBLACK, gen=(on, v))
    u, dd, io, rec, sweeps = oracle.rb_f64_general(g, RU, RD, on, v, iters, 1.9, brk, tol)
    assert got[3:] == (rec, sweeps)
    assert np.array_equal(got[0], u) and np.array_equal(got[1], dd) and np.array_equal(got[2], io)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GEN_CASES)
def test_cli_general_boundary(cuda, tmp_path, name):
    """V1's dispatcher: a general boundary also runs the general solver and
    writes fastmode_nodebasedGeneralGS.out -- with the reference's digits and
    break sweep."""
    z = dict(load_golden(name))
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
    r = subprocess.run([SIM, "-v1", str(tmp_path / "cfg")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    out = r.stdout
    assert "using general GS-method" in out, out[-400:]
    text = open(tmp_path / "fastmode_nodebasedGeneralGS.out").read().split("\n")
    k = text.index("iout = ")
    assert text[k + 1].split() == list(z["iout_txt"])
    err, vup, vdn, iout = _parse_out(str(tmp_path / "fastmode_nodebasedGeneralGS.out"), n, m)
    assert np.array_equal(vup, z["vup"]) and np.array_equal(vdn, z["vdown"])
    gen_out = out.split("using general GS-method", 1)[1]
    if int(z["record"]) >= 0:
        assert f"threshould break, iter times =  {int(z['record'])}" in gen_out
    else:
        assert "threshould break" not in gen_out
