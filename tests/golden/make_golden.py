#!/usr/bin/env python3
"""Write the golden vectors of tests/golden/ by running the reference CLI.

The reference `sim` binaries (V1 and V2) are compiled from the reference's own
source files by ``make -C oracle ref`` into oracle/_ref/ (never committed).
This script writes seeded configs -- explicit ``setCellR`` / ``setLineV`` lines,
since the reference has no seeded random generator -- runs each binary in a
scratch directory and stores what it printed (``fastmode_nodebasedGS.out``,
``default.log`` or stdout) as small .npz fixtures.  Only data leaves this
script: the inputs (states, voltages, resistances) and the outputs.

Usage: python tests/golden/make_golden.py [--only PREFIX]   (needs /root/reference)
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")

STATES4 = [2e6, 1.6e6, 1.3e6, 1e6]
ONLY = ""  # --only PREFIX: write only the fixtures whose name starts with PREFIX


def parse_out(path, n, m):
    """Parse fastmode_nodebasedGS.out (format of concrete_command.h:1321-1353)."""
    with open(path) as f:
        lines = f.read().split("\n")
    errsum = float(lines[0].split("=")[1])
    k = lines.index("vup = ")
    vup = [[float(x) for x in lines[k + 1 + i].split()] for i in range(n)]
    k = lines.index("vdown = ")
    vdn = [[float(x) for x in lines[k + 1 + i].split()] for i in range(n)]
    k = lines.index("iout = ")
    iout = [float(x) for x in lines[k + 1].split()]
    # keep the printed text of iout too: the strings are the exact digits
    iout_txt = lines[k + 1].split()
    return errsum, np.array(vup), np.array(vdn), np.array(iout), iout_txt


def write_config(path, version, n, m, states, cell_R, r, vin, vout, gs_args,
                 seg_up=(), seg_down=(), solve_cmd="nodebasedGSMethod", rload_left=None):
    vin = [float(v) for v in vin]
    vout = [float(v) for v in vout]
    r = float(r)
    L = []
    L.append("// generated by tests/golden/make_golden.py")
    L.append(f"arraySize {n} {m}")
    if version == 1:
        L.append("selector no")
    L.append(f"line_resistance {r!r}")
    for (x, y, z) in seg_up:
        L.append(f"setLineR up {x} {y} {z!r}")
    for (x, y, z) in seg_down:
        L.append(f"setLineR down {x} {y} {z!r}")
    if rload_left is not None:
        L.append(f"rload left {float(rload_left)!r}")
    L.append("setUseLine left -1")
    L.append("setUseLine down -1")
    L.append("senseBitlineI down -1")
    L.append(f"setLineV left -1 {vin[0]!r}")
    for i, v in enumerate(vin):
        if v != vin[0]:
            L.append(f"setLineV left {i} {v!r}")
    L.append(f"setLineV down -1 {vout[0]!r}")
    for j, v in enumerate(vout):
        if v != vout[0]:
            L.append(f"setLineV down {j} {v!r}")
    kw = "cellRStates" if version == 2 else "cellRstates"
    L.append(f"{kw} {len(cell_R)} " + " ".join(repr(float(x)) for x in cell_R))
    base = int(np.bincount(states.ravel()).argmax())
    L.append(f"setCellR -1 -1 {base}")
    for i in range(n):
        for j in range(m):
            if states[i, j] != base:
                L.append(f"setCellR {i} {j} {int(states[i, j])}")
    L.append(solve_cmd + " " + " ".join(str(a) for a in gs_args))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def run_case(name, version, n, m, states, cell_R, r, vin, vout, gs_args, seg_up=(),
             seg_down=(), keep_fields=True, solve_cmd="nodebasedGSMethod", rload_left=None):
    if ONLY and not name.startswith(ONLY):
        return
    exe = os.path.join(REF_BIN, f"sim_v{version}")
    ir2s = solve_cmd == "fastsolve"
    with tempfile.TemporaryDirectory() as td:
        cfg = os.path.join(td, "cfg")
        write_config(cfg, version, n, m, states, cell_R, r, vin, vout, gs_args, seg_up, seg_down,
                     solve_cmd, rload_left)
        res = subprocess.run([exe, cfg], cwd=td, capture_output=True, text=True, timeout=3600)
        text = res.stdout
        if "Error" in text:
            raise RuntimeError(f"{name}: reference reported an error:\n{text[:300]}")
        if version == 2:
            with open(os.path.join(td, "default.log")) as f:
                text = f.read()
        outp = os.path.join(td, "fastmode.out" if ir2s else "fastmode_nodebasedGS.out")
        errsum, vup, vdn, iout, iout_txt = parse_out(outp, n, m)
    if ir2s:
        mrec = re.findall(r"ir2s fast iter times = (\d+)", text)
    else:
        mrec = re.findall(r"threshould break, iter times =\s+(\d+)", text)
    record = int(mrec[-1]) if mrec else -1
    dur = [float(x) for x in re.findall(r"(?:Duration )?time = ([0-9.eE+-]+) s", text)]
    RU = np.full((n, m + 1), float(r))
    RD = np.full((m, n + 1), float(r))
    for (x, y, z) in seg_up:
        RU[x, y] = z
    for (x, y, z) in seg_down:
        RD[x, y] = z
    data = dict(version=version, n=n, m=m, states=states.astype(np.uint8),
                cell_R=np.array(cell_R, float), RU=RU, RD=RD, vin=np.array(vin, float),
                vout=np.array(vout, float), iters=int(gs_args[0]),
                mode=0 if ir2s else int(gs_args[2]),
                omega_text="" if ir2s else str(gs_args[1]),
                enable_break=int(gs_args[2]) if ir2s else int(len(gs_args) > 3 and gs_args[3] == "yes"),
                tol=float(gs_args[3] if ir2s else (gs_args[4] if len(gs_args) > 4 else 0.0)),
                command=solve_cmd, rload_left=float("nan") if rload_left is None else float(rload_left),
                record=record, errsum=errsum, iout=iout, iout_txt=np.array(iout_txt),
                solve_seconds=np.array(dur))
    if keep_fields:
        data["vup"] = vup
        data["vdown"] = vdn
    else:
        rows = np.array([0, n // 2, n - 1])
        data["rows"] = rows
        data["vup_rows"] = vup[rows]
        data["vdown_rows"] = vdn[rows]
        data["vup_rowsum"] = vup.sum(1)
        data["vdown_rowsum"] = vdn.sum(1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
    print(f"{name}: v{version} {n}x{m} record={record} iout[0]={iout[0]:.6e} "
          f"t={dur}", flush=True)


DIRS = ("up", "left", "down", "right")


def run_general(name, n, m, states, cell_R, r, on, v, gs_args):
    """A general-boundary case through V1 (nodeBasedGeneralGS, chosen by the
    dispatcher when the top or right edge is driven or the left or bottom edge
    is not): per-edge setUseLine / setNotUseLine and setLineV."""
    if ONLY and not name.startswith(ONLY):
        return
    exe = os.path.join(REF_BIN, "sim_v1")
    L = ["// generated by tests/golden/make_golden.py", f"arraySize {n} {m}", "selector no", f"line_resistance {float(r)!r}"]
    for e, dname in enumerate(DIRS):
        L.append(f"setNotUseLine {dname} -1")
        idx = [k for k in range(len(on[e])) if on[e][k]]
        if idx:
            L.append(f"setUseLine {dname} " + " ".join(str(k) for k in idx))
        for k, x in enumerate(v[e]):
            L.append(f"setLineV {dname} {k} {float(x)!r}")
    L.append("senseBitlineI down -1")
    L.append(f"cellRstates {len(cell_R)} " + " ".join(repr(float(x)) for x in cell_R))
    for i in range(n):
        for j in range(m):
            L.append(f"setCellR {i} {j} {int(states[i, j])}")
    L.append("nodebasedGSMethod " + " ".join(str(a) for a in gs_args))
    with tempfile.TemporaryDirectory() as td:
        cfg = os.path.join(td, "cfg")
        with open(cfg, "w") as f:
            f.write("\n".join(L) + "\n")
        res = subprocess.run([exe, cfg], cwd=td, capture_output=True, text=True, timeout=3600)
        text = res.stdout
        if "using general GS-method" not in text:
            raise RuntimeError(f"{name}: the reference did not take the general solver:\n{text[-400:]}")
        text = text.split("using general GS-method", 1)[1]
        errsum, vup, vdn, iout, iout_txt = parse_out(os.path.join(td, "fastmode_nodebasedGeneralGS.out"), n, m)
    mrec = re.findall(r"threshould break, iter times =\s+(\d+)", text)
    record = int(mrec[0]) if mrec else -1
    np.savez_compressed(os.path.join(HERE, name + ".npz"), version=1, n=n, m=m, states=states.astype(np.uint8),
                        cell_R=np.array(cell_R, float), r=float(r),
                        on=np.array([np.pad(np.asarray(x, int), (0, max(n, m) - len(x))) for x in on]),
                        v=np.array([np.pad(np.asarray(x, float), (0, max(n, m) - len(x))) for x in v]),
                        iters=int(gs_args[0]), omega_text=str(gs_args[1]),
                        enable_break=int(len(gs_args) > 3 and gs_args[3] == "yes"),
                        tol=float(gs_args[4]) if len(gs_args) > 4 else 0.0, record=record, errsum=errsum,
                        vup=vup, vdown=vdn, iout=iout, iout_txt=np.array(iout_txt), command="general")
    print(f"{name}: v1 general {n}x{m} record={record} iout[0]={iout[0]:.6e}", flush=True)


def main():
    global ONLY
    if len(sys.argv) > 2 and sys.argv[1] == "--only":
        ONLY = sys.argv[2]
    if not (os.path.exists(os.path.join(REF_BIN, "sim_v2"))
            and os.path.exists(os.path.join(REF_BIN, "sim_v1"))):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)

    # Config 1: 128x128, 4 states, r = 2.93, all wordlines 1 V from the left.
    rng = np.random.default_rng(20240601)
    st = rng.integers(0, 4, (128, 128))
    ones, zeros = [1.0] * 128, [0.0] * 128
    for tol in ("1e-8", "1e-12"):
        run_case(f"c1_v2_tol{tol}", 2, 128, 128, st, STATES4, 2.93, ones, zeros,
                 [100000, 1.95, 0, "yes", tol], keep_fields=False)
    run_case("c1_v1_tol1e-8", 1, 128, 128, st, STATES4, 2.93, ones, zeros,
             [100000, 1.95, 0, "yes", "1e-8"], keep_fields=False)

    # Config-2 distribution: vin in {0, 0.2} V, r ~ U[1, 5] ohm, omega 1.97.
    for k in range(8):
        rng = np.random.default_rng(1000 + k)
        st = rng.integers(0, 4, (128, 128))
        r = float(np.round(rng.uniform(1.0, 5.0), 6))
        vin = list(np.where(rng.random(128) < 0.5, 0.2, 0.0))
        run_case(f"c2_v2_{k}", 2, 128, 128, st, STATES4, r, vin, zeros,
                 [100000, 1.97, 0, "yes", "1e-10"], keep_fields=False)

    # Other sizes (V2 handles square arrays only: RULine aliases RDLine).
    for n, w, args in ((64, 1.95, [100000, 1.95, 0, "yes", "1e-10"]),
                       (256, 1.98, [100000, 1.98, 0, "yes", "1e-9"]),
                       (512, 1.99, [150, 1.99, 0, "no", "0"])):
        rng = np.random.default_rng(7 + n)
        st = rng.integers(0, 4, (n, n))
        vin = list(np.where(rng.random(n) < 0.5, 0.2, 0.0))
        run_case(f"size{n}_v2", 2, n, n, st, STATES4, 2.93, vin, [0.0] * n, args,
                 keep_fields=False)

    # Small square cases through V2 with full fields; fixed counts and breaks.
    smalls = [(2, 2, [100000, 1.95, 0, "yes", "1e-12"]), (3, 3, [17, 1.5, 0]),
              (4, 4, [100000, 1.8, 0, "yes", "1e-11"]), (8, 8, [37, 1.9, 0]),
              (16, 16, [100000, 1.9, 0, "yes", "1e-10"]), (8, 8, [0, 1.9, 0])]
    for idx, (n, m, args) in enumerate(smalls):
        rng = np.random.default_rng(300 + idx)
        st = rng.integers(0, 4, (n, m))
        vin = list(np.round(rng.uniform(0, 1, n), 3))
        vout = list(np.round(rng.uniform(0, 0.05, m), 3))
        run_case(f"small{idx}_v2_{n}x{m}", 2, n, m, st, STATES4, 2.93, vin, vout, args)

    # Non-square and per-segment resistances through V1 (ω parsed with stof).
    v1cases = [(3, 5, [100000, 1.7, 0, "yes", "1e-12"], (), ()),
               (8, 4, [100000, 1.8, 0, "yes", "1e-12"], ((1, 2, 7.5), (5, 0, 1.25)),
                ((0, 3, 9.0), (2, 8, 0.5))),
               (5, 3, [23, 1.6, 0], ((4, 3, 2.0),), ((1, 1, 4.0),)),
               (2, 7, [100000, 1.9, 0, "yes", "1e-12"], (), ((6, 2, 11.0),))]
    for idx, (n, m, args, su, sd) in enumerate(v1cases):
        rng = np.random.default_rng(500 + idx)
        st = rng.integers(0, 4, (n, m))
        vin = list(np.round(rng.uniform(0, 1, n), 3))
        vout = list(np.round(rng.uniform(0, 0.05, m), 3))
        run_case(f"small{idx}_v1_{n}x{m}", 1, n, m, st, STATES4, 2.93, vin, vout, args,
                 seg_up=su, seg_down=sd)

    # V1 mode bit 0: the "sci_china" initial state (2dArray.cc:1425-1535), then
    # sweeps; with a break; with mode bit 1 (the initial state reported); none.
    sci_cases = [(6, 9, [60, 1.8, 1]), (16, 16, [100000, 1.9, 1, "yes", "1e-12"]), (7, 5, [10, 1.7, 3]),
                 (12, 12, [0, 1.9, 1])]
    for idx, (n, m, args) in enumerate(sci_cases):
        rng = np.random.default_rng(1100 + idx)
        st = rng.integers(0, 4, (n, m))
        vin = list(np.round(rng.uniform(0, 1, n), 3))
        run_case(f"sci{idx}_v1_{n}x{m}", 1, n, m, st, STATES4, 2.93, vin, [0.0] * m, args)

    # IR2S (`fastsolve <iters> 0 <enable_break> <break_th>`) through V1:
    # fixed sweep counts, the break rule, a source/sink load, non-square.
    ir2s_cases = [(8, 8, 2.93, [1, 0, 0, "1e-9"], None), (16, 16, 2.93, [5, 0, 0, "1e-9"], 5.0),
                  (12, 20, 4.1, [3, 0, 0, "1e-9"], None), (32, 32, 2.93, [1000, 0, 1, "1e-12"], None),
                  (64, 64, 1.5, [1000, 0, 1, "1e-9"], 0.5)]
    for idx, (n, m, r, args, rl) in enumerate(ir2s_cases):
        rng = np.random.default_rng(700 + idx)
        st = rng.integers(0, 4, (n, m))
        vin = list(np.round(rng.uniform(0, 1, n), 3))
        vout = list(np.round(rng.uniform(0, 0.05, m), 3))
        run_case(f"ir2s{idx}_v1_{n}x{m}", 1, n, m, st, STATES4, r, vin, vout, args,
                 solve_cmd="fastsolve", rload_left=rl)
    # General boundaries (V1 nodeBasedGeneralGS): which edges drive, and at what voltage.
    gen_cases = [
        (8, 8, "both_sides", [100000, 1.9, 0, "yes", "1e-12"]),
        (6, 10, "two_sinks", [57, 1.7, 0]),
        (12, 7, "write_scheme", [100000, 1.85, 0, "yes", "1e-11"]),
        (16, 16, "all_driven", [100000, 1.9, 0, "yes", "1e-12"]),
        (5, 4, "open_bottoms", [200, 1.5, 0]),
    ]
    for idx, (n, m, kind, args) in enumerate(gen_cases):
        rng = np.random.default_rng(900 + idx)
        st = rng.integers(0, 4, (n, m))
        on = [np.zeros(m, int), np.ones(n, int), np.ones(m, int), np.zeros(n, int)]
        v = [np.zeros(m), np.round(rng.uniform(0, 1, n), 3), np.zeros(m), np.zeros(n)]
        if kind == "both_sides":
            on[3][:] = 1
            v[3] = np.round(rng.uniform(0, 1, n), 3)
        elif kind == "two_sinks":
            on[0][:] = 1
            v[0] = np.round(rng.uniform(0, 0.05, m), 3)
        elif kind == "write_scheme":
            on[1][:] = 0
            on[1][[2, 5]] = 1
            v[1][[2, 5]] = [1.2, 0.6]
            on[3][[2, 9]] = 1
            v[3][[2, 9]] = [0.6, 0.3]
            on[2][:] = 0
            on[2][[1, 4]] = 1
            v[2][[1, 4]] = [0.0, 0.6]
            on[0][[4]] = 1
            v[0][[4]] = [0.3]
        elif kind == "all_driven":
            for e in range(4):
                on[e][:] = 1
            v = [np.round(rng.uniform(0, 0.1, m), 3), np.round(rng.uniform(0.5, 1, n), 3),
                 np.round(rng.uniform(0, 0.1, m), 3), np.round(rng.uniform(0.5, 1, n), 3)]
        elif kind == "open_bottoms":
            on[2][[1, 3]] = 0
            on[0][[1, 3]] = 1
            v[0][[1, 3]] = [0.05, 0.0]
        run_general(f"gen{idx}_v1_{n}x{m}", n, m, st, STATES4, 2.93, on, v, args)

    rng = np.random.default_rng(777)
    st = rng.integers(0, 4, (128, 128))
    vin = list(np.where(rng.random(128) < 0.5, 0.2, 0.0))
    run_case("ir2s5_v1_128x128", 1, 128, 128, st, STATES4, 3.5, vin, [0.0] * 128, [6, 0, 0, "1e-9"],
             keep_fields=False, solve_cmd="fastsolve")


if __name__ == "__main__":
    sys.exit(main())
