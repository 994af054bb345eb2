"""Build recipe of the fadesim HIP library (gfx950).

``python -m fadesim_amd.build`` compiles every source under fadesim_amd/csrc
into ``fadesim_amd/lib/libfadesim.so`` (the C ABI of include/fadesim.h) and
``fadesim_amd/lib/sim`` (the command-language CLI).  Always rebuilds.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: every fused multiply-add in the kernels is written out
# (fmaf), so CPU replays of the kernels can match them bit for bit.
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIP_FLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(REPO, "include"), "-I", CSRC,
]

LIB_SOURCES = ["common.cpp", "rbsor.hip", "gs64.hip", "gslex.hip", "hook.hip", "nn.hip", "cells.hip", "ragged.hip", "ir2s.hip"]
CLI_SOURCES = ["sim_main.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{' '.join(cmd)}\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cli = [s for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs) + len(cli))) as ex:
        objs = list(ex.map(_compile, srcs))
        cli_objs = list(ex.map(_compile, cli))
    so = os.path.join(LIB_DIR, "libfadesim.so")
    tmp = so + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stderr[-6000:]}")
    os.replace(tmp, so)
    if cli_objs:
        exe = os.path.join(LIB_DIR, "sim")
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", exe + ".tmp"] + cli_objs + [
            "-L", LIB_DIR, "-lfadesim", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI link failed:\n{' '.join(cmd)}\n{r.stderr[-6000:]}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {so}")
    return so


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
