"""Autotuner of the mixed-size solver's per-size schedule.

For each crossbar size (64, 128, 256, 512) it times the persistent solve of a
sample of the uneven-size workload (fadesim_amd.workloads.uneven_batch) on the
GPU over a grid of candidates and keeps the fastest candidate whose solves
all reach the tolerance: first the number of wire-solve (IR2S) iterations
applied to the initial guess (the same count on the refinement correction)
and their damping -- the contraction per iteration grows with r g n^2, so
small crossbars want few and 512x512 many, damped; then, with that guess,
the correction's own count, which only has to take the refined error from
the switch point to the tolerance (a negative damping is the Chebyshev
schedule over that multiple of the estimated spectral radius, computed per
crossbar in the kernel); then the relaxation factor
w(k) = 2 / (1 + k sin(pi / (2n + 1))) and the bounds of the interval between
convergence tests.  The table is stored per GPU architecture under
fadesim_amd/tuning/<arch>.json and read by RaggedPlan.

python -m fadesim_amd.autotune [--per-size 16,32,64,256] [--tol 1e-6] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time
from typing import Optional

import torch

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
KAPPAS = (0.55, 0.7, 0.83, 1.0, 1.25)
CHECKS = ((32, 16, 256), (64, 32, 256), (32, 32, 512))
IR2S = tuple(range(17))                  # iterations on the initial guess (and the correction's)
IR2S_OMEGA = (1.0, 0.85, 0.75, 0.7, 0.6,  # their damping: 2/(2+rho) for coupling radius rho
              -1.0, -1.15, -1.3)          # negative: Chebyshev schedule over |w| x the estimated radius
CORR = (1, 2, 3, 4, 5, 6, 8, 10, 12)      # iterations on the refinement correction


def arch_name(device) -> str:
    props = torch.cuda.get_device_properties(device)
    return getattr(props, "gcnArchName", "unknown").split(":")[0]


def table_path(device) -> str:
    return os.path.join(TUNING_DIR, arch_name(device) + ".json")


def load_table(device) -> Optional[dict]:
    """The tuned per-size controls for this device's architecture, or None."""
    try:
        p = table_path(device)
    except Exception:
        return None
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("classes")


def omega_of(n: int, kappa: float) -> float:
    return 2.0 / (1.0 + kappa * math.sin(math.pi / (2 * n + 1)))


def tune_size(n: int, count: int, tol: float, dev, log=print, reps: int = 2) -> dict:
    from .ragged import RaggedPlan
    from .workloads import uneven_batch
    inst = uneven_batch({n: count}, seed=4242 + n, device=dev)

    def trial(kappa, entry):
        plan = RaggedPlan(inst["rows"], device=dev, tuning={str(n): entry})
        res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=tol)  # warm-up, and correctness
        ok = int((res.status != 0).sum()) == 0
        times = []
        for _ in range(reps):
            plan.solve(inst["g"], inst["r"], inst["vin"], tol=tol, timing=True)
            times.append(plan.last_ms())
        ms = min(times)
        sw = float(res.sweeps.double().mean())
        log(f"  n={n} kappa={kappa:.2f} w={entry['omega']:.5f} checks={entry['check_first']}/"
            f"{entry['check_min']}/{entry['check_max']} ir2s={entry['ir2s_iters']}/{entry['corr_iters']}"
            f" w_ir2s={entry['ir2s_omega']:.2f}: "
            f"{ms:.3f} ms, mean sweeps {sw:.0f}{'' if ok else ' (not converged)'}")
        return (ms if ok else math.inf, kappa, entry, sw)

    def make(kappa, checks, k, w, kc=None):
        cf, cmin, cmax = checks
        return {"omega": omega_of(n, kappa), "check_first": cf, "check_min": cmin, "check_max": cmax,
                "ir2s_iters": k, "corr_iters": max(2, k) if kc is None else kc, "ir2s_omega": w}

    stage1 = [trial(1.0, make(1.0, CHECKS[0], k, w)) for k in IR2S for w in IR2S_OMEGA if k > 0 or w == 1.0]
    e1 = min(stage1, key=lambda x: x[0])[2]
    k_best, w_best = e1["ir2s_iters"], e1["ir2s_omega"]
    results = list(stage1)
    # the correction's count on its own, with the guess's count and one step
    # either side
    if k_best > 0:
        for k in sorted({max(2, k_best - 2), k_best, k_best + 2}):
            for kc in CORR:
                if kc < k_best + 2 and not (k == k_best and kc == max(2, k_best)):
                    results.append(trial(1.0, make(1.0, CHECKS[0], k, w_best, kc)))
        e1 = min(results, key=lambda x: x[0])[2]
        k_best, kc_best = e1["ir2s_iters"], e1["corr_iters"]
    else:
        kc_best = None
    for kappa in KAPPAS:
        for checks in CHECKS:
            if kappa == 1.0 and checks == CHECKS[0]:
                continue
            results.append(trial(kappa, make(kappa, checks, k_best, w_best, kc_best)))
    best = min(results, key=lambda x: x[0])
    out = dict(best[2])
    out.update({"kappa": best[1], "ms": best[0], "mean_sweeps": best[3], "sample": count})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-size", default="256,128,64,64",
                    help="instances per size 64,128,256,512 in the timing sample")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    counts = [int(x) for x in args.per_size.split(",")]
    log = lambda s: print(s, flush=True)
    t0 = time.time()
    classes = {}
    for n, c in zip((64, 128, 256, 512), counts):
        log(f"size {n}: {c} instances")
        classes[str(n)] = tune_size(n, c, args.tol, dev, log)
        log(f"size {n}: best {classes[str(n)]}")
    table = {"arch": arch_name(dev), "device": torch.cuda.get_device_name(dev), "tol": args.tol,
             "seconds": time.time() - t0, "candidates": {"kappa": KAPPAS, "checks": CHECKS, "ir2s_iters": IR2S, "ir2s_omega": IR2S_OMEGA,
                            "corr_iters": CORR}, "classes": classes}
    out = args.out or table_path(dev)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    log("wrote " + out)


if __name__ == "__main__":
    main()
