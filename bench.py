#!/usr/bin/env python3
"""Throughput of the batched IR-drop solve on one MI355X.

Workload (one "step"): one batched solve of 4096 independent 128x128
crossbars -- 4 cell states {2e6, 1.6e6, 1.3e6, 1e6} ohm drawn uniformly, wire
resistance ~ U[1, 5] ohm per crossbar, wordline inputs in {0, 0.2} V with
p = 0.5, bitlines grounded -- to <= 1e-6 relative bitline-current error
(FP32 red-black SOR sweeps, FP64 residual test / refinement / readout).
Inputs are synthetic, seeded, and resident in HBM before the timed region.

Reported: whole-job solves/s of three initial guesses for the same solve --
zero drops (plain SOR), K wire-solve (IR2S) iterations inside the solve
kernel (K from the autotuned table), and, when trained weights are present
(fadesim_amd/weights), the learned warm start with the network's inference
inside the timed step plus one IR2S iteration on its field (`value`, the
metric's NN-warm-start figure, the network running inside the solve's
guess kernel; "nn_3k" the same start from separate features and network
kernels, bit for bit the same outputs; "nn_unet16" the same with the 4-level
full-resolution network); mean sweeps of each; the roofline of the
step's dominant kernel (the guess-first kernel when every crossbar is
certified at its first test, else the sweeping kernel) from HIP events on
its launch stream, with the HBM / LDS / VALU figures of the rocprofv3
counter runs summarised in profiles/, and plain SOR against the FP32 VALU
peak (its binding resource); the CPU baseline (the
oracle's restatement of the reference's lexicographic solver, OpenMP over
instances) on a bounded sample, and the measured error of every variant on
that sample against the exact FP64 oracle (checker of the cpu_baseline leg).

Extra objects in the same line (not part of `value`): "uneven" -- 1024
crossbars each of 64x64, 128x128, 256x256 and 512x512 (shuffled, seed 33)
solved in one persistent launch and as a captured graph replay; "mlp" -- the
784-256-10 crossbar MLP on 1024 images (sinh cells, every tile solve
batched), plain vs IR2S- and NN-warm-started tile solves; "hook" -- the
PimTorch hook drop-in with the reference's own iteration (256 FP32 tiles of
128x128, 600 sweeps), beside the oracle's restatement of that iteration on
one core (bit for bit the same results).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
                [--no-uneven] [--no-mlp] [--no-hook] [--mlp-images N] [--only plain,ir2s,nn]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
STATES4 = np.array([2e6, 1.6e6, 1.3e6, 1e6])
M = N = 128


def make_batch(B, seed):
    rng = np.random.default_rng(seed)
    st = rng.integers(0, 4, (B, M, N))
    g = (1.0 / STATES4[st]).astype(np.float32)
    r = rng.uniform(1.0, 5.0, B)
    vin = np.where(rng.random((B, M)) < 0.5, 0.2, 0.0)
    return g, r, vin


def algorithmic_bytes(passes_sum, B, warm=False, coarse=False):
    """Streaming-schedule bytes: 4 B of conductance per cell per pass over the
    crossbar (a red-black sweep or a wire-solve iteration), plus the inputs and
    the FP64 readout (4*M + 8*N per crossbar), plus the FP32 warm-start field
    (2*M*N*4) when the solve reads one from HBM, plus the network's coarse
    output ((M/4)*(N/4) pixels x 16 BF16 channels) when it is combined at
    load."""
    per = 4.0 * M + 8.0 * N + (8.0 * M * N if warm else 0.0) + (2.0 * 16 * (M // 4) * (N // 4) if coarse else 0.0)
    return 4.0 * M * N * passes_sum + B * per


def load_counters(variant, kernel):
    """rocprofv3 counter figures of `kernel` in the run of `variant`, as
    summarised in profiles/ (HBM bytes per launch from FETCH_SIZE /
    WRITE_SIZE scaled by the same counters' reading of a plain device copy;
    LDS bytes and VALU / wait fractions from the SQ counters), or {}."""
    p = os.path.join(REPO, "profiles", "solve_128x128_b4096_counters.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        t = json.load(f)
    return t.get("variants", {}).get(variant, {}).get(kernel, {})


# Per-cell FP32 operations of one red-black sweep (both nodes of a cell):
# tt = fma(h, p + q, rho) and x' = fma(w/2, n1 + n2, fma(1 - w, x, tt)),
# 8 flops per node (an FMA counted as 2).
SWEEP_FLOPS_PER_CELL = 16
# MI355X vector FP32 peak (packed FP32 included; this build emits no packed
# FP32 instructions, so half of it is the reachable rate) and LDS bandwidth
# (128 bytes per clock per CU, 256 CUs, 2.4 GHz).
VALU_FP32_PEAK_TFLOPS = 157.3
LDS_PEAK_GBS = 128 * 256 * 2.4
# VALU issue rate: each SIMD issues one wave64 VALU instruction every 4
# cycles (16 lanes per cycle), 4 SIMDs per CU, 256 CUs, 2.4 GHz -- in lane
# operations per second (FP64 and transcendental instructions take longer,
# so the reachable rate of FP64-heavy code is lower)
VALU_LANE_OPS_PEAK = 64 * 256 * 2.4e9


def valu_roofline(ctr, kernel_ms):
    """Achieved VALU lane operations per second of a kernel (SQ_INSTS_VALU of
    the counter runs x 64 lanes, over the kernel time measured here) against
    the issue peak, or None without counters."""
    if "sq_insts_valu" not in ctr:
        return None
    ach = ctr["sq_insts_valu"] * 64 / (kernel_ms * 1e-3)
    return {"bound": "valu", "achieved": ach / 1e12, "peak": VALU_LANE_OPS_PEAK / 1e12,
            "unit": "T lane-ops/s", "frac": ach / VALU_LANE_OPS_PEAK,
            "model": "SQ_INSTS_VALU x 64 lanes per launch (rocprofv3 counters in profiles/) / kernel time; "
                     "peak: one wave64 VALU instruction per 4 cycles per SIMD, 4 SIMDs x 256 CUs x 2.4 GHz"}


# Time of the restatement (oracle_nodegs, one core) over the reference V2 `sim`
# binary's own "Duration time" for the same 128x128 solve (config 1, 1507
# sweeps): 0.679 s / 0.577 s, both measured in the build container (Intel
# Xeon, gcc -O2 / the reference's -O3 CMake build).
RESTATEMENT_OVER_REFERENCE_TIME = 1.18


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, r, vin):
    """The oracle's restatement of the reference's lexicographic solver
    (rule of the reference: break at max|dv| < 1e-9, w = 1.97 -- the first
    decade that meets 1e-6) on a bounded sample of the timed batch, OpenMP
    over instances; and the exact FP64 oracle answers of that sample, the
    checker of every GPU variant's error."""
    import oracle
    B = g.shape[0]
    t0 = time.perf_counter()
    iout, sweeps, threads = oracle.nodegs_batch(g, r, vin, 100000, 1.97, 1e-9)
    dt = time.perf_counter() - t0
    iex = oracle.exact_batch(g, r, vin)
    err = float(np.max(np.abs(iout - iex) / np.abs(iex)))
    value = B / dt
    return {"value": value, "unit": "solves/s", "cores": int(threads), "kind": "port", "cpu_model": cpu_model(),
            "restatement_over_reference_time": RESTATEMENT_OVER_REFERENCE_TIME,
            "reference_equivalent_value": value * RESTATEMENT_OVER_REFERENCE_TIME,
            "sample": f"the first {B} instances of the timed batch, lexicographic SOR w=1.97 break "
                      f"max|dv|<1e-9, mean {sweeps.mean():.0f} sweeps, {dt:.2f} s wall on {threads} threads, "
                      f"max rel I_out err vs exact FP64 {err:.2e}"}, iex


def bench_uneven(args, dev):
    """1024 crossbars of each of 64, 128, 256, 512 (the uneven-size workload)."""
    import torch
    from fadesim_amd.ragged import RaggedPlan
    from fadesim_amd.workloads import uneven_batch
    counts = {64: 1024, 128: 1024, 256: 1024, 512: 1024}
    inst = uneven_batch(counts, seed=33, device=dev)
    plan = RaggedPlan(inst["rows"], device=dev)
    res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=args.tol)
    if int((res.status != 0).sum()) != 0:
        raise RuntimeError("uneven batch: not every crossbar converged")
    steps = max(1, min(args.steps, 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    for _ in range(steps):
        plan.solve(inst["g"], inst["r"], inst["vin"], tol=args.tol, timing=True, out=res)
        kms.append(plan.last_ms())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    graph, gres = plan.capture(inst["g"], inst["r"], inst["vin"], tol=args.tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    dtg = (time.perf_counter() - t0) / steps
    if not torch.equal(gres.iout, res.iout):
        raise RuntimeError("uneven batch: graph replay differs from the eager solve")
    sizes = torch.tensor(inst["rows"], device=dev)
    sw = res.sweeps.double()
    K = len(inst["rows"])
    return {"workload": "1024 each of 64x64, 128x128, 256x256, 512x512 crossbars, shuffled; g log-uniform "
                        "[5e-7, 1e-6] S, r ~ U[1,5] ohm, vin = 0.3 k/7 V, one persistent launch",
            "crossbars": K, "solves_per_s": K / dt, "ms_per_step": 1e3 * dt, "kernel_ms": float(np.mean(kms)),
            "graph_replay_ms": 1e3 * dtg, "graph_solves_per_s": K / dtg,
            "mean_sweeps": {str(n): float(sw[sizes == n].mean()) for n in (64, 128, 256, 512)},
            "plan": plan.stats, "tuned": bool(plan.tuning)}


def bench_hook(args, dev):
    """The PimTorch hook drop-in (ir_drop_gs) with the reference's own
    iteration: 256 FP32 tiles of 128x128 (16 inputs x 4 x 4 tiles), 600
    sweeps at beta 1.90 (PimTorch's defaults), against the oracle's C
    restatement of the same iteration on one core (bit for bit the same)."""
    import torch
    from fadesim_amd import hook
    rng = np.random.default_rng(7)
    bsize, numX, numY, n, sweeps = 16, 4, 4, 128, 600
    G = (1.0 / rng.choice(STATES4, (numX, n, numY * n))).astype(np.float32)
    x = np.where(rng.random((bsize, numX, n)) < 0.5, 0.2, 0.0).astype(np.float32)
    tG, tx = torch.tensor(G, device=dev), torch.tensor(x, device=dev)
    call = lambda: hook.ir_drop_gs(tx, tG, bsize, n, n, numX, numY, sweeps, 1 / 2.93, 1.9, False, 0.0, False, [])
    out = call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        out = call()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    tiles = bsize * numX * numY
    res = {"workload": f"{tiles} FP32 tiles of 128x128 (PimTorch hook ir_drop_gs), {sweeps} sweeps, beta 1.90",
           "ms_per_call": 1e3 * dt, "tiles_per_s": tiles / dt}
    if not args.no_cpu:
        # CPU baseline and checker: the oracle's restatement on one input
        # vector's 4 x 4 tiles
        import oracle
        t1 = time.perf_counter()
        ref, _ = oracle.hook_gs(x[:1], G, 1, n, n, numX, numY, sweeps, 1 / 2.93, 1.9, False, 0.0, False, [])
        dc = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": numX * numY / dc, "unit": "tiles/s", "cores": 1, "kind": "port",
                               "sample": f"{numX * numY} tiles (one input vector), the same iteration"}
        res["bitwise_equal_to_cpu"] = bool(np.array_equal(out[:1].cpu().numpy(), ref))
    return res


def bench_mlp(args, dev, net, k_ir2s, w_ir2s):
    """784-256-10 crossbar MLP (8-bit, 2 bits/cell, 128x128 PN tiles, sinh
    cells a = 2.5/V): tile solves plain, IR2S-warm-started and NN-warm-started."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(5)
    imgs = torch.rand(args.mlp_images, 784, generator=gen).to(dev)
    out = {"workload": f"784-256-10 MLP, {args.mlp_images} images U[0,1], 8-bit inputs bit-serial 0/0.2 V, "
                       "8-bit PN weights 2 bits/cell on 128x128 tiles, sinh cells a=2.5/V, tiles to 1e-6"}
    res = {}
    for name in ("plain", "ir2s", "nn") if net is not None else ("plain", "ir2s"):
        P = pim.mlp_inference_params(warm_start=net if name == "nn" else None, tol=args.tol,
                                     ir2s_iters={"plain": 0, "ir2s": k_ir2s, "nn": 1}[name],
                                     ir2s_omega=w_ir2s if name == "ir2s" else 1.0)
        model = pim.MLP(params=P, seed=0).to(dev).eval()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            logits = model(imgs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = dict(P.stats)
        res[name] = logits
        out[name] = {"seconds": dt, "images_per_s": args.mlp_images / dt, "tile_solves": st.get("solves"),
                     "tile_solves_per_s": st.get("solves", 0) / dt, "sweeps": st.get("sweeps"),
                     "max_picard_steps": st.get("picard_steps")}
    P0 = pim.PhyArrParams(Ron=1e6, Roff=2e6, cellBits=2, arrRowSize=128, arrColSize=128, inputMaxVoltage=0.2,
                          ir_mode="none")
    with torch.no_grad():
        ideal = pim.MLP(params=P0, seed=0).to(dev).eval()(imgs)
    for name in ("ir2s", "nn"):
        if name in res:
            d = (res[name] - res["plain"]).abs()
            out[f"{name}_vs_plain_logits_max_abs_diff"] = float(d.max())
            out[f"{name}_vs_plain_argmax_agreement"] = float(
                (res[name].argmax(1) == res["plain"].argmax(1)).double().mean())
    out["ir_drop_effect_vs_ideal_argmax_agreement"] = float(
        (res["plain"].argmax(1) == ideal.argmax(1)).double().mean())
    return out


def ir2s_schedule_128(dev):
    """IR2S iterations and their damping for 128x128 from the autotuned table
    (4 undamped without one; a negative damping is the Chebyshev schedule)."""
    from fadesim_amd import autotune
    t = (autotune.load_table(dev) or {}).get("128", {})
    return int(t.get("ir2s_iters", 4)) or 4, float(t.get("ir2s_omega", 1.0)) or 1.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-hook", action="store_true")
    ap.add_argument("--mlp-images", type=int, default=1024)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--only", default="", help="comma list of plain,ir2s,nn,nn_3k,nn_unet16: run only these solve variants")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solves shard "
              "without exchange across GPUs; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n}), flush=True)
        return
    from fadesim_amd.solver import last_guess_ms, last_kernel_ms, solve_batch
    from fadesim_amd import warmstart

    dev = torch.device("cuda:0")
    B = args.batch
    g, r, vin = make_batch(B, seed=2024)
    tg = torch.tensor(g, device=dev)
    tr = torch.tensor(r, device=dev)
    tv = torch.tensor(vin, device=dev)
    net = warmstart.load_default(dev)
    K, W = ir2s_schedule_128(dev)
    only = set(x for x in args.only.split(",") if x)

    # name: (IR2S iterations of the solve, passes over the conductance per
    # crossbar before the sweeps, warm field read from HBM, coarse output
    # read from HBM, step)
    variants = {
        # zero-drop initial guess, red-black SOR sweeps
        "plain": (0, 0, False, False, lambda timing=False: solve_batch(tg, tr, tv, tol=args.tol, timing=timing)),
        # K wire-solve (IR2S) iterations on the zero guess inside the solve kernel
        "ir2s": (K, K, False, False, lambda timing=False: solve_batch(tg, tr, tv, tol=args.tol, ir2s_iters=K,
                                                                    corr_iters=K, ir2s_omega=W, timing=timing)),
    }
    if net is not None:
        # the network's start formed inside the solve's guess kernel (pooled
        # features and the base field = two wire-solve passes from zero, the
        # network with its activations in LDS, the combine), one IR2S
        # iteration on it, then the same test / SOR
        fused = net.solves_in_kernel
        if fused:
            variants["nn"] = (1, 3, False, False, lambda timing=False: solve_batch(
                tg, tr, tv, net=net, tol=args.tol, ir2s_iters=1, corr_iters=K, timing=timing))
        # the same start from separate kernels: pooled features, network, and
        # the network output combined while the solve loads it with the base
        # field, which the solve forms itself (bit for bit the same outputs)
        variants["nn_3k" if fused else "nn"] = (1, 3, False, True, lambda timing=False: solve_batch(
            tg, tr, tv, **net.warm_kwargs(tg, tr, tv), tol=args.tol, ir2s_iters=1, corr_iters=K, test_first=True,
            timing=timing))
    # the 4-level full-resolution encoder-decoder (~340k parameters), for comparison
    net16 = None
    if os.path.exists(os.path.join(warmstart.WEIGHTS_DIR, "warmstart_128_unet16.bin")):
        net16 = warmstart.WarmStartNet.load(dev, name="warmstart_128_unet16")
        variants["nn_unet16"] = (1, 1, True, False, lambda timing=False: solve_batch(
            tg, tr, tv, warm=net16.predict(tg, tr, tv), tol=args.tol, ir2s_iters=1, corr_iters=K, test_first=True,
            timing=timing))
    if only:
        variants = {k: v for k, v in variants.items() if k in only}

    def run(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms, gms, sweeps, res = [], [], [], None
        for _ in range(args.steps):
            res = step(timing=True)
            kms.append(last_kernel_ms())
            gms.append(last_guess_ms())
            sweeps.append(res.sweeps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt, kms, gms, torch.stack(sweeps).to(torch.float64), res

    meas, last = {}, {}
    spec = {}
    for name, (k_ir2s, pre_passes, warm_rd, coarse_rd, step) in variants.items():
        spec[name] = (pre_passes, warm_rd, coarse_rd)
        dt, kms, gms, sw, res = run(step)
        last[name] = res
        if int((res.status != 0).sum()) != 0:
            raise RuntimeError(f"{name}: not every crossbar converged")
        if res.rejected is not None and int(res.rejected.sum()) != 0:
            raise RuntimeError(f"{name}: the warm start was rejected for {int(res.rejected.sum())} crossbars")
        kernel_ms = float(np.mean(kms))
        guess_ms = float(np.mean(gms)) if min(gms) > 0 else None
        passes = float(sw.sum()) / args.steps + pre_passes * B
        ach = algorithmic_bytes(passes, B, warm=warm_rd, coarse=coarse_rd) / (kernel_ms * 1e-3) / 1e9
        meas[name] = {"solves_per_s": B * args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                      "mean_sweeps": float(sw.mean()), "ir2s_iters": k_ir2s,
                      "ir2s_omega": W if name == "ir2s" else 1.0, "kernel_ms": kernel_ms,
                      "guess_kernel_ms": guess_ms, "kernel_GBps_algorithmic": ach}
    # the NN-warm step reported as `value`: the faster of its two forms (the
    # network inside the guess kernel, or separate features and network
    # kernels; the same outputs bit for bit)
    nn_forms = [k for k in ("nn", "nn_3k") if k in meas]
    nn_head = max(nn_forms, key=lambda k: meas[k]["solves_per_s"]) if nn_forms else None
    head = nn_head or next(k for k in ("ir2s", "plain", "nn_unet16") if k in meas)
    hm = meas[head]
    out = {"metric": "IR-drop solves/s at 1e-6 rel I_out err: NN-warm-start vs plain SOR, 1 GPU",
           "value": hm["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": hm["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{B} x 128x128 crossbars, 4 states {{2e6,1.6e6,1.3e6,1e6}} ohm, "
                                  "r ~ U[1,5] ohm, vin in {0,0.2} V, to 1e-6 rel I_out err",
                      "batch": B, "rows": M, "cols": N, "tol": args.tol,
                      "initial_guess": {"nn": "nn-warm-start (+1 IR2S iteration)", "ir2s": f"{K} IR2S iterations",
                                        "plain": "zero-drop",
                                        "nn_3k": "nn-warm-start from separate kernels (+1 IR2S iteration)",
                                        "nn_unet16": "unet16 nn-warm-start (+1 IR2S iteration)"}[head]},
           "plain": meas.get("plain"), "ir2s_warm": meas.get("ir2s"),
           "warm_start": dict(meas["nn"], network=net.describe()) if "nn" in meas else None,
           "warm_start_three_kernels": meas.get("nn_3k"),
           "warm_start_unet16": dict(meas["nn_unet16"], network=net16.describe()) if "nn_unet16" in meas else None}
    if "plain" in meas and nn_head:
        nm = meas[nn_head]
        ps, ns = meas["plain"]["mean_sweeps"], nm["mean_sweeps"]
        # SOR sweeps per solve, plain vs NN-warm (the NN-warm solve also runs one
        # wire-solve pass); the ratio is null when the NN-warm solve needs none
        out["sweeps_per_solve"] = {"plain": ps, "nn": ns}
        out["sweep_reduction"] = ps / ns if ns > 0 else None
        # passes over the crossbar: SOR sweeps, plus the IR2S iteration the
        # NN-warm solve runs (each costs about three sweeps of kernel time)
        out["pass_reduction"] = ps / (ns + nm["ir2s_iters"])
        out["kernel_time_ratio"] = meas["plain"]["kernel_ms"] / nm["kernel_ms"]
        out["nn_vs_plain_speedup"] = nm["solves_per_s"] / meas["plain"]["solves_per_s"]
    if "plain" in meas and "ir2s" in meas:
        out["ir2s_vs_plain_speedup"] = meas["ir2s"]["solves_per_s"] / meas["plain"]["solves_per_s"]
    # the dominant kernel of the reported step: guess_kernel when the step's
    # solves end at their first test (every crossbar of the nn and ir2s
    # variants), timed alone between HIP events on its stream
    gk = hm["guess_kernel_ms"] is not None and hm["mean_sweeps"] == 0
    gname = "guess_net_kernel" if head == "nn" and net is not None and net.solves_in_kernel else "guess_kernel"
    kname = (gname + "<Geo<4,8,32,16>>") if gk else "rbsor_kernel<Geo<4,8,32,16>>"
    kms_dom = hm["guess_kernel_ms"] if gk else hm["kernel_ms"]
    pre_passes, warm_rd, coarse_rd = spec[head]
    passes_dom = pre_passes * B + hm["mean_sweeps"] * B
    ach_dom = algorithmic_bytes(passes_dom, B, warm=warm_rd, coarse=coarse_rd) / (kms_dom * 1e-3) / 1e9
    ctr = load_counters(head, gname if gk else "rbsor_kernel")
    out["roofline"] = {"bound": "hbm", "achieved": ach_dom, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": ach_dom / HBM_PEAK_GBS,
                       "traffic": ctr.get("hbm_bytes_per_launch"),
                       "kernel": kname, "kernel_ms": kms_dom, "step_kernels_ms": hm["kernel_ms"],
                       "bytes_model": "(sweeps + passes before them) * 4*M*N + 4*M + 8*N per crossbar "
                                      "(conductance streamed once per sweep or wire-solve pass), + 8*M*N for "
                                      "a warm-start field read from HBM, + 2*16*(M/4)*(N/4) for a coarse "
                                      "network output read from HBM",
                       "lds": ({"achieved": ctr["lds_bytes_per_launch"] / (kms_dom * 1e-3) / 1e9,
                                "peak": LDS_PEAK_GBS, "unit": "GB/s",
                                "frac": ctr["lds_bytes_per_launch"] / (kms_dom * 1e-3) / 1e9 / LDS_PEAK_GBS}
                               if "lds_bytes_per_launch" in ctr else None),
                       "valu_busy": ctr.get("valu_busy"), "wait_frac": ctr.get("wait_frac"),
                       "counters": "profiles/solve_128x128_b4096_counters.md" if ctr else None}
    # the resource that binds the dominant kernel: VALU issue (the wire-solve
    # iterations in FP32, the test in FP64)
    out["roofline_valu"] = valu_roofline(ctr, kms_dom)
    if "plain" in meas:
        pm = meas["plain"]
        tflops = SWEEP_FLOPS_PER_CELL * M * N * pm["mean_sweeps"] * B / (pm["kernel_ms"] * 1e-3) / 1e12
        pc = load_counters("plain", "rbsor_kernel")
        # plain SOR keeps the conductance in registers across its sweeps: its
        # binding resource is VALU issue, not HBM
        out["roofline_plain_sor"] = {"bound": "valu", "achieved": tflops, "peak": VALU_FP32_PEAK_TFLOPS,
                                     "unit": "TFLOP/s", "frac": tflops / VALU_FP32_PEAK_TFLOPS,
                                     "flops_model": f"{SWEEP_FLOPS_PER_CELL} FP32 flops per cell per sweep",
                                     "kernel_ms": pm["kernel_ms"], "traffic": pc.get("hbm_bytes_per_launch"),
                                     "valu_busy": pc.get("valu_busy"),
                                     "valu_issue": valu_roofline(pc, pm["kernel_ms"]),
                                     "hbm_GBps_measured": (pc["hbm_bytes_per_launch"] / (pm["kernel_ms"] * 1e-3) / 1e9
                                                           if "hbm_bytes_per_launch" in pc else None)}
    if not args.no_cpu:
        S = min(args.cpu_sample, B)
        cpu, iex = cpu_baseline(g[:S], r[:S], vin[:S])
        out["cpu_baseline"] = cpu
        # error of every variant against the exact FP64 oracle: the output of
        # its last timed step (the very call that was timed) on the sample
        out["max_rel_err"] = {name: float(np.max(np.abs(res.iout[:S].cpu().numpy() - iex) / np.abs(iex)))
                              for name, res in last.items()}
        out["max_rel_err_sample"] = f"the first {S} instances of the timed batch, last timed step"
    else:
        out["cpu_baseline"] = None
    for key, flag, fn in (("uneven", args.no_uneven, lambda: bench_uneven(args, dev)),
                          ("mlp", args.no_mlp, lambda: bench_mlp(args, dev, net, K, W)),
                          ("hook", args.no_hook, lambda: bench_hook(args, dev))):
        if flag:
            continue
        try:
            out[key] = fn()
        except Exception as e:  # an extra workload never hides the headline line
            out[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
