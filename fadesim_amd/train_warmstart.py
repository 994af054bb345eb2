"""Train and evaluate the learned warm start on the GPU.

python -m fadesim_amd.train_warmstart [--n 50000] [--epochs 8] [--batch 64]
    [--out DIR] [--eval 4096] [--max-seconds 600]

Data: seeded instances of the 4096 x 128x128 benchmark distribution (4 cell
states, r ~ U[1,5] ohm, vin in {0, 0.2} V), seeds [1e6, 1e6 + n); targets are
the drop fields of fadesim_solve_batch converged to 1e-8.  Evaluation on
held-out seeds [1e7, 1e7 + eval): mean sweeps to 1e-6 and solve throughput
with the zero-drop guess, the second-order physics field alone, the network,
the network followed by one wire-solve (IR2S) iteration, and four IR2S
iterations alone -- printed and written as JSON next to the weights.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import torch

from . import warmstart
from .solver import solve_batch

TRAIN_SEED0 = 1_000_000
EVAL_SEED0 = 10_000_000


def evaluate(net, n_eval, dev, log=print):
    g, r, vin = warmstart.make_instances(n_eval, EVAL_SEED0, dev)
    res = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    plain, tp = timed(lambda: solve_batch(g, r, vin, tol=1e-6))
    res["plain"] = {"mean_sweeps": plain.sweeps.double().mean().item(), "seconds": tp,
                    "solves_per_s": n_eval / tp}

    def physics_only():
        feat, base, scale = net.features(g, r, vin)
        return solve_batch(g, r, vin, warm=base, tol=1e-6)
    phys, tph = timed(physics_only)
    res["physics_2nd_order"] = {"mean_sweeps": phys.sweeps.double().mean().item(), "seconds": tph,
                                "solves_per_s": n_eval / tph}
    nn, tn = timed(lambda: solve_batch(g, r, vin, warm=net.predict(g, r, vin), tol=1e-6))
    res["nn"] = {"mean_sweeps": nn.sweeps.double().mean().item(), "seconds": tn, "solves_per_s": n_eval / tn}
    # the network's field refined by one wire-solve (IR2S) iteration and tested before any sweep
    nn1, tn1 = timed(lambda: solve_batch(g, r, vin, warm=net.predict(g, r, vin), tol=1e-6, ir2s_iters=1,
                                         corr_iters=4, test_first=True))
    res["nn_ir2s1"] = {"mean_sweeps": nn1.sweeps.double().mean().item(), "seconds": tn1,
                       "solves_per_s": n_eval / tn1}
    ir4, ti4 = timed(lambda: solve_batch(g, r, vin, tol=1e-6, ir2s_iters=4, corr_iters=4))
    res["ir2s4"] = {"mean_sweeps": ir4.sweeps.double().mean().item(), "seconds": ti4, "solves_per_s": n_eval / ti4}
    w = net.predict(g, r, vin)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        net.predict(g, r, vin)
    torch.cuda.synchronize()
    res["nn_inference_seconds"] = (time.perf_counter() - t0) / 3
    for k in ("plain", "physics_2nd_order", "nn", "nn_ir2s1", "ir2s4"):
        log(f"eval {k}: mean sweeps {res[k]['mean_sweeps']:.1f}, {res[k]['solves_per_s']:.0f} solves/s")
    log(f"eval nn inference: {1e3 * res['nn_inference_seconds']:.2f} ms for {n_eval}")
    res["n_eval"] = n_eval
    res["statuses_ok"] = bool(int(sum((x.status != 0).sum() for x in (plain, phys, nn, nn1, ir4))) == 0)
    del w
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=50000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=warmstart.WEIGHTS_DIR)
    ap.add_argument("--eval", type=int, default=4096)
    ap.add_argument("--max-seconds", type=float, default=600)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--arch", default="coarse16f", choices=sorted(warmstart.ARCHITECTURES))
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    log = lambda s: print(s, flush=True)
    net = warmstart.WarmStartNet(dev, seed=args.seed, arch=args.arch)
    log(f"network: {net.n_params()} parameters")
    t0 = time.time()
    data = warmstart.generate_dataset(net, args.n, TRAIN_SEED0)
    torch.cuda.synchronize()
    log(f"dataset: {args.n} solver solutions in {time.time() - t0:.1f}s; target rms "
        f"{data['target'].pow(2).mean().sqrt().item():.3e}")
    net.train(data, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log, seed=args.seed,
              max_seconds=args.max_seconds)
    net.info.update({"train_instances": args.n, "train_seeds": [TRAIN_SEED0, TRAIN_SEED0 + args.n],
                     "target_tol": 1e-8})
    del data
    torch.cuda.empty_cache()
    res = evaluate(net, args.eval, dev, log)
    net.info["eval"] = {k: v for k, v in res.items()}
    net.save(directory=args.out)
    with open(os.path.join(args.out, warmstart.DEFAULT_NAME + "_eval.json"), "w") as f:
        json.dump(res, f, indent=1)
    log("saved " + os.path.join(args.out, warmstart.DEFAULT_NAME + ".bin"))


if __name__ == "__main__":
    sys.exit(main())
