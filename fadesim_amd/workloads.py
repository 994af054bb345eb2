"""Seeded synthetic crossbar workloads (inputs resident on the GPU).

* ``uniform_batch``: B crossbars of one size, 4 cell states
  {2e6, 1.6e6, 1.3e6, 1e6} ohm drawn uniformly, wire resistance U[1, 5] ohm,
  wordline inputs 0 or 0.2 V with p = 0.5 (the 4096 x 128x128 benchmark
  distribution).
* ``uneven_batch``: a shuffled mix of sizes -- conductances log-uniform in
  [5e-7, 1e-6] S, wire resistance U[1, 5] ohm, inputs 0.3 k / 7 V with k
  uniform on 0..7 (8-level quantized) -- packed for RaggedPlan.
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np
import torch

STATES4 = (2e6, 1.6e6, 1.3e6, 1e6)


def uniform_batch(B: int, n: int, seed: int, device, m: int = None):
    m = n if m is None else m
    rng = np.random.default_rng(seed)
    st = rng.integers(0, 4, (B, n, m))
    g = (1.0 / np.asarray(STATES4)[st]).astype(np.float32)
    r = rng.uniform(1.0, 5.0, B)
    vin = np.where(rng.random((B, n)) < 0.5, 0.2, 0.0)
    return (torch.tensor(g, device=device), torch.tensor(r, device=device), torch.tensor(vin, device=device))


def uneven_batch(counts: Dict[int, int], seed: int, device, numpy_too: bool = False):
    """counts: {size: number of crossbars}.  Returns a dict with rows (list,
    shuffled order), packed g (float32), r (float64 [K]), vin (float64 packed)
    on `device`, and optionally the per-instance numpy arrays."""
    rng = np.random.default_rng(seed)
    sizes = []
    for n, c in sorted(counts.items()):
        sizes += [int(n)] * int(c)
    order = rng.permutation(len(sizes))
    sizes = [sizes[k] for k in order]
    gs, vins = [], []
    lo, hi = math.log(5e-7), math.log(1e-6)
    for n in sizes:
        gs.append(np.exp(rng.uniform(lo, hi, (n, n))).astype(np.float32))
        vins.append(0.3 * rng.integers(0, 8, n) / 7.0)
    r = rng.uniform(1.0, 5.0, len(sizes))
    out = {"rows": sizes,
           "g": torch.tensor(np.concatenate([x.ravel() for x in gs]) if gs else np.zeros(0, np.float32),
                             device=device),
           "r": torch.tensor(r, device=device),
           "vin": torch.tensor(np.concatenate(vins) if vins else np.zeros(0), device=device)}
    if numpy_too:
        out.update({"g_list": gs, "vin_list": vins, "r_np": r})
    return out
