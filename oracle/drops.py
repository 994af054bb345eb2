"""Drop-variable IR2S iteration in FP64 numpy -- the CPU statement of the
batched solver's in-kernel initial guess (and of its damped variant).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the product.

With drops p = vin_i - u_ij, q = d_ij - vout_j and uniform wire resistance r
(source and sink segments r as well), one iteration solves every wordline and
bitline exactly for the present cell currents s = g (c - p - q), c = vin - vout:
    p_ij = r sum_j' (min(j, j') + 1) s_ij'       (ir_drop_fast_solver.h:44-55)
    q_ij = r sum_i' (n - max(i, i')) s_i'j       (ir_drop_fast_solver.h:57-70)
written here as two cumulative sums per line.  Damping w blends the new drops
with the old: x <- (1 - w) x + w x_new (w = 1 is the reference iteration).
Pinned against oracle.ir2s (the statement-by-statement restatement of the
reference, itself pinned bitwise against the reference binary) and against
the exact solve (fixed point) in tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np


def wire_drops(s: np.ndarray, r: float):
    """Drops of both wire layers carrying the cell currents s [n, m]."""
    seg_p = np.cumsum(s[:, ::-1], axis=1)[:, ::-1]      # wordline segment j: currents at columns >= j
    p = r * np.cumsum(seg_p, axis=1)
    seg_q = np.cumsum(s, axis=0)                         # bitline segment below row i: rows <= i
    q = r * np.cumsum(seg_q[::-1], axis=0)[::-1]
    return p, q


def spectral_estimate(g, r: float) -> float:
    """r gbar (1/(4 sin^2(pi/(4m+2))) + 1/(4 sin^2(pi/(4n+2)))): the largest
    eigenvalue of s -> g (P + Q) s for uniform g (the wordline Green's matrix
    min(j,j')+1 and the bitline one n - max(i,i') have largest eigenvalues
    1/(4 sin^2(pi/(4m+2))) and 1/(4 sin^2(pi/(4n+2))); with uniform g the
    operator is r g (P x I + I x Q)), used with the crossbar's mean
    conductance for any g -- the bound of the Chebyshev schedule."""
    g = np.asarray(g, dtype=np.float64)
    n, m = g.shape
    sc, sr = np.sin(np.pi / (4 * m + 2)), np.sin(np.pi / (4 * n + 2))
    return float(r * g.mean() * (0.25 / sc ** 2 + 0.25 / sr ** 2))


def chebyshev_schedule(K: int, rho: float):
    """Damping of steps k = 0..K-1: 1 / mu_k at the Chebyshev nodes
    mu_k = 1 + rho (1 + cos(pi (2k+1) / (2K))) / 2 of [1, 1 + rho]."""
    return [1.0 / (1.0 + rho * 0.5 * (1.0 + np.cos(np.pi * (2 * k + 1) / (2 * K)))) for k in range(int(K))]


def ir2s_drops(g, r: float, vin, vout=None, iters: int = 1, omega: float = 1.0, p0=None, q0=None):
    """`iters` damped IR2S iterations from (p0, q0) (zero drops by default).
    omega < 0: the Chebyshev schedule over [0, |omega| spectral_estimate].
    Returns (p, q, iout) with iout_j = sum_i g_ij (c_ij - p_ij - q_ij)."""
    g = np.asarray(g, dtype=np.float64)
    n, m = g.shape
    vin = np.asarray(vin, dtype=np.float64).reshape(n)
    vout = np.zeros(m) if vout is None else np.asarray(vout, dtype=np.float64).reshape(m)
    c = vin[:, None] - vout[None, :]
    p = np.zeros((n, m)) if p0 is None else np.asarray(p0, dtype=np.float64).copy()
    q = np.zeros((n, m)) if q0 is None else np.asarray(q0, dtype=np.float64).copy()
    if omega < 0:
        steps = chebyshev_schedule(iters, -omega * spectral_estimate(g, r))
    else:
        steps = [omega] * int(iters)
    for w in steps:
        pn, qn = wire_drops(g * (c - p - q), float(r))
        p = (1.0 - w) * p + w * pn
        q = (1.0 - w) * q + w * qn
    return p, q, (g * (c - p - q)).sum(0)


def coupling_radius(g, r: float, iters: int = 80, seed: int = 0) -> float:
    """Spectral radius of s -> g (P + Q) s (the undamped iteration's
    contraction) by power iteration."""
    g = np.asarray(g, dtype=np.float64)
    e = np.random.default_rng(seed).standard_normal(g.shape)
    lam = 0.0
    for _ in range(iters):
        p, q = wire_drops(e, float(r))
        e2 = g * (p + q)
        lam = np.linalg.norm(e2) / np.linalg.norm(e)
        e = e2 / np.linalg.norm(e2)
    return float(lam)
