"""Exact FP64 solve of large crossbars by block Gauss-Seidel over the two
wire layers (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the GPU solvers on crossbars too
large for the banded Cholesky of oracle.exact (its cost grows as n m^3).

Same network as oracle.exact (Appendix-A KCL of crossbarHspice/V2/
concrete_command.h:1226-1252 with uniform wire resistance r), in drop
variables p = vin_i - u_ij, q = d_ij - vout_j, c = vin_i - vout_j:
  wordline row i:  (g + a_j + a_j+1) p_ij - a p_i,j-1 - a p_i,j+1 = g (c - q)   (p_i,-1 = 0, open right end)
  bitline col j:   (g + b_i + b_i+1) q_ij - b q_i-1,j - b q_i+1,j = g (c - p)   (open top, q_n,j = 0)
Each half step solves one layer exactly (one tridiagonal system per row or
column, Thomas algorithm vectorised across lines) with the other layer
frozen: block Gauss-Seidel on an SPD 2x2 block system, which converges for
any such system; here the rate is about (r g n^2)^2 per step, so a few
dozen steps reach FP64 round-off.  The readout is I_j = sum_i g (c - p - q).
Pinned against oracle.exact on small arrays (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def _thomas(diag: np.ndarray, off: float, rhs: np.ndarray) -> np.ndarray:
    """Solve tridiagonal systems along the last axis: diag[..., k] x_k +
    off (x_k-1 + x_k+1) = rhs[..., k] (off constant)."""
    n = diag.shape[-1]
    cp = np.empty_like(diag)
    dp = np.empty_like(rhs)
    cp[..., 0] = off / diag[..., 0]
    dp[..., 0] = rhs[..., 0] / diag[..., 0]
    for k in range(1, n):
        den = diag[..., k] - off * cp[..., k - 1]
        cp[..., k] = off / den
        dp[..., k] = (rhs[..., k] - off * dp[..., k - 1]) / den
    x = np.empty_like(rhs)
    x[..., n - 1] = dp[..., n - 1]
    for k in range(n - 2, -1, -1):
        x[..., k] = dp[..., k] - cp[..., k] * x[..., k + 1]
    return x


def line_solve(g, r: float, vin, vout=None, tol: float = 1e-15, max_iter: int = 2000):
    """Returns (p, q, iout, iterations) for one crossbar g [n, m] (S)."""
    g = np.asarray(g, dtype=np.float64)
    n, m = g.shape
    vin = np.asarray(vin, dtype=np.float64).reshape(n)
    vout = np.zeros(m) if vout is None else np.asarray(vout, dtype=np.float64).reshape(m)
    a = 1.0 / float(r)
    c = vin[:, None] - vout[None, :]
    drow = g + 2 * a          # wordline diagonal: source at j = 0, open at j = m-1
    drow[:, m - 1] -= a
    dcol = (g + 2 * a).T      # bitline diagonal (columns as rows): open top i = 0, sink below i = n-1
    dcol[:, 0] -= a
    p = np.zeros((n, m))
    q = np.zeros((n, m))
    I_prev = None
    for it in range(1, max_iter + 1):
        p = _thomas(drow, -a, g * (c - q))
        q = _thomas(dcol, -a, (g * (c - p)).T).T
        I = (g * (c - p - q)).sum(0)
        if I_prev is not None:
            scale = max(np.abs(I).max(), 1e-300)
            if np.abs(I - I_prev).max() <= tol * scale:
                return p, q, I, it
        I_prev = I
    return p, q, I, max_iter
