"""The PimTorch hook's unconverged iterates, pinned by linear algebra rather
than by a second transcription of its loops.

One sweep of ir_drop_gs (split_arr_matmul.cu:27-119, host loop :246-306)
visits the wordline nodes row by row left to right, using the new left
neighbour, the old right neighbour and the old bitline node, then the bitline
nodes column by column bottom to top, using the new lower neighbour, the old
upper one and the new wordline node, each update relaxed by beta.  That is
SOR for the KCL system A x = b with the unknowns in that visiting order:
with A = D - E - F (D diagonal, E the couplings to unknowns visited earlier,
F to those visited later),
    (D - beta E) x' = ((1 - beta) D + beta F) x + beta b.
This file builds A from the network (wire conductance g_wire for every
segment, source and sink loads g_wire, top and right ends open, sink at 0 V),
iterates the splitting with dense solves from the hook's initial state
(u = input, d = 0), and compares k sweeps of the CPU replay (CPU test) and of
the GPU kernel (GPU test) with it.  Sinh cells: before sweep k the cell
conductance is 1/R(V) of the state two sweeps old (the initial state for the
first two), with the cell model of split_arr_matmul.cu:142-196; the readout
uses the conductance of the last sweep.  Orders a plausible mistake would
produce (bitlines top to bottom, wordlines right to left, the bitline pass
reading the old wordline values) are shown to give different iterates, so the
comparison would catch them.
"""
import numpy as np
import pytest

import oracle


def _sinh_G(V, z, nl):
    V = np.where((V < 1e-5) & (V > -1e-5), np.where(V > 0, 1e-5, -1e-5), V)
    V = np.abs(V)
    return 1.0 / ((1 - z) * V / (nl[3] * np.sinh(nl[2] * V)) + z * V / (nl[5] * np.sinh(nl[4] * V)))


def _layer_split(n_lines, length, gw, g_cell, first_load, last_load, forward):
    """D, E, F of one layer's chains (n_lines independent chains of `length`
    nodes along axis 1 of g_cell), flattened line-major.  first_load /
    last_load: conductance from the chain's first / last node to its driven
    end (0: open).  forward: visited from index 0 upward."""
    N = n_lines * length
    D = np.zeros(N)
    E = np.zeros((N, N))
    F = np.zeros((N, N))
    for a in range(n_lines):
        for k in range(length):
            p = a * length + k
            D[p] = g_cell[a, k] + (first_load if k == 0 else gw) + (last_load if k == length - 1 else gw)
            for q_k in (k - 1, k + 1):
                if 0 <= q_k < length:
                    q = a * length + q_k
                    earlier = q_k < k if forward else q_k > k
                    (E if earlier else F)[p, q] = gw
    return D, E, F


def hook_dense(inp, Glin, gw, beta, k, nl=None, order="reference"):
    """Bitline currents after k sweeps of the hook's iteration on one n x m
    crossbar, from the matrix splitting above.  order: "reference",
    "top_down" (bitlines top to bottom), "right_to_left" (wordlines),
    "old_wordline" (bitline pass reads the old wordline values)."""
    n, m = Glin.shape
    u = np.repeat(inp[:, None], m, 1).astype(np.float64)
    d = np.zeros((n, m))
    z = None
    if nl is not None:
        z = np.clip((1.0 / Glin - nl[0]) / (nl[1] - nl[0]), 0, None)
    hist = [(u.copy(), d.copy()), (u.copy(), d.copy())]
    G = Glin.astype(np.float64)
    for _ in range(k):
        if nl is not None:
            u2, d2 = hist[-2]
            G = _sinh_G(u2 - d2, z, nl)
        # wordline layer: chains are rows; left end driven through g_wire, right end open
        Du, Eu, Fu = _layer_split(n, m, gw, G, gw, 0.0, order != "right_to_left")
        bu = np.zeros((n, m))
        bu[:, 0] = gw * inp
        rhs = ((1 - beta) * Du * u.ravel() + beta * (Fu @ u.ravel()) + beta * (bu + G * d).ravel())
        un = np.linalg.solve(np.diag(Du) - beta * Eu, rhs).reshape(n, m)
        # bitline layer: chains are columns (transpose); top end open, bottom sinks through g_wire
        Dd, Ed, Fd = _layer_split(m, n, gw, G.T, 0.0, gw, order == "top_down")
        usrc = u if order == "old_wordline" else un
        rhs = ((1 - beta) * Dd * d.T.ravel() + beta * (Fd @ d.T.ravel()) + beta * (G * usrc).T.ravel())
        dn = np.linalg.solve(np.diag(Dd) - beta * Ed, rhs).reshape(m, n).T
        u, d = un, dn
        hist.append((u.copy(), d.copy()))
    return ((u - d) * G).sum(0)


def _case(n, m, seed, nonlinear):
    rng = np.random.default_rng(seed)
    gw = 1 / 2.93
    G = 1.0 / rng.uniform(1e3, 2e4, (n, m))  # strong cells: the orders differ visibly
    nl = None
    if nonlinear:
        # sinh cells a = 2.5 / V with chord resistances Ron = 1 kOhm, Roff = 20 kOhm at 1 V
        from fadesim_amd import pim
        k2 = 2 * np.cosh(2.5 / 2)
        a_on, b_on = pim.get_nonlinear_params_from_K2(1.0, 1e3, k2)
        a_off, b_off = pim.get_nonlinear_params_from_K2(1.0, 2e4, k2)
        nl = [1e3, 2e4, a_on, b_on, a_off, b_off]
    inp = np.round(rng.uniform(0.05, 0.3, n), 3)
    return inp, G, gw, nl


CASES = [(2, 3, 1, 1.0, False), (3, 2, 2, 1.9, False), (4, 4, 7, 1.9, False), (5, 3, 3, 1.5, False),
         (4, 4, 6, 1.9, True), (3, 5, 9, 1.7, True)]


@pytest.mark.parametrize("n,m,k,beta,nonlinear", CASES)
def test_oracle_replay_equals_the_splitting(n, m, k, beta, nonlinear):
    inp, G, gw, nl = _case(n, m, n * 10 + m, nonlinear)
    ref = hook_dense(inp, G, gw, beta, k, nl)
    out, sweeps = oracle.hook_gs(inp[None, None], G[None], 1, n, m, 1, 1, k, gw, beta, False, 0.0, nl is not None,
                                 nl or [])
    assert sweeps == k
    np.testing.assert_allclose(out.ravel(), ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n,m,k,beta,nonlinear", CASES)
def test_wrong_orders_give_other_iterates(n, m, k, beta, nonlinear):
    inp, G, gw, nl = _case(n, m, n * 10 + m, nonlinear)
    ref = hook_dense(inp, G, gw, beta, k, nl)
    for order in ("top_down", "right_to_left", "old_wordline"):
        other = hook_dense(inp, G, gw, beta, k, nl, order)
        assert np.max(np.abs(other - ref) / np.abs(ref)) > 1e-9, order


@pytest.mark.gpu
@pytest.mark.parametrize("n,m,k,beta,nonlinear", CASES)
def test_gpu_hook_iterates_equal_the_splitting(cuda, n, m, k, beta, nonlinear):
    import torch
    from fadesim_amd import hook
    inp, G, gw, nl = _case(n, m, n * 10 + m, nonlinear)
    ref = hook_dense(inp, G, gw, beta, k, nl)
    ti = torch.tensor(inp[None, None], dtype=torch.float64, device=cuda)
    tg = torch.tensor(G[None], dtype=torch.float64, device=cuda)
    out = hook.ir_drop_gs(ti, tg, 1, n, m, 1, 1, k, gw, beta, False, 0.0, nl is not None, nl or [])
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy().ravel(), ref, rtol=1e-12, atol=0)
