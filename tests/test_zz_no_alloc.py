"""fadesim_solve_batch_noalloc (the batched solve with its workspace in a
device array of the library, no allocation) against fadesim_solve_batch:
the same outputs bit for bit, on two calls in a row, the second smaller
(the array keeps the first call's contents)."""
import numpy as np
import pytest

import oracle
from conftest import config2_instances

pytestmark = pytest.mark.gpu


def _both(cuda, g, r, vin, vout=None, **kw):
    import torch
    from fadesim_amd.solver import solve_batch
    T = lambda a: None if a is None else torch.tensor(a, device=cuda)
    args = (T(g), T(r), T(vin), T(vout))
    ref = solve_batch(*args, want_field=True, **kw)
    got = solve_batch(*args, want_field=True, noalloc=True, **kw)
    torch.cuda.synchronize()
    for name in ("iout", "sweeps", "status", "field"):
        assert torch.equal(getattr(got, name), getattr(ref, name)), name
    assert torch.equal(got.est.isnan(), ref.est.isnan())
    assert torch.equal(torch.nan_to_num(got.est), torch.nan_to_num(ref.est))
    return got


def test_noalloc_equals_the_allocating_entry_point(cuda):
    g, r, vin = config2_instances(300, 128, seed=41)
    big = _both(cuda, g, r, vin, tol=1e-6, ir2s_iters=4, corr_iters=4, test_first=True)
    assert int((big.status != 0).sum()) == 0
    iex = oracle.exact_batch(g[:4], r[:4], vin[:4])
    assert np.max(np.abs(big.iout[:4].cpu().numpy() - iex) / np.abs(iex)) <= 1e-6
    # second call: fewer, smaller crossbars with driven bitline ends, plain guess
    g2, r2, vin2 = config2_instances(40, 64, seed=42)
    vout2 = np.random.default_rng(3).uniform(0.0, 0.01, (40, 64))
    small = _both(cuda, g2, r2, vin2, vout2, tol=1e-6)
    assert int((small.status != 0).sum()) == 0


def test_noalloc_refuses_timing(cuda):
    import torch
    from fadesim_amd.solver import solve_batch
    g, r, vin = config2_instances(2, 64, seed=1)
    with pytest.raises(RuntimeError, match="FADESIM_F_TIMING"):
        solve_batch(torch.tensor(g, device=cuda), torch.tensor(r, device=cuda), torch.tensor(vin, device=cuda),
                    timing=True, noalloc=True)
