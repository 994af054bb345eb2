"""The learned warm start's base field formed by the solve itself
(solve_batch(warm_coarse=...) without warm): bit for bit the solve from the
features kernel's base field, certified or swept.

The learned warm start formed inside the solve (fadesim_nn_solve_batch):
bit for bit the outputs of the three-kernel path it fuses -- pooled features
and base field, the network, the combine while the solve loads -- with every
crossbar certified at its first test, with every crossbar swept by the
sweeping kernel from the start the guess kernel hands it, and with a network
whose guess the first test rejects; and the exact FP64 solve at the
tolerance."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _net(cuda):
    from fadesim_amd import warmstart
    net = warmstart.load_default(cuda)
    if net is None or not net.solves_in_kernel:
        pytest.skip("no committed coarse16f weights")
    return net


def _both(net, g, r, vin, **kw):
    from fadesim_amd.solver import solve_batch
    fused = solve_batch(g, r, vin, net=net, want_field=True, **kw)
    three = solve_batch(g, r, vin, **net.warm_kwargs(g, r, vin), test_first=True, want_field=True, **kw)
    return fused, three


def _assert_same(a, b):
    import torch
    torch.cuda.synchronize()
    for name in ("iout", "sweeps", "status", "est", "rejected", "field"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), name


def test_in_kernel_start_equals_three_kernel_path(cuda):
    """The headline NN-warm call (1 IR2S iteration, first test, refinement
    guess of 4 IR2S iterations) on 256 held-out instances: every output equal,
    every crossbar certified before a sweep, within 1e-6 of the exact solve."""
    from fadesim_amd import warmstart
    net = _net(cuda)
    g, r, vin = warmstart.make_instances(256, seed=30_000_000, device=cuda)
    fused, three = _both(net, g, r, vin, tol=1e-6, ir2s_iters=1, corr_iters=4)
    _assert_same(fused, three)
    assert int((fused.status != 0).sum()) == 0 and int(fused.sweeps.sum()) == 0
    assert int(fused.rejected.sum()) == 0
    k = 48
    iex = oracle.exact_batch(g[:k].cpu().numpy(), r[:k].cpu().numpy(), vin[:k].cpu().numpy())
    err = np.abs(fused.iout[:k].cpu().numpy() - iex) / np.abs(iex)
    print(f"in-kernel NN start: max rel err {err.max():.2e}")
    assert err.max() <= 1e-6


def test_in_kernel_start_swept_crossbars_equal(cuda):
    """Without the IR2S iteration and without the refinement step the
    network's guess alone reaches 1e-5 on most crossbars but not all: the
    guess kernel certifies those, hands the others' starts to the sweeping
    kernel, and every output still equals the three-kernel path's."""
    from fadesim_amd import warmstart
    net = _net(cuda)
    g, r, vin = warmstart.make_instances(64, seed=30_100_000, device=cuda)
    fused, three = _both(net, g, r, vin, tol=1e-5, ir2s_iters=0, corr_iters=4, refine=False)
    _assert_same(fused, three)
    sw = fused.sweeps.cpu().numpy()
    print(f"in-kernel NN start, no IR2S: sweeps {sw.min()}..{sw.max()}, mean {sw.mean():.1f}")
    assert int((sw > 0).sum()) >= 4 and int((fused.status != 0).sum()) == 0
    iex = oracle.exact_batch(g[:16].cpu().numpy(), r[:16].cpu().numpy(), vin[:16].cpu().numpy())
    assert np.max(np.abs(fused.iout[:16].cpu().numpy() - iex) / np.abs(iex)) <= 1e-5


def test_in_kernel_start_rejects_a_bad_network(cuda):
    """A network whose output is scaled 1e5x gives guesses the first test
    judges worse than the zero guess on most crossbars: both paths drop them
    for zero drops with the same IR2S iteration, bit for bit, report it, and
    sweep the rest from the bad guess."""
    from fadesim_amd import warmstart
    net = _net(cuda)
    out_op = [op for op in net.ops if op.name == "out"][0]
    w = net._view(net.wb, out_op)
    w.mul_(1e5)
    net._wver += 1
    g, r, vin = warmstart.make_instances(64, seed=30_200_000, device=cuda)
    fused, three = _both(net, g, r, vin, tol=1e-6, ir2s_iters=1, corr_iters=4)
    _assert_same(fused, three)
    assert int(fused.rejected.sum()) >= 16
    assert int((fused.status != 0).sum()) == 0


@pytest.mark.parametrize("case", ["headline", "swept"])
def test_base_formed_by_the_solve_equals_base_from_features(cuda, case):
    """warm_kwargs(base_in_kernel=True) -- the solve forms the base field, two
    undamped IR2S iterations from zero -- against the base written by the
    features kernel: every output equal, in the guess kernel (headline call)
    and in the sweeping kernel (no IR2S iteration, no refinement, 1e-5)."""
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    net = _net(cuda)
    g, r, vin = warmstart.make_instances(96, seed=30_300_000, device=cuda)
    kw = dict(tol=1e-6, ir2s_iters=1, corr_iters=4) if case == "headline" else dict(tol=1e-5, ir2s_iters=0,
                                                                                      refine=False)
    a = solve_batch(g, r, vin, **net.warm_kwargs(g, r, vin, base_in_kernel=True), test_first=True, want_field=True,
                    **kw)
    b = solve_batch(g, r, vin, **net.warm_kwargs(g, r, vin, base_in_kernel=False), test_first=True, want_field=True,
                    **kw)
    _assert_same(a, b)
    if case == "swept":
        assert int((a.sweeps > 0).sum()) >= 4
    assert int((a.status != 0).sum()) == 0
