"""The learned warm start's GPU kernels against library references:
convolution forward and weight gradient against torch (CPU, FP64) on the same
bf16-rounded operands; the whole network's weight gradients against torch
FP64 autograd of the same graph; the physics features against a numpy
evaluation of the superposition sums; and -- with trained weights present --
the warm start's effect on the solver on held-out seeds."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from conftest import config2_instances

pytestmark = pytest.mark.gpu


def _bf(t):
    import torch
    return t.to(torch.bfloat16)


def _torch_weight(wflat, cin, cout, kp):
    """[Cout][Kp] (k = tap*Cin + ci, bias at 9*Cin) -> torch [Cout][Cin][3][3], bias [Cout]."""
    w = wflat.view(cout, kp)
    wt = w[:, :9 * cin].reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).contiguous()
    return wt, w[:, 9 * cin].clone()


@pytest.mark.parametrize("H,W,cin,cout", [(16, 16, 8, 16), (32, 32, 16, 32), (16, 32, 32, 64), (16, 16, 96, 96),
                                          (32, 16, 64, 96)])
def test_conv_fwd_and_wgrad_match_torch(cuda, H, W, cin, cout):
    import torch
    import torch.nn.functional as F
    from fadesim_amd import _native as nat
    from fadesim_amd.warmstart import kpad
    B = 2
    kp = kpad(cin)
    gen = torch.Generator().manual_seed(H * 100 + cin)
    x = _bf(torch.randn(B, H, W, cin, generator=gen))
    w = torch.zeros(cout, kp)
    w[:, :9 * cin + 1] = torch.randn(cout, 9 * cin + 1, generator=gen) * 0.2
    w = _bf(w)
    xd, wd = x.to(cuda), w.to(cuda)
    y = torch.empty((B, H, W, cout), dtype=torch.bfloat16, device=cuda)
    s = ctypes.c_void_p(nat.stream_handle(cuda))
    L = nat.lib()
    nat.check(L.fadesim_nn_conv3x3_fwd(B, H, W, cin, cout, kp, ctypes.c_void_p(xd.data_ptr()),
                                       ctypes.c_void_p(wd.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1, s), "fwd")
    wt, bias = _torch_weight(w.double(), cin, cout, kp)
    xt = x.double().permute(0, 3, 1, 2).requires_grad_(False)
    wt.requires_grad_(True)
    bias.requires_grad_(True)
    ref = F.relu(F.conv2d(xt, wt, bias, padding=1)).permute(0, 2, 3, 1)
    got = y.cpu().double()
    assert torch.max(torch.abs(got - ref.detach())) <= 1e-2 * torch.max(torch.abs(ref)) + 1e-3
    # weight gradient (bias included) for an upstream gradient dz (ReLU mask applied by the caller)
    dz = _bf(torch.randn(B, H, W, cout, generator=gen)) * (ref.detach() > 0)
    dz = _bf(dz)
    dw = torch.zeros((cout, kp), dtype=torch.float32, device=cuda)
    nat.check(L.fadesim_nn_conv3x3_wgrad(B, H, W, cin, cout, kp, ctypes.c_void_p(xd.data_ptr()),
                                         ctypes.c_void_p(dz.to(cuda).data_ptr()), ctypes.c_void_p(dw.data_ptr()), s),
              "wgrad")
    pre = F.conv2d(xt, wt, bias, padding=1).permute(0, 2, 3, 1)
    (pre * dz.double()).sum().backward()
    gw = dw.cpu().double()
    gref = torch.zeros(cout, kp, dtype=torch.float64)
    gref[:, :9 * cin] = wt.grad.permute(0, 2, 3, 1).reshape(cout, 9 * cin)
    gref[:, 9 * cin] = bias.grad
    assert torch.allclose(gw, gref, rtol=2e-3, atol=2e-3 * gref.abs().max())
    # data gradient through the flipped weights equals torch's conv input gradient
    kpt = kpad(cout)
    if cin % 16 == 0:
        wT = torch.empty((cin, kpt), dtype=torch.bfloat16, device=cuda)
        nat.check(L.fadesim_nn_flip_weights(cin, cout, kp, kpt, ctypes.c_void_p(wd.data_ptr()),
                                            ctypes.c_void_p(wT.data_ptr()), s), "flip")
        dx = torch.empty((B, H, W, cin), dtype=torch.bfloat16, device=cuda)
        nat.check(L.fadesim_nn_conv3x3_fwd(B, H, W, cout, cin, kpt, ctypes.c_void_p(dz.to(cuda).data_ptr()),
                                           ctypes.c_void_p(wT.data_ptr()), ctypes.c_void_p(dx.data_ptr()), 0, s),
                  "dgrad")
        xt2 = x.double().permute(0, 3, 1, 2).requires_grad_(True)
        (F.conv2d(xt2, wt.detach(), bias.detach(), padding=1).permute(0, 2, 3, 1) * dz.double()).sum().backward()
        gx = xt2.grad.permute(0, 2, 3, 1)
        assert torch.max(torch.abs(dx.cpu().double() - gx)) <= 1e-2 * torch.max(torch.abs(gx)) + 1e-3


class _Bf16(object):
    """Rounds to bf16 where the kernels store bf16: the value going forward and
    the gradient coming back (torch autograd in FP64 otherwise)."""

    @staticmethod
    def fn():
        import torch

        class F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, x):
                return x.to(torch.bfloat16).to(x.dtype)

            @staticmethod
            def backward(ctx, g):
                return g.to(torch.bfloat16).to(g.dtype)
        return F.apply


def _torch_net(net, x_nhwc, wflat):
    """Same graph as fadesim_amd.warmstart.architecture() in torch FP64 (NCHW),
    rounded to bf16 wherever the kernels store bf16: every activation, every
    gradient an op passes back to its input, and the sum of the gradients of
    an activation used twice."""
    import torch.nn.functional as F
    from fadesim_amd.warmstart import kpad
    rnd = _Bf16.fn()
    acts = {"x": rnd(x_nhwc.permute(0, 3, 1, 2))}
    for op in net.ops:
        if op.kind == "conv":
            o = net.offsets[op.name]
            wt, b = _torch_weight(wflat[o:o + op.cout * kpad(op.cin)], op.cin, op.cout, kpad(op.cin))
            y = F.conv2d(rnd(acts[op.src]), wt, b, padding=1)
            acts[op.name] = rnd(F.relu(y) if op.relu else y)
        elif op.kind == "pool":
            acts[op.name] = rnd(F.avg_pool2d(rnd(acts[op.src]), 2))
        else:
            acts[op.name] = rnd(F.interpolate(rnd(acts[op.src]), scale_factor=2, mode="nearest") + rnd(acts[op.skip]))
    return acts["out"].permute(0, 2, 3, 1)


@pytest.mark.parametrize("arch", ["unet16", "coarse32", "coarse32f", "coarse16f"])
def test_network_gradients_match_torch_autograd(cuda, arch):
    """Weight gradients of the whole network (MFMA forward, weight- and
    data-gradient kernels, pooling, upsampling, skip sums) against torch FP64
    autograd of the same graph with the kernels' bf16 storage reproduced:
    what remains is FP32 accumulation order and the odd bf16 rounding it
    flips, so each layer's gradient agrees to a few parts in a thousand."""
    import torch
    from fadesim_amd import _native as nat
    from fadesim_amd.warmstart import OUT_C, WarmStartNet
    net = WarmStartNet(cuda, seed=3, arch=arch)
    B = 2
    R = 128 >> net.out_level
    Rin = 128 >> net.in_level
    gen = torch.Generator().manual_seed(5)
    x = _bf(torch.rand(B, Rin, Rin, 8, generator=gen))
    tgt = torch.randn(B, 2, R, R, generator=gen).float()
    acts = net.forward(x.to(cuda), keep=True)
    out = acts["out"]
    dout = torch.empty_like(out)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    nat.check(nat.lib().fadesim_nn_mse(B, R, R, ctypes.c_void_p(out.data_ptr()), OUT_C,
                                       ctypes.c_void_p(tgt.to(cuda).data_ptr()), ctypes.c_void_p(dout.data_ptr()),
                                       ctypes.c_void_p(loss.data_ptr()), ctypes.c_void_p(nat.stream_handle(cuda))),
              "mse")
    grad = torch.zeros(net.nparams_padded, dtype=torch.float32, device=cuda)
    net.backward(acts, dout, grad)
    torch.cuda.synchronize()
    wref = net.wb.cpu().double().requires_grad_(True)
    o = _torch_net(net, x.double(), wref)
    # the kernel's loss: sum over pixels and both output channels of d^2 / (2 P), P = pixels
    l = ((o[..., 0:2].permute(0, 3, 1, 2) - tgt.double()) ** 2).sum() / (2 * B * R * R)
    l.backward()
    assert abs(loss.item() - l.item()) <= 1e-3 * l.item()
    g = grad.cpu().double()
    worst = 0.0
    for op in net.ops:
        if op.kind != "conv":
            continue
        o0 = net.offsets[op.name]
        n = op.cout * (((9 * op.cin + 1 + 31) // 32) * 32)
        a, b = g[o0:o0 + n], wref.grad[o0:o0 + n]
        rel = float((a - b).norm() / (b.norm() + 1e-300))
        worst = max(worst, rel)
        assert rel < 2e-2, (op.name, rel)
    print(f"{arch}: largest relative gradient error of a layer {worst:.2e}")


def _np_first_order(g, r, c):
    n, m = g.shape
    I = g * c
    j = np.arange(m)
    W = np.minimum(j[:, None], j[None, :]) + 1.0
    i = np.arange(n)
    Wq = (n - np.maximum(i[:, None], i[None, :])).astype(float)
    return r * (I @ W.T), r * (Wq @ I)


def test_physics_features(cuda):
    import torch
    from fadesim_amd.warmstart import WarmStartNet
    net = WarmStartNet(cuda)
    g, r, vin = config2_instances(3, 128, seed=4)
    tg, tr, tv = (torch.tensor(a, device=cuda) for a in (g, r, vin))
    feat, base, scale = net.features(tg, tr, tv)
    base = base.cpu().numpy()
    for b in range(3):
        gd = g[b].astype(np.float64)
        c = np.repeat(vin[b][:, None], 128, 1)
        p1, q1 = _np_first_order(gd, r[b], c)
        p2, q2 = _np_first_order(gd, r[b], c - p1 - q1)
        np.testing.assert_allclose(base[b, 0], p2, rtol=1e-4, atol=1e-6 * np.abs(p2).max())
        np.testing.assert_allclose(base[b, 1], q2, rtol=1e-4, atol=1e-6 * np.abs(q2).max())
        s2 = max(np.abs(p2 - p1).max(), np.abs(q2 - q1).max())
        assert scale[b].item() == pytest.approx(s2, rel=1e-4)
        # the second-order field is a far better initial guess than zero
        iex = oracle.exact_batch(g[b:b + 1], r[b], vin[b:b + 1])[0]
        i2 = (gd * (c - p2 - q2)).sum(0)
        assert np.max(np.abs(i2 - iex) / np.abs(iex)) < 3e-3


def test_trained_warm_start_cuts_sweeps(cuda):
    import torch
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    net = warmstart.load_default(cuda)
    if net is None:
        pytest.skip("no trained weights committed")
    g, r, vin = warmstart.make_instances(256, seed=10_000_000, device=cuda)
    plain = solve_batch(g, r, vin, tol=1e-6)
    warm = solve_batch(g, r, vin, warm=net.predict(g, r, vin), tol=1e-6)
    assert int((warm.status != 0).sum()) == 0
    sp, sw = plain.sweeps.double().mean().item(), warm.sweeps.double().mean().item()
    # the trained network's field needs well under half of plain SOR's sweeps
    # (about a sixth when measured)
    assert sw < 0.4 * sp, (sp, sw)
    assert int(warm.rejected.sum()) == 0
    k = 8
    iex = oracle.exact_batch(g[:k].cpu().numpy(), r[:k].cpu().numpy(), vin[:k].cpu().numpy())
    err = np.max(np.abs(warm.iout[:k].cpu().numpy() - iex) / np.abs(iex))
    assert err <= 1e-6, err


@pytest.mark.parametrize("level", [1, 2])
def test_combine_up_matches_bilinear(cuda, level):
    """Upsampling combine of a coarse network output against torch's bilinear
    interpolation (half-pixel centres, edge clamping) in FP32."""
    import torch
    import torch.nn.functional as F
    from fadesim_amd import _native as nat
    B, M, N, C = 3, 128, 64, 16
    gen = torch.Generator().manual_seed(level)
    out = torch.randn(B, M >> level, N >> level, C, generator=gen).to(torch.bfloat16).to(cuda)
    base = torch.randn(B, 2, M, N, generator=gen).to(cuda)
    scale = torch.rand(B, generator=gen).to(cuda)
    warm = torch.empty(B, 2, M, N, device=cuda)
    rc = nat.lib().fadesim_nn_combine_up(B, M, N, ctypes.c_void_p(base.data_ptr()), ctypes.c_void_p(scale.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), C, level, ctypes.c_void_p(warm.data_ptr()),
                                         ctypes.c_void_p(nat.stream_handle(cuda)))
    nat.check(rc, "combine_up")
    up = F.interpolate(out[..., :2].permute(0, 3, 1, 2).float(), scale_factor=1 << level, mode="bilinear",
                       align_corners=False)
    ref = base + scale.view(B, 1, 1, 1) * up
    torch.testing.assert_close(warm, ref, rtol=1e-5, atol=1e-5)


def test_coarse_network_predicts_full_field(cuda):
    import torch
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    net = warmstart.WarmStartNet(cuda, arch="coarse32", seed=1)
    assert net.out_level == 2
    g, r, vin = warmstart.make_instances(8, seed=3, device=cuda)
    w = net.predict(g, r, vin)
    assert w.shape == (8, 2, 128, 128) and bool(torch.isfinite(w).all())
    # an untrained network's field is still only a warm start: the solve converges
    res = solve_batch(g, r, vin, warm=w, tol=1e-6, ir2s_iters=1, corr_iters=4)
    assert int((res.status != 0).sum()) == 0


def test_pooled_features_match_pooled_full_features(cuda):
    """fadesim_nn_features_pooled (register-tile wire solves, FP32) against the
    full-resolution features (FP64 scans) averaged over 4x4 cells, and its
    second-order field against the closed form."""
    import torch
    import torch.nn.functional as F
    from fadesim_amd.warmstart import WarmStartNet
    full = WarmStartNet(cuda, arch="coarse32")
    pooled = WarmStartNet(cuda, arch="coarse32f")
    assert (full.in_level, pooled.in_level) == (0, 2)
    g, r, vin = config2_instances(4, 128, seed=9)
    tg, tr, tv = (torch.tensor(a, device=cuda) for a in (g, r, vin))
    f0, b0, s0 = full.features(tg, tr, tv)
    f2, b2, s2 = pooled.features(tg, tr, tv)
    assert f2.shape == (4, 32, 32, 8)
    torch.testing.assert_close(s2, s0, rtol=1e-4, atol=0)
    for k in range(2):
        scale = b0[:, k].abs().amax(dim=(1, 2)).view(4, 1, 1)
        assert float(((b2[:, k] - b0[:, k]).abs() / scale).max()) < 1e-5
    ref = F.avg_pool2d(f0.float().permute(0, 3, 1, 2), 4).permute(0, 2, 3, 1)
    # features: bf16 values averaged (FP32) vs averages rounded to bf16
    torch.testing.assert_close(f2.float(), ref, rtol=2e-2, atol=2e-3)
    # a trained coarse network runs on them end to end
    w = pooled.predict(tg, tr, tv)
    assert w.shape == (4, 2, 128, 128) and bool(torch.isfinite(w).all())


@pytest.mark.parametrize("arch", ["coarse32f", "coarse16f"])
def test_fused_coarse_network_equals_layered(cuda, arch):
    """The one-kernel coarse network (activations in LDS) against the same
    network run layer by layer: the same GEMM fragments and order, so the
    outputs agree to the last bf16 bit."""
    import torch
    from fadesim_amd.warmstart import WarmStartNet
    net = WarmStartNet(cuda, arch=arch, seed=3)
    assert net.fused
    gen = torch.Generator().manual_seed(5)
    x = (torch.randn(37, 32, 32, 8, generator=gen) * 0.5).to(torch.bfloat16).to(cuda)
    a = net.forward(x)
    b = net.forward_fused(x)
    torch.cuda.synchronize()
    assert a.shape == b.shape == (37, 32, 32, 16)
    assert torch.equal(a, b), float((a.float() - b.float()).abs().max())


@pytest.mark.parametrize("arch", ["coarse32f", "coarse32", "coarse16f"])
def test_warm_start_combined_at_load_equals_combined_field(cuda, arch):
    """solve_batch(warm=base, warm_coarse=out, warm_scale=s) forms the warm
    start while loading with the arithmetic of fadesim_nn_combine_up: the
    initial field (zero sweeps) is that kernel's field bit for bit, and a
    converged solve from it is the same solve."""
    import torch
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    net = warmstart.WarmStartNet(cuda, arch=arch, seed=11)
    g, r, vin = warmstart.make_instances(96, seed=20_000_000, device=cuda)
    warm = net.predict(g, r, vin)
    kw = net.warm_kwargs(g, r, vin)
    assert "warm_coarse" in kw and kw["warm_level"] == net.out_level >= 1
    f0 = solve_batch(g, r, vin, **kw, fixed_sweeps=0, want_field=True)
    torch.cuda.synchronize()
    assert torch.equal(f0.field, warm)
    a = solve_batch(g, r, vin, warm=warm, tol=1e-6, ir2s_iters=1, corr_iters=4, test_first=True, want_field=True)
    b = solve_batch(g, r, vin, **kw, tol=1e-6, ir2s_iters=1, corr_iters=4, test_first=True, want_field=True)
    torch.cuda.synchronize()
    assert int((b.status != 0).sum()) == 0
    assert torch.equal(a.sweeps, b.sweeps)
    assert torch.equal(a.iout, b.iout)
    assert torch.equal(a.field, b.field)


def test_warm_coarse_arguments_checked(cuda):
    import torch
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    g, r, vin = warmstart.make_instances(4, seed=5, device=cuda)
    base = torch.zeros((4, 2, 128, 128), dtype=torch.float32, device=cuda)
    out = torch.zeros((4, 32, 32, 16), dtype=torch.bfloat16, device=cuda)
    s = torch.ones(4, dtype=torch.float32, device=cuda)
    with pytest.raises(ValueError):  # no scale
        solve_batch(g, r, vin, warm=base, warm_coarse=out, warm_scale=None, warm_level=2)
    with pytest.raises(ValueError):
        solve_batch(g, r, vin, warm=base, warm_coarse=out, warm_scale=s, warm_level=1)
    with pytest.raises(RuntimeError):
        # level 0 (a full-resolution "coarse" grid) is outside warm_level's range 1..6
        solve_batch(g, r, vin, warm=base, warm_coarse=torch.zeros((4, 128, 128, 2), dtype=torch.bfloat16,
                                                                   device=cuda), warm_scale=s, warm_level=0)


@pytest.mark.parametrize("level,ch", [(1, 2), (1, 4), (3, 16)])
def test_warm_coarse_levels_at_load(cuda, level, ch):
    """Every level the LDS staging takes at 128x128 (level 1: a 64x64 grid,
    2 x 64 x 64 floats of the 12800 available): the initial field formed at
    load equals fadesim_nn_combine_up's field bit for bit."""
    import torch
    from fadesim_amd import _native as nat
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    B, M, N = 5, 128, 128
    g, r, vin = warmstart.make_instances(B, seed=9, device=cuda)
    gen = torch.Generator().manual_seed(level * 10 + ch)
    out = (torch.randn(B, M >> level, N >> level, ch, generator=gen) * 1e-2).to(torch.bfloat16).to(cuda)
    base = (torch.randn(B, 2, M, N, generator=gen) * 1e-3).to(cuda)
    scale = torch.rand(B, generator=gen).to(cuda)
    warm = torch.empty(B, 2, M, N, device=cuda)
    nat.check(nat.lib().fadesim_nn_combine_up(B, M, N, ctypes.c_void_p(base.data_ptr()),
                                              ctypes.c_void_p(scale.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                              ch, level, ctypes.c_void_p(warm.data_ptr()),
                                              ctypes.c_void_p(nat.stream_handle(cuda))), "combine_up")
    f0 = solve_batch(g, r, vin, warm=base, warm_coarse=out, warm_scale=scale, warm_level=level, fixed_sweeps=0,
                     want_field=True)
    torch.cuda.synchronize()
    assert torch.equal(f0.field, warm)


def _headline_nn_call(net, g, r, vin, K):
    """The NN-warm step that bench.py times: the network's output combined
    with the base field while the solve loads it, one IR2S iteration on it,
    the first test before any sweep, K IR2S iterations for the refinement
    correction's guess."""
    from fadesim_amd.solver import solve_batch
    return solve_batch(g, r, vin, **net.warm_kwargs(g, r, vin), tol=1e-6, ir2s_iters=1, corr_iters=K,
                       test_first=True)


def test_headline_nn_call_meets_the_tolerance(cuda):
    """bench.py's NN-warm call on 128 held-out instances (seeds outside the
    training range) against the exact FP64 solve: within 1e-6 relative in
    every bitline, the guess kept for every instance."""
    import torch
    from fadesim_amd import warmstart
    net = warmstart.load_default(cuda)
    if net is None:
        pytest.skip("no trained weights committed")
    g, r, vin = warmstart.make_instances(128, seed=30_000_000, device=cuda)
    res = _headline_nn_call(net, g, r, vin, 4)
    torch.cuda.synchronize()
    assert int((res.status != 0).sum()) == 0
    assert int(res.rejected.sum()) == 0
    iex = oracle.exact_batch(g.cpu().numpy(), r.cpu().numpy(), vin.cpu().numpy())
    err = np.abs(res.iout.cpu().numpy() - iex) / np.abs(iex)
    print(f"headline NN-warm call: max rel err {err.max():.2e}, sweeps {np.bincount(res.sweeps.cpu().numpy())}")
    assert err.max() <= 1e-6


@pytest.mark.parametrize("kind", ["scaled1000", "noise"])
def test_bad_warm_start_falls_back_to_the_zero_guess(cuda, kind):
    """A guess estimated worse than the zero guess at the first test (after
    its IR2S iteration, which shrinks any error by the coupling's spectral
    radius, 0.01-0.05 here) is dropped: the solve continues exactly as the
    solve from zero drops with the same IR2S iterations (bit for bit), and
    says so in `rejected`."""
    import torch
    from fadesim_amd import warmstart
    from fadesim_amd.solver import solve_batch
    net = warmstart.load_default(cuda)
    if net is None:
        pytest.skip("no trained weights committed")
    g, r, vin = warmstart.make_instances(64, seed=31_000_000, device=cuda)
    w = net.predict(g, r, vin)
    if kind == "scaled1000":
        bad = 1000.0 * w
    else:
        bad = torch.randn(w.shape, generator=torch.Generator().manual_seed(1)).to(cuda) * (300 * w.abs().max())
    kw = dict(tol=1e-6, ir2s_iters=1, corr_iters=4, test_first=True, want_field=True)
    res = solve_batch(g, r, vin, warm=bad, **kw)
    zero = solve_batch(g, r, vin, **kw)
    good = solve_batch(g, r, vin, warm=w, **kw)
    torch.cuda.synchronize()
    assert int(res.rejected.sum()) == 64 and int(good.rejected.sum()) == 0
    assert int((res.status != 0).sum()) == 0
    assert torch.equal(res.iout, zero.iout) and torch.equal(res.sweeps, zero.sweeps)
    assert torch.equal(res.field, zero.field)
    iex = oracle.exact_batch(g[:8].cpu().numpy(), r[:8].cpu().numpy(), vin[:8].cpu().numpy())
    assert np.max(np.abs(res.iout[:8].cpu().numpy() - iex) / np.abs(iex)) <= 1e-6
