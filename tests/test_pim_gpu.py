"""PimTorch-level crossbar MVM through the GPU solvers: the converged batched
solve against the restatement driven by exact FP64 tile solves, the
reference-iteration mode against the restatement driven by the CPU hook
replay, and the sinh-nonlinear Picard solve against the converged fixed
point of the CPU hook replay."""
import numpy as np
import pytest

import oracle
from oracle import pim as opim

pytestmark = pytest.mark.gpu


def _exact_currents(r, C):
    def cur(vin, G):
        bsize, numX, R = vin.shape
        numY = G.shape[2] // C
        slots = G.reshape(numX, R, numY, C).transpose(0, 2, 1, 3).reshape(numX * numY, R, C)
        b, x, y = np.meshgrid(np.arange(bsize), np.arange(numX), np.arange(numY), indexing="ij")
        b, x, y = b.ravel(), x.ravel(), y.ravel()
        v = vin[b, x].astype(np.float64)
        out = np.zeros((bsize, numX, numY, C))
        nz = v.any(1)
        if nz.any():
            out[b[nz], x[nz], y[nz]] = oracle.exact_slots(slots, r, (x * numY + y)[nz], v[nz])
        return out.reshape(bsize, numX, numY * C)
    return cur


def _tie_bound(info, rescale, window):
    total = 0.0
    for (i, j, sign), a in info["adc_in"].items():
        frac = np.abs(a - np.floor(a) - 0.5)
        n = int((frac < window).sum())
        total += n * rescale * 2.0 ** (i + j + info["s1"] + info["s2"])
    return total


def _draws(sigma, seed=557):
    """Variation draws in call order from a generator seeded like the
    reference's network generator (commonConst.py:122-130)."""
    import torch
    gen = torch.Generator().manual_seed(seed)

    def vary(G):
        q = torch.randn(G.shape, generator=gen)
        return (torch.from_numpy(np.ascontiguousarray(G)) / torch.exp(q * sigma)).numpy()
    return vary


@pytest.mark.parametrize("sigma", [0.0, 0.2])
def test_converged_linear_layer_matches_exact_restatement(cuda, sigma):
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(11)
    insize, outsize, bsize = 200, 40, 2
    x = torch.rand(bsize, insize, generator=gen)
    w = torch.randn(insize, outsize, generator=gen)
    # 64x64 tiles keep the exact CPU solves short; the pipeline is the same
    P = pim.mlp_inference_params(nonlinear_cell_enable=False, arrRowSize=64, arrColSize=64)
    P.sigma = sigma
    pim.variations.reset_all()
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, params=P).cpu().numpy()
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), cell_bits=2, rows=64, cols=64, vmax=0.2,
                                  ron=P.Ron, roff=P.Roff, currents=_exact_currents(P.r_wire, 64),
                                  vary=_draws(sigma) if sigma else None)
    # the solve is within 1e-6 of the exact currents: an ADC input a moves by at
    # most ~1e-6 * 1023, so codes within that window of a tie may round either way
    bound = _tie_bound(info, 64 * 3 / 1023, 2e-3)
    assert np.max(np.abs(out - ref)) <= bound + 1e-5 * np.abs(ref).max()
    assert P.stats["solves"] > 0


@pytest.mark.parametrize("sigma,order", [(0.0, "batchBybatch"), (0.3, "nodeBynode")])
def test_reference_mode_layer_matches_hook_restatement(cuda, sigma, order):
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(12)
    insize, outsize, bsize = 70, 20, 2
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="reference", ir_drop_iter_time=40, sigma=sigma,
                         ir_drop_GS_method_mode=getattr(pim.GS_method, order))
    pim.variations.reset_all()
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, params=P).cpu().numpy()

    def hook_currents(vin, G):
        bsize_, numX, R = vin.shape
        numY = G.shape[2] // 64
        o, _ = oracle.hook_gs(vin.astype(np.float32), G.astype(np.float32), bsize_, R, 64, numX, numY,
                              40, 1 / 2.93, 1.90, False, 1e-7, False, [])
        return o
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), rows=64, cols=64, currents=hook_currents,
                                  vary=_draws(sigma) if sigma else None)
    bound = _tie_bound(info, 64 / 1023, 1e-4)
    assert np.max(np.abs(out - ref)) <= bound + 1e-6 * np.abs(ref).max()


def test_nonlinear_picard_matches_hook_fixed_point(cuda):
    import torch
    from fadesim_amd import pim
    from fadesim_amd.solver import solve_batch_nonlinear
    P = pim.mlp_inference_params()
    rng = np.random.default_rng(5)
    B, n = 3, 64
    s = rng.integers(0, 4, (B, n, n))
    G = (s.astype(np.float32) * np.float32(P.cellDeltaConduct) + np.float32(P.cellMinConduct)).astype(np.float32)
    vin = (rng.random((B, n)) < 0.5) * 0.2
    res = solve_batch_nonlinear(torch.tensor(G, device=cuda), P.r_wire, torch.tensor(vin, device=cuda),
                                params=P.nonlinear_params, tol=1e-6)
    assert int((res.status != 0).sum()) == 0 and res.outer >= 2
    got = res.iout.cpu().numpy()
    for b in range(B):
        ref, _ = oracle.hook_gs(vin[b][None, None].astype(np.float64), G[b][None].astype(np.float64), 1, n, n, 1, 1,
                                200000, 1 / P.r_wire, 1.9, True, 1e-14, True, P.nonlinear_params)
        ref = ref.reshape(n)
        assert np.max(np.abs(got[b] - ref)) / np.max(np.abs(ref)) <= 1e-6


def _mlp_layer1_tiles(P, picks):
    """Tile solves of layer 1 of the 784-256-10 MLP (the 128x128 sinh-cell
    workload): conductances [K, 128, 128] float32 and drives [K, 128] of the
    (input bit, image, row tile, weight slice, sign, column tile) picks, formed
    as splitArr_matmul_nn forms them."""
    import torch
    import torch.nn.functional as F
    from fadesim_amd import pim
    model = pim.MLP(params=P, seed=0)
    w = model.layers[0].fp_weight.detach().t()
    qw, _ = pim.quantization_tensor(w, 8, pim.TensorType.PN)
    arrays = [qw.clamp(min=0), (-qw).clamp(min=0)]
    imgs = torch.rand(8, 784, generator=torch.Generator().manual_seed(21))
    qin, _ = pim.quantization_tensor(imgs, 8, pim.TensorType.Normal)
    R = C = 128
    qpad = F.pad(qin, (0, 7 * R - 784))
    Gs, vs = [], []
    for (i, b, x, j, s, y) in picks:
        sl = F.pad((arrays[s] >> (2 * j)) & 3, (0, 2 * C - 256, 0, 7 * R - 784))[x * R:(x + 1) * R, y * C:(y + 1) * C]
        Gs.append(sl.to(torch.float32).mul(P.cellDeltaConduct).add(P.cellMinConduct))
        vs.append(((qpad[b, x * R:(x + 1) * R] >> i) & 1).to(torch.float64) * P.inputMaxVoltage)
    return torch.stack(Gs).numpy(), torch.stack(vs).numpy()


@pytest.mark.parametrize("guess", ["plain", "nn", "ir2s"])
def test_sinh_tiles_of_the_mlp_match_the_nonlinear_oracle(cuda, guess):
    """128x128 tiles of the MLP's first layer with sinh cells (a = 2.5 / V):
    the Picard solve from the plain, the learned and the IR2S initial guess
    reaches the oracle's converged nonlinear state within 1e-6 relative in
    every bitline."""
    import torch
    from fadesim_amd import pim, warmstart
    from fadesim_amd.solver import solve_batch_nonlinear
    P = pim.mlp_inference_params()
    picks = [(0, 0, 0, 0, 0, 0), (3, 1, 2, 1, 1, 0), (5, 2, 4, 2, 0, 1), (2, 3, 6, 3, 1, 1),
             (2, 4, 1, 3, 0, 1), (4, 5, 5, 0, 1, 0), (1, 6, 3, 2, 1, 1), (6, 7, 0, 1, 0, 0)]
    G, vin = _mlp_layer1_tiles(P, picks)
    assert vin.any(1).all()
    kw = {}
    if guess == "nn":
        net = warmstart.load_default(cuda)
        if net is None:
            pytest.skip("no trained weights committed")
        kw = dict(warm_fn=lambda Gz, rw, v: net.predict(Gz, rw, v))
    elif guess == "ir2s":
        kw = dict(ir2s_iters=4, corr_iters=4, test_first=True)
    res = solve_batch_nonlinear(torch.tensor(G, device=cuda), P.r_wire, torch.tensor(vin, device=cuda),
                                params=P.nonlinear_params, tol=1e-6, **kw)
    torch.cuda.synchronize()
    assert int((res.status != 0).sum()) == 0 and res.outer >= 2
    ref, its = oracle.exact_nonlinear_batch(G, P.r_wire, vin, P.nonlinear_params)
    err = np.abs(res.iout.cpu().numpy() - ref) / np.abs(ref)
    print(f"{guess}: Picard steps {res.outer}, sweeps {res.total_sweeps.tolist()}, max rel err {err.max():.2e}")
    assert err.max() <= 1e-6


def _tie_accounting(x, w, Ip, Iw, out_p, out_w, P, outsize):
    """Checks that the two runs of one layer on the same input differ only by
    ADC codes at rounding ties: every code that differs moves by one and the
    rounding boundary lies between its two pre-rounding values, and those
    codes, weighted as splitArr_matmul_nn weights them, account for the whole
    difference of the layer outputs.  Returns the number of differing codes."""
    import torch
    from fadesim_amd import pim
    _, s1 = pim.quantization_tensor(x, 8, pim.TensorType.Normal)
    _, s2 = pim.quantization_tensor(w, 8, pim.TensorType.PN)
    solved = Ip != 0
    rel = ((Ip - Iw).abs() / Ip.abs().clamp_min(1e-300))[solved]
    assert rel.max().item() <= 2e-6  # both within 1e-6 of the converged state
    ap = Ip.to(torch.float32).div(P.adcUseI).mul(1023)
    aw = Iw.to(torch.float32).div(P.adcUseI).mul(1023)
    dc = aw.round() - ap.round()
    moved = dc != 0
    assert dc.abs().max().item() <= 1
    lo, hi = torch.minimum(ap, aw)[moved], torch.maximum(ap, aw)[moved]
    edge = torch.floor(lo) + 0.5
    assert bool(((lo <= edge) & (edge <= hi)).all())
    nbit, nslice = Ip.shape[0], Ip.shape[1]
    i = torch.arange(nbit, device=Ip.device).view(-1, 1, 1)
    j = torch.arange(nslice, device=Ip.device).view(1, -1, 1)
    sgn = torch.tensor([1.0, -1.0], device=Ip.device).view(1, 1, -1)
    wt = torch.pow(2.0, (i + 2 * j + s1 + s2).double()) * sgn * torch.where(i == nbit - 1, -1.0, 1.0)
    rescale = P.arrRowSize * 3 / 1023
    delta = torch.einsum("ijsbxyc,ijs->byc", dc.double(), wt).reshape(Ip.shape[3], -1)[:, :outsize] * rescale
    torch.testing.assert_close((out_w - out_p).double(), delta, rtol=0, atol=1e-5 * out_p.abs().max().item())
    return int(moved.sum())


def test_mlp_1024_images_nn_warm_vs_plain(cuda):
    """The MLP workload of 1024 images with sinh cells, plain against NN-warm
    tile solves, layer by layer on the same inputs: every difference is an ADC
    rounding tie (_tie_accounting), and the classes agree on every image."""
    import torch
    from fadesim_amd import pim, warmstart
    net = warmstart.load_default(cuda)
    if net is None:
        pytest.skip("no trained weights committed")
    imgs = torch.rand(1024, 784, generator=torch.Generator().manual_seed(3)).to(cuda)
    P = pim.mlp_inference_params()
    model = pim.MLP(params=P, seed=0).to(cuda).eval()
    L1, L2 = model.layers

    def run(layer, x, warm):
        P.record, P.warm_start, P.stats = [], (net if warm else None), {}
        with torch.no_grad():
            out = layer(x)
        return out, P.record[0], P.stats["sweeps"]

    o1p, I1p, sw_p = run(L1, imgs, False)
    o1w, I1w, sw_w = run(L1, imgs, True)
    n1 = _tie_accounting(imgs, L1.fp_weight.t(), I1p, I1w, o1p - L1.fp_bias, o1w - L1.fp_bias, P, 256)
    a_p, a_w = torch.relu(o1p), torch.relu(o1w)
    o2p, I2p, _ = run(L2, a_p, False)
    o2x, I2x, _ = run(L2, a_p, True)
    n2 = _tie_accounting(a_p, L2.fp_weight.t(), I2p, I2x, o2p - L2.fp_bias, o2x - L2.fp_bias, P, 10)
    o2w, _, _ = run(L2, a_w, True)   # the NN-warm model end to end
    agree = (o2p.argmax(1) == o2w.argmax(1)).float().mean().item()
    print(f"layer-1 sweeps plain {sw_p} NN-warm {sw_w}; tie codes {n1} + {n2}; "
          f"logits max |diff| {(o2p - o2w).abs().max().item():.3e} of {o2p.abs().max().item():.3e}")
    assert agree == 1.0
    assert sw_w < sw_p


@pytest.mark.parametrize("mode", ["reference", "ir2s"])
def test_ou_mode_layer_matches_hook_restatement(cuda, mode):
    """OU mode through the hook's OU solvers (ir_gs_ou / ir2s_ou) against the
    restatement driven by their CPU replays."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(13)
    insize, outsize, bsize, ou = 70, 20, 2, (16, 8)
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    iters = 30 if mode == "reference" else 4
    P = pim.PhyArrParams(ir_mode=mode, ir_drop_iter_time=iters, useOUSize=True, OUSize=ou)
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, params=P).cpu().numpy()

    def hook_currents(vin, G):
        bsize_, numX, R = vin.shape
        numY = G.shape[2] // 64
        args = (vin.astype(np.float32), G.astype(np.float32), bsize_, R, 64, numX, numY, iters)
        if mode == "reference":
            o, _ = oracle.hook_gs_ou(*args, 1 / 2.93, 1.90, False, 1e-7, False, [], ou_rows=ou[0], ou_cols=ou[1])
        else:
            o, _ = oracle.hook_ir2s_ou(*args, 2.93, False, 1e-7, False, [], ou_rows=ou[0], ou_cols=ou[1])
        return o.reshape(bsize_, -1, G.shape[2])
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), rows=64, cols=64, currents=hook_currents, ou_rows=ou[0])
    bound = _tie_bound(info, ou[0] / 1023, 1e-4)
    assert np.max(np.abs(out - ref)) <= bound + 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("mode,ou", [("reference", None), ("ir2s", None), ("reference", (16, 8)), ("ir2s", (8, 8))])
def test_ref_arrays_match_hook_restatement(cuda, mode, ou):
    """Reference-column arrays through the hook's solvers (R x (C+2) tiles, or
    C + OUSize[1] columns in OU mode) against the restatement driven by their
    CPU replays."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(17)
    insize, outsize, bsize = 70, 20, 2
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    iters = 30 if mode == "reference" else 4
    P = pim.PhyArrParams(ir_mode=mode, ir_drop_iter_time=iters, useOUSize=ou is not None, OUSize=ou or (8, 8))
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, pim.phyArrMode.Ref,
                                 params=P).cpu().numpy()
    E = ou[1] if ou else 2

    def hook_currents(vin, G):
        bsize_, numX, R = vin.shape
        Cx = 64 + E
        numY = G.shape[2] // Cx
        args = (vin.astype(np.float32), G.astype(np.float32), bsize_, R, Cx, numX, numY, iters)
        extra = dict(ou_rows=ou[0], ou_cols=ou[1]) if ou else {}
        if mode == "reference":
            o, _ = oracle.hook_gs_ou(*args, 1 / 2.93, 1.90, False, 1e-7, False, [], **extra)
        else:
            o, _ = oracle.hook_ir2s_ou(*args, 2.93, False, 1e-7, False, [], **extra)
        return o.reshape(bsize_, -1, G.shape[2])
    ref, info = opim.split_matmul_ref(x.numpy(), w.numpy(), rows=64, cols=64, currents=hook_currents,
                                      ou_rows=ou[0] if ou else None, ou_cols=ou[1] if ou else None)
    per_read = ou[0] if ou else 64
    bound = _tie_bound(info, per_read / 1023 * 129, 1e-4)
    assert np.max(np.abs(out - ref)) <= bound + 1e-6 * np.abs(ref).max()


def test_conv2d_sp_on_the_gpu_matches_restatement(cuda):
    """A Conv2d_sp layer in eval mode with the hook's line-GS crossbars."""
    import torch
    from fadesim_amd import pim
    torch.manual_seed(5)
    P = pim.PhyArrParams(ir_mode="reference", ir_drop_iter_time=20)
    conv = pim.Conv2d_sp(4, 6, 3, padding=1, params=P).to(cuda).eval()
    x = torch.rand(2, 4, 6, 5, device=cuda)
    with torch.no_grad():
        y = conv(x).cpu().numpy()

    def hook_currents(vin, G):
        bsize_, numX, R = vin.shape
        numY = G.shape[2] // 64
        o, _ = oracle.hook_gs(vin.astype(np.float32), G.astype(np.float32), bsize_, R, 64, numX, numY, 20,
                              1 / 2.93, 1.90, False, 1e-7, False, [])
        return o
    ref = opim.conv2d_sp(x.cpu().numpy(), conv.fp_weight.detach().cpu().numpy(), conv.fp_bias.detach().cpu().numpy(),
                         (1, 1), (1, 1), (1, 1), rows=64, cols=64, currents=hook_currents)
    assert y.shape == (2, 6, 6, 5)
    np.testing.assert_allclose(y, ref, rtol=0, atol=2e-3 * np.abs(ref).max())


def test_ou_node_by_node_layer_runs_each_ou_block_alone(cuda):
    """OU mode with the nodeBynode order (ir_drop_solve.py:315-360): every OU
    block iterated as a crossbar of its own with plain wire ends (no source or
    sink loads) -- the restatement driven by the hook replay on the blocks."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(13)
    insize, outsize, bsize, our, ouc = 70, 20, 2, 16, 8
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="reference", ir_drop_iter_time=30, useOUSize=True, OUSize=(our, ouc),
                         ir_drop_GS_method_mode=pim.GS_method.nodeBynode)
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, params=P).cpu().numpy()

    def block_currents(vin, G):
        bsize_, numX, R = vin.shape
        nrb = numX * (R // our)
        vb = vin.astype(np.float32).reshape(bsize_, nrb, our)
        Gb = G.astype(np.float32).reshape(nrb, our, G.shape[2])
        o, _ = oracle.hook_gs(vb, Gb, bsize_, our, ouc, nrb, G.shape[2] // ouc, 30, 1 / 2.93, 1.90, False, 1e-7,
                              False, [])
        return o
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), rows=64, cols=64, currents=block_currents, ou_rows=our)
    bound = _tie_bound(info, our / 1023, 1e-4)
    assert np.max(np.abs(out - ref)) <= bound + 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(64, 66), (100, 40), (128, 130), (200, 250)])
def test_padded_crossbars_match_exact(cuda, shape):
    """Crossbars of shapes no kernel takes, solved as the smallest larger
    crossbar with the added cells off (solver.solve_padded: a batched shape,
    or a 256x256 mixed-size square), against the exact FP64 solve of the
    original shape at 1e-6 relative in every bitline; conductance slots
    shared through gidx."""
    import torch
    from fadesim_amd.solver import embedding_shape, solve_padded
    R, C = shape
    rng = np.random.default_rng(R * 1000 + C)
    S, B = 3, 8
    g = np.exp(rng.uniform(np.log(5e-7), np.log(1e-6), (S, R, C))).astype(np.float32)
    r = rng.uniform(1.0, 5.0, B)
    vin = 0.3 * rng.integers(0, 8, (B, R)) / 7.0
    gidx = rng.integers(0, S, B).astype(np.int32)
    res = solve_padded(torch.tensor(g, device=cuda), torch.tensor(r, device=cuda), torch.tensor(vin, device=cuda),
                       gidx=torch.tensor(gidx, device=cuda), tol=1e-6, ir2s_iters=4, corr_iters=4, test_first=True)
    torch.cuda.synchronize()
    assert (res.status.cpu().numpy() == 0).all()
    iex = oracle.exact_batch(g[gidx], r, vin)
    err = np.abs(res.iout.cpu().numpy() - iex) / np.abs(iex)
    print(f"{R}x{C} as {embedding_shape(R, C)}: max rel err {err.max():.2e}")
    assert err.max() <= 1e-6


def test_converged_ref_layer_matches_exact_restatement(cuda):
    """Reference-column arrays (R x (C+2) tiles) in the converged mode, solved
    as 64x128 crossbars with the added cells off, against the restatement
    driven by exact FP64 solves of the 64x66 tiles."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(21)
    insize, outsize, bsize = 130, 40, 2
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.mlp_inference_params(nonlinear_cell_enable=False, arrRowSize=64, arrColSize=64)
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, pim.phyArrMode.Ref,
                                 params=P).cpu().numpy()
    ref, info = opim.split_matmul_ref(x.numpy(), w.numpy(), cell_bits=2, rows=64, cols=64, vmax=0.2,
                                      ron=P.Ron, roff=P.Roff, currents=_exact_currents(P.r_wire, 66))
    bound = _tie_bound(info, 64 * 3 / 1023 * 129, 2e-3)
    assert np.max(np.abs(out - ref)) <= bound + 1e-5 * np.abs(ref).max()
    assert P.stats["solves"] > 0


@pytest.mark.parametrize("ref_cols", [False, True])
def test_converged_ou_layer_matches_converged_ou_iterations(cuda, ref_cols):
    """OU mode in the converged model (every OU block solved as its whole tile
    with the other cells off, 16x8 blocks of 64x64 tiles, PN arrays or
    reference-column arrays of 64 + 8 columns) against the restatement
    driven by the reference's OU iteration (ir_gs_ou's CPU replay in FP64,
    the block's own wire loads) run until no node moves by 1e-15 V."""
    import torch
    from fadesim_amd import pim
    gen = torch.Generator().manual_seed(23)
    insize, outsize, bsize, ou = 70, 20, 2, (16, 8)
    x = torch.rand(bsize, insize, generator=gen) - 0.3
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="converged", useOUSize=True, OUSize=ou)
    mode = pim.phyArrMode.Ref if ref_cols else pim.phyArrMode.PN
    out = pim.splitArr_matmul_nn(x.to(cuda), w.to(cuda), 8, 1, 10, 10, 8, mode, params=P).cpu().numpy()
    Cx = 64 + (ou[1] if ref_cols else 0)

    def converged_ou(vin, G):
        bsize_, numX, R = vin.shape
        numY = G.shape[2] // Cx
        o, sw = oracle.hook_gs_ou(vin.astype(np.float64), G.astype(np.float64), bsize_, R, Cx, numX, numY, 100000,
                                  1 / P.r_wire, 1.5, True, 1e-15, False, [], ou_rows=ou[0], ou_cols=ou[1])
        return o.reshape(bsize_, -1, G.shape[2])
    if ref_cols:
        ref, info = opim.split_matmul_ref(x.numpy(), w.numpy(), rows=64, cols=64, currents=converged_ou,
                                          ou_rows=ou[0], ou_cols=ou[1])
        bound = _tie_bound(info, ou[0] / 1023 * 129, 2e-3)
    else:
        ref, info = opim.split_matmul(x.numpy(), w.numpy(), rows=64, cols=64, currents=converged_ou, ou_rows=ou[0])
        bound = _tie_bound(info, ou[0] / 1023, 2e-3)
    assert np.max(np.abs(out - ref)) <= bound + 1e-5 * np.abs(ref).max()
    assert P.stats["solves"] > 0
