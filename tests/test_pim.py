"""PimTorch-level crossbar MVM (fadesim_amd.pim) on the CPU: quantization
against hand-computed cases and the numpy restatement, the ideal-crossbar
pipeline against the restatement and against the integer product it
approximates, and the nonlinear-cell constants against their closed forms."""
import math

import numpy as np
import pytest
import torch

from oracle import pim as opim
from fadesim_amd import pim


def test_quantization_hand_cases():
    x = torch.tensor([0.5, -1.0, 0.25, 0.0])
    q, s = pim.quantization_tensor(x, 8, pim.TensorType.Normal)
    assert s == -6 and q.tolist() == [32, -64, 16, 0]          # s = ceil(log2(1/127))
    q, s = pim.quantization_tensor(x, 8, pim.TensorType.PN)
    assert s == -7 and q.tolist() == [64, -128, 32, 0]         # s = ceil(log2(1/255))
    # round half to even: 2.5 -> 2, 3.5 -> 4 at resolution 1 (max 127 -> s = 0)
    q, s = pim.quantization_tensor(torch.tensor([127.0, 2.5, 3.5, -2.5]), 8, pim.TensorType.Normal)
    assert s == 0 and q.tolist() == [127, 2, 4, -2]
    # all-zero tensor: max taken as 1e-12
    q, s = pim.quantization_tensor(torch.zeros(3), 8, pim.TensorType.Normal)
    assert s == math.ceil(math.log2(1e-12 / 127)) and q.tolist() == [0, 0, 0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quantization_matches_restatement(seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(1000) * 10 ** rng.uniform(-3, 3)).astype(np.float32)
    for bits in (4, 8):
        for kind, tt in (("normal", pim.TensorType.Normal), ("pn", pim.TensorType.PN)):
            q, s = pim.quantization_tensor(torch.from_numpy(x), bits, tt)
            qo, so = opim.quantize(x, bits, kind)
            assert s == so and np.array_equal(q.numpy(), qo)


def _near_tie_bound(info, rescale, outsize):
    """Largest output change that ADC codes within 1e-4 of a rounding boundary
    can cause (each such code may round either way)."""
    total = 0.0
    for (i, j, sign), a in info["adc_in"].items():
        frac = np.abs(a - np.floor(a) - 0.5)
        n = int((frac < 1e-4).sum())
        total += n * rescale * 2.0 ** (i + j + info["s1"] + info["s2"])
    return total


@pytest.mark.parametrize("insize,outsize,rows,cell_bits,seed", [(100, 30, 64, 1, 0), (200, 70, 64, 2, 1),
                                                                (130, 10, 128, 2, 2)])
def test_ideal_crossbar_matches_restatement(insize, outsize, rows, cell_bits, seed):
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand(3, insize, generator=gen) * 2 - 0.5
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="none", cellBits=cell_bits, arrRowSize=rows, arrColSize=rows)
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=P)
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), cell_bits=cell_bits, rows=rows, cols=rows,
                                  vmax=P.inputMaxVoltage, ron=P.Ron, roff=P.Roff)
    rescale = rows * (2 ** cell_bits - 1) / 1023
    bound = _near_tie_bound(info, rescale, outsize)
    assert np.max(np.abs(out.numpy() - ref)) <= bound + 1e-6 * np.abs(ref).max()


def test_ideal_crossbar_approximates_integer_product():
    """With no IR drop the crossbar computes the quantized product up to ADC
    rounding: |out - q_x q_w 2^(s1+s2)| <= sum over (bit, slice, sign) of
    numX * outsize-independent half-code errors."""
    gen = torch.Generator().manual_seed(7)
    insize, outsize, rows = 150, 20, 64
    x = torch.rand(4, insize, generator=gen)
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="none")
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=P).double()
    qx, s1 = pim.quantization_tensor(x, 8, pim.TensorType.Normal)
    qw, s2 = pim.quantization_tensor(w, 8, pim.TensorType.PN)
    exact = (qx.double() @ qw.double()) * 2.0 ** (s1 + s2)
    numX = -(-insize // rows)
    rescale = rows / 1023
    # Gmin leakage cancels between the pos and neg arrays; each code rounds by <= 1/2
    bound = sum(2 * numX * 0.5 * rescale * 2.0 ** (i + j + s1 + s2) for i in range(8) for j in range(8))
    assert torch.max(torch.abs(out - exact)).item() <= bound
    assert torch.max(torch.abs(out - exact)).item() < 0.05 * exact.abs().max().item()


def test_sinh_parameters_closed_form():
    P = pim.mlp_inference_params()
    Ron, Roff, a_on, b_on, a_off, b_off = P.nonlinear_params
    assert (Ron, Roff) == (1e6, 2e6)
    assert a_on == pytest.approx(2.5, rel=1e-12) and a_off == pytest.approx(2.5, rel=1e-12)
    # chord resistance at 1 V: V / (b sinh(a V)) = R
    assert 1.0 / (b_on * math.sinh(a_on)) == pytest.approx(Ron, rel=1e-12)
    assert 1.0 / (b_off * math.sinh(a_off)) == pytest.approx(Roff, rel=1e-12)
    assert P.adcUseI == pytest.approx(128 * (1e-6 - 5e-7) * 0.2)


def test_tile_instances_skip_zero_vectors():
    bits = torch.zeros(3, 2, 2, 8, dtype=torch.int64)
    bits[0, 1, 0, 3] = 1
    bits[2, 0, 1, 0] = 1
    ii, bb, xx, jj, ss, yy = pim.tile_instances(bits, nslice=4, numY=3)
    assert ii.numel() == 2 * 4 * 2 * 3
    assert set(zip(ii.tolist(), bb.tolist(), xx.tolist())) == {(0, 1, 0), (2, 0, 1)}


def test_converged_mode_refuses_cpu():
    P = pim.PhyArrParams(ir_mode="converged")
    with pytest.raises(RuntimeError):
        pim.splitArr_matmul_nn(torch.rand(1, 10), torch.randn(10, 3), 8, 1, 10, 10, 8, params=P)


@pytest.mark.parametrize("ou,seed", [((8, 8), 0), ((16, 4), 1)])
def test_ideal_ou_mode_matches_restatement(ou, seed):
    """OU mode with the ideal crossbar: one ADC reading per OU row block at
    full scale adcOUuseI, summed after the ADC."""
    gen = torch.Generator().manual_seed(20 + seed)
    insize, outsize, rows = 150, 40, 64
    x = torch.rand(3, insize, generator=gen) - 0.2
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="none", useOUSize=True, OUSize=ou)
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=P)
    ref, info = opim.split_matmul(x.numpy(), w.numpy(), rows=rows, cols=rows, vmax=P.inputMaxVoltage,
                                  ron=P.Ron, roff=P.Roff, currents=opim.ideal_currents_ou(ou[0]), ou_rows=ou[0])
    assert P.adcOUuseI == pytest.approx(ou[0] * (1e-3 - 1e-6))
    assert P.arrOUnum == (rows // ou[0], rows // ou[1])
    bound = _near_tie_bound(info, ou[0] / 1023, outsize)
    assert np.max(np.abs(out.numpy() - ref)) <= bound + 1e-6 * np.abs(ref).max()
    # the OU reading is a finer ADC: closer to the integer product than whole-array reads
    qx, s1 = pim.quantization_tensor(x, 8, pim.TensorType.Normal)
    qw, s2 = pim.quantization_tensor(w, 8, pim.TensorType.PN)
    exact = (qx.double() @ qw.double()) * 2.0 ** (s1 + s2)
    whole = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=pim.PhyArrParams(ir_mode="none")).double()
    assert torch.abs(out.double() - exact).mean() < torch.abs(whole - exact).mean()


def test_ou_size_checks():
    x, w = torch.rand(1, 10), torch.randn(10, 3)
    with pytest.raises(Exception, match="not the factor"):
        pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=pim.PhyArrParams(ir_mode="none", useOUSize=True,
                                                                                 OUSize=(6, 8)))
    with pytest.raises(Exception, match="not less than 2"):
        pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=pim.PhyArrParams(ir_mode="none", useOUSize=True,
                                                                                 OUSize=(8, 1)))
    with pytest.raises(NotImplementedError):  # converged OU blocks: linear cells only
        pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, params=pim.PhyArrParams(ir_mode="converged", useOUSize=True,
                                                                                 nonlinear_cell_enable=True))


# --- reference-column arrays and Conv2d_sp -------------------------------------

@pytest.mark.parametrize("insize,outsize,rows,cell_bits,ou,seed", [(100, 30, 64, 1, None, 0),
                                                                   (200, 70, 64, 2, None, 1),
                                                                   (150, 40, 64, 2, (8, 8), 2),
                                                                   (130, 20, 64, 1, (16, 4), 3)])
def test_ref_arrays_match_restatement(insize, outsize, rows, cell_bits, ou, seed):
    gen = torch.Generator().manual_seed(40 + seed)
    x = torch.rand(3, insize, generator=gen) * 2 - 0.5
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="none", cellBits=cell_bits, arrRowSize=rows, arrColSize=rows,
                         useOUSize=ou is not None, OUSize=ou or (8, 8))
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, pim.phyArrMode.Ref, params=P)
    cur = opim.ideal_currents_ou(ou[0]) if ou else opim.ideal_currents
    ref, info = opim.split_matmul_ref(x.numpy(), w.numpy(), cell_bits=cell_bits, rows=rows, cols=rows,
                                      vmax=P.inputMaxVoltage, ron=P.Ron, roff=P.Roff, currents=cur,
                                      ou_rows=ou[0] if ou else None, ou_cols=ou[1] if ou else None)
    per_read = ou[0] if ou else rows
    bound = _near_tie_bound(info, per_read * (2 ** cell_bits - 1) / 1023 * 129, outsize)
    assert np.max(np.abs(out.numpy() - ref)) <= bound + 1e-6 * np.abs(ref).max()


def test_ref_arrays_approximate_integer_product():
    """Reference columns remove the Gmin leakage (zero column) and the
    2^(wb-1) weight offset (ones column): the result is the two's complement
    product up to ADC rounding -- each code +-1/2, the zero-column
    subtraction +-1/2, and for slice 0 the ones column's +-1 code times 128."""
    gen = torch.Generator().manual_seed(8)
    insize, outsize, rows = 150, 20, 64
    x = torch.rand(4, insize, generator=gen)
    w = torch.randn(insize, outsize, generator=gen)
    P = pim.PhyArrParams(ir_mode="none")
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, pim.phyArrMode.Ref, params=P).double()
    qx, s1 = pim.quantization_tensor(x, 8, pim.TensorType.Normal)
    qw, s2 = pim.quantization_tensor(w, 8, pim.TensorType.Normal)
    exact = (qx.double() @ qw.double()) * 2.0 ** (s1 + s2)
    numX = -(-insize // rows)
    rescale = rows / 1023
    bound = sum(numX * (1.0 + (128.5 if j == 0 else 0.0)) * rescale * 2.0 ** (i + j + s1 + s2)
                for i in range(8) for j in range(8))
    err = torch.max(torch.abs(out - exact)).item()
    assert err <= bound
    assert err < 0.1 * exact.abs().max().item()
    # a plain offset-binary array without the reference columns would be far off
    assert torch.max(torch.abs(out - (qx.double() @ (qw.double() + 128)) * 2.0 ** (s1 + s2))).item() > 10 * err


@pytest.mark.parametrize("arr,k,stride,pad,dil,bias_on_array", [
    (pim.phyArrMode.PN, 3, 1, 1, 1, False), (pim.phyArrMode.PN, 2, 2, 0, 1, True),
    (pim.phyArrMode.Ref, 3, 1, 2, 2, False)])
def test_conv2d_sp_matches_restatement(arr, k, stride, pad, dil, bias_on_array):
    torch.manual_seed(3)
    P = pim.PhyArrParams(ir_mode="none")
    conv = pim.Conv2d_sp(3, 5, k, stride=stride, padding=pad, dilation=dil, phy_arr_mode=arr,
                         use_fp_bias=not bias_on_array, params=P)
    x = torch.rand(2, 3, 9, 8)
    conv.train()
    torch.testing.assert_close(conv(x), F_conv2d(x, conv.fp_weight, conv.fp_bias, stride, pad, dil))
    conv.eval()
    with torch.no_grad():
        y = conv(x)
        fl = F_conv2d(x, conv.fp_weight, conv.fp_bias, stride, pad, dil)
    assert y.shape == fl.shape
    mm = opim.split_matmul if arr == pim.phyArrMode.PN else opim.split_matmul_ref
    kw = dict(rows=64, cols=64, vmax=P.inputMaxVoltage, ron=P.Ron, roff=P.Roff)
    if bias_on_array:
        cols, (Ho, Wo) = opim.unfold(x.numpy(), k, k, (stride, stride), (pad, pad), (dil, dil))
        rows = cols.transpose(0, 2, 1).reshape(-1, 3 * k * k)
        rows = np.concatenate([rows, np.ones((rows.shape[0], 1), np.float32)], 1)
        wm = np.concatenate([conv.fp_weight.detach().numpy().reshape(5, -1).T,
                             conv.fp_bias.detach().numpy()[None]], 0)
        o, _ = mm(rows, wm, **kw)
        ref = o.reshape(2, Ho * Wo, 5).transpose(0, 2, 1).reshape(2, 5, Ho, Wo)
    else:
        ref = opim.conv2d_sp(x.numpy(), conv.fp_weight.detach().numpy(), conv.fp_bias.detach().numpy(),
                             (stride, stride), (pad, pad), (dil, dil), matmul=mm, **kw)
    np.testing.assert_allclose(y.numpy(), ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    # and it is a quantized approximation of the float convolution
    assert float((y - fl).abs().max()) < 0.15 * float(fl.abs().max())


def F_conv2d(x, w, b, stride, pad, dil):
    return torch.nn.functional.conv2d(x, w, b, stride, pad, dil)


# ------------------------------------------------------------------ variation
# conductance_variation.py:6-14 with the generators of commonConst.py:122-154

def _draws(sigma, seed=557):
    """The variation hook of the restatement: a CPU generator seeded like the
    reference's network generator (555 + 2), one draw per call."""
    gen = torch.Generator().manual_seed(seed)

    def vary(G):
        q = torch.randn(G.shape, generator=gen)
        return (torch.from_numpy(np.ascontiguousarray(G)) / torch.exp(q * sigma)).numpy()
    return vary


def test_variation_draw_is_the_seeded_lognormal():
    """The factor is exp(sigma q), q one randn draw per cell from the network
    generator (seed 557 after reset): the first draw equals that generator's
    first randn, and over many cells log(g0 / g) has mean 0 and standard
    deviation sigma (sigma is the std of q; a variance would give 0.3**0.5)."""
    pim.variations.reset_all()
    g0 = torch.full((2, 64, 128), 1e-6, dtype=torch.float32)
    g = pim.add_variation(g0, 0.3, pim.variations.rand_gen_network)
    q = torch.randn(g0.shape, generator=torch.Generator().manual_seed(557))
    assert torch.equal(g, g0 / torch.exp(q * 0.3))
    lq = torch.log(g0.double() / g.double())
    assert abs(lq.mean().item()) < 0.01 and abs(lq.std().item() - 0.3) < 0.01
    # the generator advances: the next call draws new values
    assert not torch.equal(pim.add_variation(g0, 0.3, pim.variations.rand_gen_network), g)
    # reset restarts the sequence
    pim.variations.network_gen_reset()
    assert torch.equal(pim.add_variation(g0, 0.3, pim.variations.rand_gen_network), g)


@pytest.mark.parametrize("arr,ou,seed", [(pim.phyArrMode.PN, None, 0), (pim.phyArrMode.PN, (8, 8), 1),
                                         (pim.phyArrMode.Ref, None, 2)])
def test_variation_matches_restatement(arr, ou, seed):
    """sigma != 0 through the ideal crossbar: one draw per (input bit, slice,
    array) over the padded [numX, R, numY*C] array, in the reference's loop
    order -- the restatement makes the draws in that loop itself."""
    gen = torch.Generator().manual_seed(seed)
    insize, outsize, rows = 100, 30, 64
    x = torch.rand(3, insize, generator=gen) * 2 - 0.5
    w = torch.randn(insize, outsize, generator=gen)
    kw = dict(useOUSize=True, OUSize=ou) if ou else {}
    P = pim.PhyArrParams(ir_mode="none", cellBits=2, arrRowSize=rows, arrColSize=rows, sigma=0.2, **kw)
    pim.variations.reset_all()
    out = pim.splitArr_matmul_nn(x, w, 8, 1, 10, 10, 8, arrType=arr, params=P).numpy()
    mm = opim.split_matmul if arr == pim.phyArrMode.PN else opim.split_matmul_ref
    okw = dict(cell_bits=2, rows=rows, cols=rows, vmax=P.inputMaxVoltage, ron=P.Ron, roff=P.Roff)
    if ou:
        okw.update(currents=opim.ideal_currents_ou(ou[0]), ou_rows=ou[0])
        if arr == pim.phyArrMode.Ref:
            okw["ou_cols"] = ou[1]
    ref, info = mm(x.numpy(), w.numpy(), vary=_draws(0.2), **okw)
    per_read = ou[0] if ou else rows
    bound = _near_tie_bound(info, per_read * 3 / 1023, outsize)
    assert np.max(np.abs(out - ref)) <= bound + 1e-6 * np.abs(ref).max()
    # the variation is visible, and a different draw order (slices outside
    # input bits) gives another answer: the comparison pins the order
    ref0, _ = mm(x.numpy(), w.numpy(), **okw)
    tol = bound + 1e-6 * np.abs(ref).max()
    assert np.max(np.abs(ref - ref0)) > 3 * tol
    ns = 2 if arr == pim.phyArrMode.PN else 1
    calls = iter(range(8 * 4 * ns))

    def vary_swapped(G):
        # the draw this call would get if slices were the outer loop
        i, rest = divmod(next(calls), 4 * ns)
        j, s = divmod(rest, ns)
        g2 = torch.Generator().manual_seed(557)
        for _ in range((j * 8 + i) * ns + s):
            torch.randn(G.shape, generator=g2)
        q = torch.randn(G.shape, generator=g2)
        return (torch.from_numpy(np.ascontiguousarray(G)) / torch.exp(q * 0.2)).numpy()
    wrong, _ = mm(x.numpy(), w.numpy(), vary=vary_swapped, **okw)
    assert np.max(np.abs(wrong - ref)) > 3 * tol


def test_gs_orders_give_the_same_iterates():
    """PimTorch's torch fallback (ir_drop_solve.py:191-292), both update orders
    transcribed with torch ops on the CPU: nodeBynode (:201-249) and
    batchBybatch (:250-292) give the same iterates bit for bit, so either
    order is the iteration the hook kernel runs; against the hook's CPU
    replay (the CUDA hook's arithmetic) they agree to FP32 rounding (the
    torch and CUDA forms round differently).  The OU form of nodeBynode
    (:315-360) is the whole-array iteration on each OU block."""
    import oracle
    gen = torch.Generator().manual_seed(4)
    bsize, numX, numY, R, C, it, beta = 2, 2, 2, 6, 5, 7, 1.9
    r_wire = 2.93
    inv = (torch.rand(bsize, numX, R, generator=gen) * 0.3).float()
    x = (1.0 / (1e3 + torch.rand(numX, R, numY * C, generator=gen) * 2e4)).float()
    g_wire = 1.0 / r_wire

    def start(inv, x, numX, numY, R, C):
        last_vup = inv.unsqueeze(-1).unsqueeze(-1).repeat(1, 1, 1, numY, C)
        last_vdown = x.new_zeros((inv.shape[0], numX, R, numY, C))
        return last_vup, last_vdown, last_vup.clone(), last_vdown.clone(), x.reshape(numX, R, numY, C)

    def node_by_node(inv, x, numX, numY, R, C):
        last_vup, last_vdown, now_vup, now_vdown, xx = start(inv, x, numX, numY, R, C)
        for _ in range(it):
            for i in range(R - 1, -1, -1):
                for j in range(C):
                    left = inv[..., i].unsqueeze(-1).repeat(1, 1, numY) if j == 0 else now_vup[:, :, i, :, j - 1]
                    lu, ld = last_vup[:, :, i, :, j], last_vdown[:, :, i, :, j]
                    xij = xx[:, i, :, j]
                    xo = xij.mul(ld)
                    lu1b, ld1b = lu.mul(1.0 - beta), ld.mul(1.0 - beta)
                    if j == C - 1:
                        now_vup[:, :, i, :, j] = (left.mul(g_wire) + xo).mul(beta).div(xij.add(g_wire)) + lu1b
                    else:
                        right = last_vup[:, :, i, :, j + 1]
                        now_vup[:, :, i, :, j] = ((left + right).mul(g_wire) + xo).mul(beta).div(
                            xij.add(2 * g_wire)) + lu1b
                    right = inv.new_zeros((inv.shape[0], numX, numY)) if i == R - 1 else now_vdown[:, :, i + 1, :, j]
                    xo = xij.mul(now_vup[:, :, i, :, j])
                    if i == 0:
                        now_vdown[:, :, i, :, j] = (right.mul(g_wire) + xo).mul(beta).div(xij.add(g_wire)) + ld1b
                    else:
                        left = last_vdown[:, :, i - 1, :, j]
                        now_vdown[:, :, i, :, j] = ((left + right).mul(g_wire) + xo).mul(beta).div(
                            xij.add(2 * g_wire)) + ld1b
            last_vup, now_vup = now_vup, last_vup
            last_vdown, now_vdown = now_vdown, last_vdown
        return (last_vup - last_vdown).mul(xx).sum(2).reshape(inv.shape[0], numX, -1)

    def batch_by_batch(inv, x, numX, numY, R, C):
        last_vup, last_vdown, now_vup, now_vdown, xx = start(inv, x, numX, numY, R, C)
        for _ in range(it):
            left = inv.unsqueeze(-1).repeat(1, 1, 1, numY)
            for j in range(C - 1):
                right = last_vup[..., j + 1]
                sumg = xx[..., j].add(g_wire * 2)
                now_vup[..., j] = ((left + right).mul(g_wire) + xx[..., j] * last_vdown[..., j]).mul(beta).div(
                    sumg) + last_vup[..., j].mul(1.0 - beta)
                left = now_vup[..., j]
            sumg = xx[..., -1].add(g_wire)
            now_vup[..., -1] = (left.mul(g_wire) + xx[..., -1] * last_vdown[..., -1]).mul(beta).div(sumg) + \
                last_vup[..., -1].mul(1.0 - beta)
            down = inv.new_zeros((inv.shape[0], numX, numY, C))
            for i in range(R - 1, 0, -1):
                up = last_vdown[:, :, i - 1, ...]
                sumg = xx[:, i, ...].add(g_wire * 2)
                now_vdown[:, :, i, ...] = ((up + down).mul(g_wire) + xx[:, i, ...] * now_vup[:, :, i, ...]).mul(
                    beta).div(sumg) + last_vdown[:, :, i, ...].mul(1.0 - beta)
                down = now_vdown[:, :, i, ...]
            sumg = xx[:, 0, ...].add(g_wire)
            now_vdown[:, :, 0, ...] = (down.mul(g_wire) + xx[:, 0, ...] * now_vup[:, :, 0, ...]).mul(beta).div(
                sumg) + last_vdown[:, :, 0, ...].mul(1.0 - beta)
            last_vup, now_vup = now_vup, last_vup
            last_vdown, now_vdown = now_vdown, last_vdown
        return (last_vup - last_vdown).mul(xx).sum(2).reshape(inv.shape[0], numX, -1)

    nb = node_by_node(inv, x, numX, numY, R, C).numpy()
    bb = batch_by_batch(inv, x, numX, numY, R, C).numpy()
    assert np.array_equal(nb, bb)
    hk, _ = oracle.hook_gs(inv.numpy(), x.numpy(), bsize, R, C, numX, numY, it, g_wire, beta, False, 0.0,
                           False, [])
    np.testing.assert_allclose(nb, hk, rtol=2e-6, atol=0)
    # OU blocks (3 x 5 of the 6 x 5 array here): nodeBynode on each block alone
    our, nrb = 3, numX * (R // 3)
    inv_b, x_b = inv.reshape(bsize, nrb, our), x.reshape(nrb, our, numY * C)
    nb_ou = node_by_node(inv_b, x_b, nrb, numY, our, C).numpy()
    assert np.array_equal(nb_ou, batch_by_batch(inv_b, x_b, nrb, numY, our, C).numpy())
    hk_ou, _ = oracle.hook_gs(inv_b.numpy(), x_b.numpy(), bsize, our, C, nrb, numY, it, g_wire, beta, False, 0.0,
                              False, [])
    np.testing.assert_allclose(nb_ou, hk_ou, rtol=2e-6, atol=0)
