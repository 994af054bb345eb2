"""PimTorch crossbar MVM path on the GPU: the caller side of the IR-drop hook.

Restates the fixed-point, bit-sliced crossbar matrix multiply of PimTorch --
``quantization_tensor`` (pimtorch/pimfixedpoint/fixedPoint/nn/fixedPointArithmetic.py:19-96),
``splitArr_matmul_nn`` (splitArrayArithmetic.py:1161-1260), ``matmul`` /
``IRdrop_matmul`` (splitArrayArithmetic.py:442-1018, PN and reference-column
arrays, whole-array and OU mode), ``Linear_sp`` (modules/linear_split.py:
13-78) and ``Conv2d_sp`` (modules/conv_split.py:31-136) -- with the same
arithmetic: Normal (two's complement) quantized inputs applied bit-serially,
PN weights split into positive/negative magnitudes and sliced ``cellBits``
per cell, conductance ``slice*dG + Gmin`` in the float type, per-tile ADC
``round(I/adcUseI*(2^adc-1))``, sum over row tiles, rescale and shift-add
with the MSB subtracted (complement code).

MI355X-first difference: instead of one hook call per (input bit, weight
slice, pos/neg) -- 2 * inputBits * weightBits/cellBits calls per layer --
every tile solve of the layer goes into batched solves: the distinct
conductance tiles become slots of ``solve_batch(..., gidx=...)`` and every
(bit, slice, sign, image, row tile, column tile) with a non-zero input vector
is one instance (an all-zero input vector gives exactly zero current and is
not solved).  ``PhyArrParams.ir_mode`` selects the crossbar model:

* ``"converged"`` -- fadesim_solve_batch to ``tol`` (linear cells), or the
  Picard iteration of solve_batch_nonlinear (sinh cells); optionally seeded
  by the learned warm start (``warm_start``);
* ``"reference"`` -- the reference's own line-GS iteration through
  fadesim_amd.hook.ir_drop_gs (``ir_drop_iter_time`` sweeps at
  ``ir_drop_GS_method_beta_scale``), one call per (bit, slice, sign) as
  PimTorch makes them;
* ``"ir2s"``      -- the reference's IR2S iteration through
  fadesim_amd.hook.ir2s (``ir_drop_mode.ir2s``, splitArrayArithmetic.py:466-476);
* ``"none"``      -- the ideal crossbar (``no_ir_drop``, ir_drop_solve.py:12-15).

OU mode (``useOUSize``, commonConst.py:238-247): only one ``OUSize[0] x
OUSize[1]`` block of a tile is active at a time, so every block's currents are
read by the ADC on their own (full scale ``adcOUuseI``) and summed after it
(splitArrayArithmetic.py:524-605, 833-840, 968-976).  Its crossbar models are
the hook's OU solvers (``"reference"``: ir_gs_ou, ``"ir2s"``: ir2s_ou) and the
ideal per-block product (``"none"``: no_ir_drop_cal_ou, ir_drop_solve.py:21-30).
"""
from __future__ import annotations

import math
from enum import Enum
from typing import Optional, Tuple

import torch
from torch import Tensor, nn
import torch.nn.functional as F


class TensorType(Enum):
    Normal = 0  # two's complement, n bits: -2^(n-1) .. 2^(n-1)-1
    PN = 2      # positive/negative magnitudes, n bits: -(2^n-1) .. 2^n-1


class phyArrMode(Enum):
    PN = 0
    Ref = 1


class GS_method(Enum):
    """Update order of PimTorch's torch line-GS fallback (commonConst.py:174-177,
    ir_drop_solve.py:191-292).  Whole arrays: both orders give the same
    iterates -- row by row from the bottom, every wordline node reads its new
    left and old right neighbour and its old bitline node, every bitline node
    its new lower and old upper neighbour and its new wordline node, in the
    same float expressions -- so both run the hook kernel.  OU blocks: the
    nodeBynode form (:315-360) leaves out the OU's source and sink loads (plain
    wire ends), i.e. the whole-array iteration on each OU block."""
    nodeBynode = 1
    batchBybatch = 2


class variations:
    """The seeded CPU generators of the conductance variation
    (commonConst.py:122-154): network arrays draw from rand_gen_network (seed
    + 2); reset_all / network_gen_reset restart them."""
    var_defaultSeed = 555
    rand_gen_memCMOS = torch.Generator()
    rand_gen_mirrored = torch.Generator()
    rand_gen_network = torch.Generator()

    @staticmethod
    def network_gen_reset(seed: int = 555):
        variations.rand_gen_network.manual_seed(seed + 2)

    @staticmethod
    def mirrored_gen_reset(seed: int = 555):
        variations.rand_gen_mirrored.manual_seed(seed + 1)

    @staticmethod
    def memCMOS_gen_reset(seed: int = 555):
        variations.rand_gen_memCMOS.manual_seed(seed)

    @staticmethod
    def reset_all(seed: int = 555):
        variations.memCMOS_gen_reset(seed)
        variations.mirrored_gen_reset(seed)
        variations.network_gen_reset(seed)


variations.reset_all()


def add_variation(conductance: Tensor, sigma: float = 0.0, rand_gen: Optional[torch.Generator] = None) -> Tensor:
    """Lognormal cell variation (conductance_variation.py:6-14): R = R0 exp(q),
    q ~ N(0, sigma^2) per cell, so g = g0 / exp(q).  q is drawn on the CPU
    (torch.randn from rand_gen, FP32, in the tensor's shape) and the factor
    moved to the tensor's device, as the reference does, so a seeded
    generator gives the reference's draws."""
    factor = torch.randn(conductance.shape, generator=rand_gen).mul_(sigma).exp_()
    return conductance / factor.to(conductance.device)


class ConstForSplitArray:
    """Default bit widths of commonConst.py:45-50."""
    input_bit_width = 8
    output_bit_width = 10
    weight_bit_width = 8
    adc_bit_width = 10
    dac_bit_width = 1


def get_nonlinear_params_from_K2(nonlinear_V: float, R: float, k2: float) -> Tuple[float, float]:
    """(a, b) of I = b sinh(a V) with chord resistance R at nonlinear_V and
    I(V)/I(V/2) = k2 (conductance_variation.py:29-33)."""
    x = (k2 + math.sqrt(k2 * k2 - 4)) / 2
    a = 2 * math.log(x) / nonlinear_V
    b = nonlinear_V / R / math.sinh(a * nonlinear_V)
    return a, b


class PhyArrParams:
    """The crossbar constants of PimTorch's phyArrParams (commonConst.py:187-325)
    that the MVM path reads, with the same defaults.  Derived values are
    properties so that changing Ron/Roff/cellBits/arrRowSize keeps them
    consistent."""

    def __init__(self, **kw):
        self.Ron = 1e3
        self.Roff = 1e6
        self.sigma = 0.0
        self.cellBits = 1
        self.arrRowSize = 64
        self.arrColSize = 64
        self.inputMaxVoltage = 1.0
        self.r_wire = 2.93
        self.ir_mode = "reference"            # converged | reference | ir2s | none
        self.useOUSize = False                # OU mode: OUSize blocks activated one at a time
        self.OUSize = (8, 8)
        self.ir_drop_iter_time = 600
        self.ir_drop_GS_method_beta_scale = 1.90
        self.ir_drop_GS_method_mode = GS_method.batchBybatch  # update order (both orders: same iterates)
        self.ir_drop_less_than_thresholdBreak_enable = False
        self.ir_drop_less_than_threshold = 1e-7
        self.nonlinear_cell_enable = False
        self.nonlinear_params = []            # Ron, Roff, a_on, b_on, a_off, b_off
        self.tol = 1e-6                       # converged mode: relative I_out tolerance
        self.warm_start = None                # converged mode: WarmStartNet or None
        self.record = None                    # a list: every layer call appends its tile currents
        self.ir2s_iters = 0                   # converged mode: IR2S iterations on each solve's initial guess
        self.ir2s_omega = 1.0                 # their damping in (0, 1]; negative: the Chebyshev schedule
        self.max_batch = 65536                # instances per batched solve
        self.stats = {}                       # converged mode: solves, sweeps, Picard steps (accumulated)
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    @property
    def cellMaxConduct(self):
        return 1.0 / self.Ron

    @property
    def cellMinConduct(self):
        return 1.0 / self.Roff

    @property
    def cellDeltaConduct(self):
        return (self.cellMaxConduct - self.cellMinConduct) / (2 ** self.cellBits - 1)

    @property
    def adcUseI(self):
        return self.arrRowSize * (self.cellMaxConduct - self.cellMinConduct) * self.inputMaxVoltage

    @property
    def adcOUuseI(self):
        return self.OUSize[0] * (self.cellMaxConduct - self.cellMinConduct) * self.inputMaxVoltage

    @property
    def arrOUnum(self):
        return (self.arrRowSize // self.OUSize[0], self.arrColSize // self.OUSize[1])

    def check_ou(self):
        """The OU size checks of commonConst.py:244-247."""
        if self.arrRowSize % self.OUSize[0] != 0 or self.arrColSize % self.OUSize[1] != 0:
            raise Exception("[Error] ou size is not the factor of the crx size.")
        if self.OUSize[1] < 2 or self.OUSize[0] < 2:
            raise Exception("[Error] ou size is not less than 2.")

    def set_sinh_cells(self, a: float, nonlinear_V: float = 1.0):
        """Sinh cells I = b sinh(a V) for both states, b from the chord
        resistance (Ron / Roff) at nonlinear_V as get_nonlinear_params_from_K2
        does; a = 2 ln(x)/V  <=>  K2 = 2 cosh(a V / 2)."""
        k2 = 2 * math.cosh(a * nonlinear_V / 2)
        a_on, b_on = get_nonlinear_params_from_K2(nonlinear_V, self.Ron, k2)
        a_off, b_off = get_nonlinear_params_from_K2(nonlinear_V, self.Roff, k2)
        self.nonlinear_cell_enable = True
        self.nonlinear_params = [self.Ron, self.Roff, a_on, b_on, a_off, b_off]
        return self


phyArrParams = PhyArrParams()


def mlp_inference_params(**kw) -> PhyArrParams:
    """The 784-256-10 MLP inference workload's crossbar: 128x128 tiles, 2 bits
    per cell on the linear map Gmin = 1/2 MOhm .. Gmax = 1/1 MOhm (the
    conductance range of the 4-state 128x128 solver workload the learned warm
    start is trained on), 2.93 ohm wires, 0 / 0.2 V bit-serial inputs and sinh
    cells with a = 2.5 / V (chord calibration at 1 V), every tile solved to
    1e-6 by the batched solver."""
    P = PhyArrParams(Ron=1e6, Roff=2e6, cellBits=2, arrRowSize=128, arrColSize=128, inputMaxVoltage=0.2,
                     r_wire=2.93, ir_mode="converged")
    P.set_sinh_cells(2.5)
    for k, v in kw.items():
        if not hasattr(P, k):
            raise AttributeError(k)
        setattr(P, k, v)
    return P


def _get_fixed_point_position(max_abs: float, bit_width: int, tensor_type: TensorType) -> int:
    """fixedPointArithmetic.py:19-27 (zero max -> 1e-12)."""
    if math.isclose(max_abs, 0.0):
        max_abs = 1e-12
    if tensor_type == TensorType.Normal:
        return math.ceil(math.log2(max_abs / ((1 << (bit_width - 1)) - 1)))
    if tensor_type == TensorType.PN:
        return math.ceil(math.log2(max_abs / ((1 << bit_width) - 1)))
    raise ValueError("Unknown tensor type : " + str(tensor_type))


def quantization_tensor(tensor: Tensor, bit_width: int, tensor_type: TensorType) -> Tuple[Tensor, int]:
    """fixedPointArithmetic.py:76-96: int = round(x / 2^s) (half to even), s from
    the largest magnitude.  Returns (int32 tensor, s)."""
    s = _get_fixed_point_position(tensor.abs().max().item(), bit_width, tensor_type)
    return tensor.div(2.0 ** s).round().to(torch.int32), s


def _tiles(n: int, size: int) -> int:
    return (n + size - 1) // size


def splitArr_matmul_nn(input: Tensor, w: Tensor, inputBits: int, dacBits: int, adcBits: int,
                       outputBits: int, weightBits: int, arrType: phyArrMode = phyArrMode.PN,
                       float_type: torch.dtype = torch.float32,
                       params: Optional[PhyArrParams] = None) -> Tensor:
    """output [bsize, outsize] = input [bsize, insize] x w [insize, outsize]
    through the simulated crossbars (complement-code inputs, 1-bit DAC).

    arrType PN: positive/negative magnitude arrays (splitArrayArithmetic.py:
    1183-1226).  arrType Ref: one array per weight slice holding the two's
    complement weight offset by 2^(weightBits-1) (:1228-1258), each tile with
    a zero column and a ones column appended (OU mode: a block of OUSize[1]
    columns, cells past the ones column at 1e-20 S; :765-785); the zero
    column's ADC code is subtracted from every column and, for slice 0, the
    ones column's code times 2^(weightBits-1) (:996-1012)."""
    P = params or phyArrParams
    ref = arrType == phyArrMode.Ref
    if arrType not in (phyArrMode.PN, phyArrMode.Ref):
        raise NotImplementedError(f"array mode {arrType}")
    if dacBits != 1:
        raise Exception("[Error] Complement code input mode, but dac is not 1 bit. dacbits = {}".format(dacBits))
    if weightBits % P.cellBits != 0:
        raise Exception("[Error] cellbit is not factor of weight_bit")
    with torch.no_grad():
        bsize, insize = input.shape
        outsize = w.shape[1]
        qin, s1 = quantization_tensor(input, inputBits, TensorType.Normal)
        R, C = P.arrRowSize, P.arrColSize
        numX, numY = _tiles(insize, R), _tiles(outsize, C)
        Rp, Cp = numX * R, numY * C
        nslice = weightBits // P.cellBits
        mask_cell = (1 << P.cellBits) - 1
        inputV = P.inputMaxVoltage / (2 ** dacBits - 1)
        if P.useOUSize:
            P.check_ou()
        if ref:
            qw, s2 = quantization_tensor(w, weightBits, TensorType.Normal)
            plus = 2 ** (weightBits - 1)
            qw = (qw + plus).clamp(max=2 ** weightBits - 1)
            arrays = [qw]
            E = P.OUSize[1] if P.useOUSize else 2  # reference columns per tile
        else:
            qw, s2 = quantization_tensor(w, weightBits, TensorType.PN)
            arrays = [qw.clamp(min=0), (-qw).clamp(min=0)]
            E = 0
        ns = len(arrays)
        Cx = C + E
        # slices [nslice][ns][Rp][Cp]; padded cells hold slice 0 (conductance Gmin)
        slices = torch.stack([torch.stack([(a >> (P.cellBits * j)) & mask_cell for a in arrays])
                              for j in range(nslice)])
        slices = F.pad(slices, (0, Cp - outsize, 0, Rp - insize)).view(nslice, ns, numX, R, numY, C)
        if ref:
            extra = slices.new_zeros((nslice, ns, numX, R, numY, E))
            extra[..., 1] = 1
            slices = torch.cat((slices, extra), -1)
        G = slices.to(float_type).mul(P.cellDeltaConduct).add(P.cellMinConduct)  # matmul :753-754, :777-783
        if ref and P.useOUSize:
            G[..., C + 2:] = 1e-20
        if P.sigma != 0.0:
            # the reference draws a fresh variation for every (input bit,
            # slice, array) call of matmul, in that loop order (:1197-1227
            # around :757-760 / :788), over the padded [numX, R, numY*Cx]
            # array: one conductance set per input bit.  (At sigma = 0 the
            # reference still draws; this path skips the draws.)
            gen = variations.rand_gen_network
            G = torch.stack([torch.stack([torch.stack([
                add_variation(G[j, s].reshape(numX, R, numY * Cx), P.sigma, gen).view(numX, R, numY, Cx)
                for s in range(ns)]) for j in range(nslice)]) for _ in range(inputBits)])
        else:
            G = G.unsqueeze(0)
        G = G.permute(0, 1, 2, 3, 5, 4, 6).contiguous()  # [nv][nslice][ns][numX][numY][R][Cx], nv = 1 or inputBits
        # input bits [inputBits][bsize][numX][R] (two's complement bits of qin)
        qpad = F.pad(qin, (0, Rp - insize))
        bits = torch.stack([(qpad >> i) & 1 for i in range(inputBits)]).view(inputBits, bsize, numX, R)
        # tile currents [inputBits][nslice][ns][bsize][numX][numY][Cx]; in OU
        # mode one entry per OU row block: [..][numX * arrOUnum[0]][numY][Cx]
        if P.ir_mode == "none":
            vin = bits.to(float_type).mul(inputV)
            Gb = G.expand(inputBits, *G.shape[1:])
            if P.useOUSize:
                our = P.OUSize[0]
                I = torch.einsum("ibxkr,ijsxykrc->ijsbxkyc", vin.view(inputBits, bsize, numX, R // our, our),
                                 Gb.reshape(inputBits, nslice, ns, numX, numY, R // our, our, Cx))
                I = I.reshape(inputBits, nslice, ns, bsize, numX * (R // our), numY, Cx)
            else:
                I = torch.einsum("ibxr,ijsxyrc->ijsbxyc", vin, Gb)
        elif P.ir_mode == "converged":
            I = _solve_converged_ou(bits, G, inputV, P) if P.useOUSize else _solve_converged(bits, G, inputV, P)
        elif P.ir_mode in ("reference", "ir2s"):
            I = _solve_reference(bits, G, inputV, P, float_type)
        else:
            raise ValueError(f"unknown ir_mode {P.ir_mode}")
        # ADC per tile or per OU (matmul :833-840), sum over row tiles / OUs and rescale (:958-976)
        adc_full, rows_per_read = (P.adcOUuseI, P.OUSize[0]) if P.useOUSize else (P.adcUseI, R)
        adc = I.to(float_type).div(adc_full).mul(2 ** adcBits - 1).round()
        if ref:  # reference columns (:996-1012)
            adc = adc - adc[..., C:C + 1]
            adc[:, 0] = adc[:, 0] - adc[:, 0, ..., C + 1:C + 2].mul(plus).round()
            adc = adc[..., :C]
        out = adc.sum(4).reshape(inputBits, nslice, ns, bsize, Cp)[..., :outsize]
        out = out.mul(rows_per_read * (2 ** dacBits - 1) * (2 ** P.cellBits - 1) / (2 ** adcBits - 1))
        out = out[:, :, 0] if ref else out[:, :, 0] - out[:, :, 1]  # [inputBits][nslice][bsize][outsize]
        output = w.new_zeros((bsize, outsize))
        for i in range(inputBits):  # splitArr_matmul_nn :1197-1227, :1245-1256
            for jj in range(nslice):
                term = out[i, jj].to(w.dtype).mul(2.0 ** (i + P.cellBits * jj + s1 + s2))
                output = output.sub(term) if i == inputBits - 1 else output.add(term)
        return output


def tile_instances(bits: Tensor, nslice: int, numY: int, ns: int = 2):
    """Instance list of a layer's crossbar solves: every (bit, image, row tile)
    with a non-zero input vector x every (slice, array, column tile), ns
    arrays per slice (2: positive and negative, 1: a reference-column array).
    Returns index tensors (i, b, x, j, s, y), each [K]."""
    dev = bits.device
    ii, bb, xx = torch.nonzero(bits.any(-1), as_tuple=True)
    J, S, Y = torch.meshgrid(torch.arange(nslice, device=dev), torch.arange(ns, device=dev),
                             torch.arange(numY, device=dev), indexing="ij")
    J, S, Y = J.reshape(-1), S.reshape(-1), Y.reshape(-1)
    n_in, n_t = ii.numel(), J.numel()
    return (ii.repeat_interleave(n_t), bb.repeat_interleave(n_t), xx.repeat_interleave(n_t),
            J.repeat(n_in), S.repeat(n_in), Y.repeat(n_in))


def _solve_converged(bits, G, inputV, P):
    """All tile solves of the layer through the batched solver; tiles of a
    shape no kernel takes (reference-column arrays: R x (C+2)) as the
    smallest larger crossbar with the added cells off (solver.solve_padded,
    linear cells)."""
    from .solver import SUPPORTED_SHAPES, solve_batch, solve_batch_nonlinear, solve_padded
    nbit, bsize, numX, R = bits.shape
    nv, nslice, ns, _, numY, _, C = G.shape
    if not G.is_cuda:
        raise RuntimeError("the converged crossbar model runs on the GPU (G is on the CPU)")
    dev = G.device
    # slot (((v*nslice + j)*ns + s)*numX + x)*numY + y, v the input bit when
    # every bit has its own conductances (variation), else 0
    slots = G.reshape(-1, R, C).to(torch.float32).contiguous()
    I = torch.zeros((nbit, nslice, ns, bsize, numX, numY, C), dtype=torch.float64, device=dev)
    ii, bb, xx, jj, ss, yy = tile_instances(bits, nslice, numY, ns)
    total = ii.numel()
    vv = ii if nv > 1 else torch.zeros_like(ii)
    gidx = ((((vv * nslice + jj) * ns + ss) * numX + xx) * numY + yy).to(torch.int32)
    r = torch.full((1,), float(P.r_wire), dtype=torch.float64, device=dev)
    net = P.warm_start
    k = int(P.ir2s_iters)
    ir2s = {"ir2s_iters": k, "corr_iters": max(2, k), "ir2s_omega": float(P.ir2s_omega),
            "test_first": True} if k > 0 else {}
    stats = P.stats  # accumulates over calls until the caller resets it
    for k in ("solves", "sweeps", "picard_steps", "calls"):
        stats.setdefault(k, 0)
    stats["solves"] += total
    for c0 in range(0, total, P.max_batch):
        c1 = min(total, c0 + P.max_batch)
        n = c1 - c0
        sl = slice(c0, c1)
        vin = bits[ii[sl], bb[sl], xx[sl]].to(torch.float64) * inputV
        gi = gidx[sl]
        rr = r.expand(n)
        # a network that solves in-kernel forms its start inside the solve
        in_kernel = net is not None and net.solves_in_kernel and (R, C) == (net.M, net.N)
        if (R, C) not in SUPPORTED_SHAPES:
            if P.nonlinear_cell_enable:
                raise NotImplementedError(f"converged sinh cells on {R} x {C} tiles (the shapes of the batched "
                                          f"solver: {sorted(SUPPORTED_SHAPES)})")
            res = solve_padded(slots, rr, vin, gidx=gi, tol=P.tol, **ir2s)
            stats["sweeps"] += int(res.sweeps.sum())
        elif P.nonlinear_cell_enable:
            warm_fn = (lambda Gz, rw, v: net.predict(Gz, rw, v)) if net is not None and not in_kernel else None
            res = solve_batch_nonlinear(slots, rr, vin, params=P.nonlinear_params, gidx=gi, tol=P.tol,
                                        warm_fn=warm_fn, net=net if in_kernel else None, **ir2s)
            stats["sweeps"] += int(res.total_sweeps.sum())
            stats["picard_steps"] = max(stats["picard_steps"], res.outer)
        else:
            if in_kernel:
                res = solve_batch(slots, rr, vin, gidx=gi, tol=P.tol, net=net, **ir2s)
            else:
                warm = net.predict(slots.index_select(0, gi.long()), rr, vin) if net is not None else None
                res = solve_batch(slots, rr, vin, warm=warm, gidx=gi, tol=P.tol, **ir2s)
            stats["sweeps"] += int(res.sweeps.sum())
        stats["calls"] += 1
        if int((res.status != 0).sum()) != 0:
            raise RuntimeError("crossbar solve did not converge")
        I[ii[sl], jj[sl], ss[sl], bb[sl], xx[sl], yy[sl]] = res.iout
    if P.record is not None:
        P.record.append(I)
    return I


def _solve_converged_ou(bits, G, inputV, P):
    """OU mode in the converged model: only one OUSize[0] x OUSize[1] block of
    a tile conducts during a read, so each block's column currents are those
    of its whole tile with every other cell off and only the block's rows
    driven -- the state the reference's OU iterations (ir_gs_ou,
    split_arr_matmul.cu:1086-1327, with the wire loads of the block's
    position) converge to.  One converged solve of the masked tiles per
    (row block, column block); the result has one entry per OU row block,
    [..][numX * R / OUSize[0]][numY][Cx] (x-major), as the OU iterations give
    it.  Linear cells."""
    if P.nonlinear_cell_enable:
        raise NotImplementedError("converged OU mode with sinh cells (the cell model of an off cell is not defined)")
    nbit, bsize, numX, R = bits.shape
    nv, nslice, ns, _, numY, _, Cx = G.shape
    our, ouc = P.OUSize
    if R % our or Cx % ouc:
        raise ValueError(f"{R} x {Cx} tiles do not split into {our} x {ouc} OU blocks")
    nrb, ncb = R // our, Cx // ouc
    I = torch.zeros((nbit, nslice, ns, bsize, numX, nrb, numY, Cx), dtype=torch.float64, device=G.device)
    record, P.record = P.record, None
    try:
        for rb in range(nrb):
            rows = slice(rb * our, (rb + 1) * our)
            bits_rb = torch.zeros_like(bits)
            bits_rb[..., rows] = bits[..., rows]
            if not bool(bits_rb.any()):
                continue
            for cb in range(ncb):
                cols = slice(cb * ouc, (cb + 1) * ouc)
                Gm = torch.zeros_like(G)
                Gm[..., rows, cols] = G[..., rows, cols]
                Ib = _solve_converged(bits_rb, Gm, inputV, P)
                I[:, :, :, :, :, rb, :, cols] = Ib[..., cols]
    finally:
        P.record = record
    I = I.reshape(nbit, nslice, ns, bsize, numX * nrb, numY, Cx)
    if P.record is not None:
        P.record.append(I)
    return I


def _solve_reference(bits, G, inputV, P, float_type):
    """The reference's iterations: one hook call per (input bit, weight slice,
    array), as splitArrayArithmetic.py:466-511 (whole arrays), :524-605 (OU
    mode) and :612-701 (reference-column arrays, C + 2 or C + OUSize[1]
    columns) make them -- line-GS (ir_drop_gs / ir_gs_ou) or IR2S (ir2s /
    ir2s_ou)."""
    from . import hook
    nbit, bsize, numX, R = bits.shape
    nv, nslice, ns, _, numY, _, C = G.shape  # C includes reference columns
    dev = G.device
    nrb = numX * (R // P.OUSize[0]) if P.useOUSize else numX
    node_ou = P.useOUSize and P.ir_mode == "reference" and P.ir_drop_GS_method_mode == GS_method.nodeBynode
    I = torch.zeros((nbit, nslice, ns, bsize, nrb, numY, C), dtype=torch.float64, device=dev)
    brk, thr = P.ir_drop_less_than_thresholdBreak_enable, P.ir_drop_less_than_threshold
    nl_on, nl = P.nonlinear_cell_enable, P.nonlinear_params
    for i in range(nbit):
        vin = bits[i].to(float_type).mul(inputV).contiguous()  # [bsize][numX][R]
        for jj in range(nslice):
            for s in range(ns):
                Gh = G[i if nv > 1 else 0, jj, s].permute(0, 2, 1, 3).reshape(numX, R, numY * C)
                Gh = Gh.to(float_type).contiguous()
                it = P.ir_drop_iter_time
                if node_ou:
                    # ir_drop_accurate_GSMethod_OU's nodeBynode form (ir_drop_solve.py:315-360): every
                    # OU block a crossbar of its own with plain wire ends
                    our, ouc = P.OUSize
                    vb = vin.reshape(bsize, nrb, our).contiguous()
                    Gb = Gh.reshape(nrb, our, numY * C).contiguous()
                    out = hook.ir_drop_gs(vb, Gb, bsize, our, ouc, nrb, numY * C // ouc, it, 1.0 / P.r_wire,
                                          P.ir_drop_GS_method_beta_scale, brk, thr, nl_on, nl)
                elif P.ir_mode == "ir2s" and P.useOUSize:
                    out = hook.ir2s_ou(vin, Gh, bsize, R, C, P.OUSize[0], P.OUSize[1], numX, numY, it, P.r_wire,
                                       brk, thr, nl_on, nl)
                elif P.ir_mode == "ir2s":
                    out = hook.ir2s(vin, Gh, bsize, R, C, numX, numY, it, P.r_wire, brk, thr, nl_on, nl)
                elif P.useOUSize:
                    out = hook.ir_gs_ou(vin, Gh, bsize, R, C, P.OUSize[0], P.OUSize[1], numX, numY, it,
                                        1.0 / P.r_wire, P.ir_drop_GS_method_beta_scale, brk, thr, nl_on, nl)
                else:
                    out = hook.ir_drop_gs(vin, Gh, bsize, R, C, numX, numY, it, 1.0 / P.r_wire,
                                          P.ir_drop_GS_method_beta_scale, brk, thr, nl_on, nl)
                I[i, jj, s] = out.view(bsize, nrb, numY, C).to(torch.float64)
    return I


class Linear_sp(nn.Module):
    """modules/linear_split.py:13-78: float training, crossbar inference in eval mode."""

    def __init__(self, in_features, out_features, bias=True,
                 input_bit_width=ConstForSplitArray.input_bit_width,
                 output_bit_width=ConstForSplitArray.output_bit_width,
                 weight_bit_width=ConstForSplitArray.weight_bit_width,
                 dac_bit_width=ConstForSplitArray.dac_bit_width,
                 adc_bit_width=ConstForSplitArray.adc_bit_width,
                 phy_arr_mode=phyArrMode.PN, fp_type=torch.float32, use_fp_bias=True, params=None):
        super().__init__()
        self.inSize, self.outSize = in_features, out_features
        self.hasBias, self.useFpBias = bias, use_fp_bias
        self.inputBits, self.outputBits, self.weightBits = input_bit_width, output_bit_width, weight_bit_width
        self.adcBits, self.dacBits = adc_bit_width, dac_bit_width
        self.phyArrMode, self.fp_type, self.params = phy_arr_mode, fp_type, params
        self.fp_weight = nn.Parameter(torch.empty((out_features, in_features), dtype=fp_type))
        self.fp_bias = nn.Parameter(torch.empty(out_features, dtype=fp_type)) if bias else None
        nn.init.kaiming_uniform_(self.fp_weight, a=math.sqrt(5))
        if self.fp_bias is not None:
            bound = 1 / math.sqrt(in_features)
            nn.init.uniform_(self.fp_bias, -bound, bound)

    def forward(self, input: Tensor) -> Tensor:
        if self.training:
            return F.linear(input, self.fp_weight, self.fp_bias)
        if self.hasBias and not self.useFpBias:
            input = torch.cat((input, input.new_ones(input.size(0), 1)), 1)
            logicArr = torch.cat((self.fp_weight.t(), self.fp_bias.unsqueeze(0)), 0)
        else:
            logicArr = self.fp_weight.t()
        output = splitArr_matmul_nn(input, logicArr, self.inputBits, self.dacBits, self.adcBits,
                                    self.outputBits, self.weightBits, self.phyArrMode, self.fp_type,
                                    params=self.params)
        if self.hasBias and self.useFpBias:
            output = output + self.fp_bias
        return output


class Conv2d_sp(nn.Module):
    """modules/conv_split.py:31-136: float training (F.conv2d); in eval mode
    the input is unfolded (im2col), multiplied through the crossbars by
    splitArr_matmul_nn with the weights as a [Cin*kh*kw, Cout] matrix, the fp
    bias added after the crossbars (or mapped as an extra weight row with
    use_fp_bias=False), and the result folded back to NCHW.  Zero padding,
    groups = 1 in eval mode (the reference's eval path has the same limits)."""

    def __init__(self, in_channels, out_channels, kernel_size, stride=1, padding=0, dilation=1, groups=1,
                 bias=True, padding_mode="zeros",
                 input_bit_width=ConstForSplitArray.input_bit_width,
                 output_bit_width=ConstForSplitArray.output_bit_width,
                 weight_bit_width=ConstForSplitArray.weight_bit_width,
                 dac_bit_width=ConstForSplitArray.dac_bit_width,
                 adc_bit_width=ConstForSplitArray.adc_bit_width,
                 phy_arr_mode=phyArrMode.PN, fp_type=torch.float32, use_fp_bias=True, params=None):
        super().__init__()
        if groups <= 0:
            raise ValueError("groups must be a positive integer")
        if in_channels % groups != 0:
            raise ValueError("in_channels must be divisible by groups")
        if out_channels % groups != 0:
            raise ValueError("out_channels must be divisible by groups")
        pair = lambda v: tuple(v) if isinstance(v, (tuple, list)) else (v, v)  # noqa: E731
        self.in_channels, self.out_channels = in_channels, out_channels
        self.kernel_size, self.stride = pair(kernel_size), pair(stride)
        self.padding, self.dilation, self.groups = pair(padding), pair(dilation), groups
        self.hasBias, self.padding_mode, self.useFpBias = bias, padding_mode, use_fp_bias
        self.inputBits, self.outputBits, self.weightBits = input_bit_width, output_bit_width, weight_bit_width
        self.adcBits, self.dacBits = adc_bit_width, dac_bit_width
        self.phyArrMode, self.fp_type, self.params = phy_arr_mode, fp_type, params
        self.fp_weight = nn.Parameter(torch.empty((out_channels, in_channels // groups, *self.kernel_size),
                                                  dtype=fp_type))
        self.fp_bias = nn.Parameter(torch.empty(out_channels, dtype=fp_type)) if bias else None
        nn.init.kaiming_uniform_(self.fp_weight, a=math.sqrt(5))
        if self.fp_bias is not None:
            fan_in = (in_channels // groups) * self.kernel_size[0] * self.kernel_size[1]
            bound = 1 / math.sqrt(fan_in)
            nn.init.uniform_(self.fp_bias, -bound, bound)

    def forward(self, input: Tensor) -> Tensor:
        if self.training:
            return F.conv2d(input, self.fp_weight, self.fp_bias, self.stride, self.padding, self.dilation,
                            self.groups)
        if self.groups != 1 or self.padding_mode != "zeros":
            raise NotImplementedError("crossbar inference of Conv2d_sp: groups = 1, zero padding")
        B, _, H, W = input.shape
        kh, kw = self.kernel_size
        cols = F.unfold(input, self.kernel_size, dilation=self.dilation, padding=self.padding, stride=self.stride)
        rows = cols.transpose(1, 2).reshape(-1, self.in_channels * kh * kw)
        wmat = self.fp_weight.reshape(self.out_channels, -1)
        if self.hasBias and not self.useFpBias:  # bias mapped onto the arrays as a ones row
            rows = torch.cat((rows, rows.new_ones(rows.size(0), 1)), 1)
            logicArr = torch.cat((wmat.t(), self.fp_bias.unsqueeze(0)), 0)
        else:
            logicArr = wmat.t()
        out = splitArr_matmul_nn(rows, logicArr, self.inputBits, self.dacBits, self.adcBits, self.outputBits,
                                 self.weightBits, self.phyArrMode, self.fp_type, params=self.params)
        if self.hasBias and self.useFpBias:
            out = out + self.fp_bias
        Ho = (H + 2 * self.padding[0] - self.dilation[0] * (kh - 1) - 1) // self.stride[0] + 1
        Wo = (W + 2 * self.padding[1] - self.dilation[1] * (kw - 1) - 1) // self.stride[1] + 1
        return out.reshape(B, Ho * Wo, self.out_channels).transpose(1, 2).reshape(B, self.out_channels, Ho, Wo)


class MLP(nn.Module):
    """Perceptron of Linear_sp layers (784-256-10 by default) with ReLU between
    them and seeded Gaussian weights."""

    def __init__(self, sizes=(784, 256, 10), params=None, seed=0):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.layers = nn.ModuleList(Linear_sp(a, b, params=params) for a, b in zip(sizes[:-1], sizes[1:]))
        with torch.no_grad():
            for L in self.layers:
                L.fp_weight.copy_(torch.randn(L.fp_weight.shape, generator=gen) / math.sqrt(L.inSize))
                L.fp_bias.copy_(torch.randn(L.fp_bias.shape, generator=gen) * 0.01)

    def forward(self, x):
        for k, L in enumerate(self.layers):
            x = L(x)
            if k < len(self.layers) - 1:
                x = torch.relu(x)
        return x
