"""PimTorch crossbar MVM path on the GPU: the caller side of the IR-drop hook.

Restates the fixed-point, bit-sliced crossbar matrix multiply of PimTorch --
``quantization_tensor`` (pimtorch/pimfixedpoint/fixedPoint/nn/fixedPointArithmetic.py:19-96),
``splitArr_matmul_nn`` (splitArrayArithmetic.py:1161-1260), ``matmul`` /
``IRdrop_matmul`` (splitArrayArithmetic.py:442-1018, PN arrays, whole-array
mode) and ``Linear_sp`` (modules/linear_split.py:13-78) -- with the same
arithmetic: Normal (two's complement) quantized inputs applied bit-serially,
PN weights split into positive/negative magnitudes and sliced ``cellBits``
per cell, conductance ``slice*dG + Gmin``, per-tile ADC
``round(I/adcUseI*(2^adc-1))``, sum over row tiles, rescale and shift-add with
the MSB subtracted (complement code).

MI355X-first difference: instead of one hook call per (input bit, weight
slice, pos/neg) -- 2 * inputBits * weightBits/cellBits calls per layer -- every
tile solve of the layer goes into ONE batched solve: the distinct conductance
tiles become slots of ``solve_batch(..., gidx=...)`` and every (bit, slice,
sign, image, row tile, column tile) is an instance (all-zero input vectors
are skipped: their current is exactly 0).  ``ir_mode`` selects the solver:

* ``"converged"`` -- fadesim_solve_batch to ``tol`` (optionally NN-warm-started);
* ``"reference"`` -- the reference's own line-GS iteration through
  fadesim_amd.hook.ir_drop_gs (iter_time sweeps at beta), per call like PimTorch;
* ``"none"``      -- ideal crossbar (``no_ir_drop``).
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


class ConstForSplitArray:
    input_bit_width = 8
    output_bit_width = 10
    weight_bit_width = 8
    adc_bit_width = 10
    dac_bit_width = 1


class PhyArrParams:
    """The crossbar constants of PimTorch's phyArrParams (commonConst.py:187-325)
    that the MVM path reads.  Derived values are properties so that changing
    Ron/Roff/cellBits/arrRowSize keeps them consistent."""

    def __init__(self, **kw):
        self.Ron = 1e3
        self.Roff = 1e6
        self.sigma = 0.0
        self.cellBits = 1
        self.arrRowSize = 64
        self.arrColSize = 64
        self.inputMaxVoltage = 1.0
        self.r_wire = 2.93
        self.ir_mode = "reference"            # converged | reference | none
        self.ir_drop_iter_time = 600
        self.ir_drop_GS_method_beta_scale = 1.90
        self.ir_drop_less_than_thresholdBreak_enable = False
        self.ir_drop_less_than_threshold = 1e-7
        self.nonlinear_cell_enable = False
        self.nonlinear_params = []
        self.tol = 1e-6                       # converged mode
        self.warm_start = None                # WarmStartNet or None (converged mode)
        self.max_batch = 131072               # instances per solve call
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


phyArrParams = PhyArrParams()


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
    through the simulated crossbars (PN arrays, complement-code inputs, 1-bit DAC)."""
    P = params or phyArrParams
    if arrType != phyArrMode.PN:
        raise NotImplementedError("only positive/negative (PN) arrays are supported")
    if dacBits != 1:
        raise Exception("[Error] Complement code input mode, but dac is not 1 bit. dacbits = {}".format(dacBits))
    if weightBits % P.cellBits != 0:
        raise Exception("[Error] cellbit is not factor of weight_bit")
    with torch.no_grad():
        dev = w.device
        bsize, insize = input.shape
        outsize = w.shape[1]
        qin, s1 = quantization_tensor(input, inputBits, TensorType.Normal)
        qw, s2 = quantization_tensor(w, weightBits, TensorType.PN)
        wpos = qw.clamp(min=0)
        wneg = (-qw).clamp(min=0)
        R, C = P.arrRowSize, P.arrColSize
        numX, numY = _tiles(insize, R), _tiles(outsize, C)
        Rp, Cp = numX * R, numY * C
        nslice = weightBits // P.cellBits
        mask_cell = (1 << P.cellBits) - 1
        inputV = P.inputMaxVoltage / (2 ** dacBits - 1)
        # conductance slots [nslice][2][numX][numY][R][C] (linear map of the slice value)
        slices = torch.stack([torch.stack([(wpos >> (P.cellBits * j)) & mask_cell,
                                           (wneg >> (P.cellBits * j)) & mask_cell]) for j in range(nslice)])
        slices = F.pad(slices, (0, Cp - outsize, 0, Rp - insize))  # padded cells hold slice 0
        G = slices.to(torch.float64) * P.cellDeltaConduct + P.cellMinConduct
        G = G.view(nslice, 2, numX, R, numY, C).permute(0, 1, 2, 4, 3, 5).contiguous()
        # input bits [inputBits][bsize][numX][R] (two's complement bits of qin)
        qpad = F.pad(qin, (0, Rp - insize))
        bits = torch.stack([(qpad >> i) & 1 for i in range(inputBits)]).view(inputBits, bsize, numX, R)
        # currents [inputBits][nslice][2][bsize][numX][numY][C]
        if P.ir_mode == "none":
            I = torch.einsum("ibxr,jsxyrc->ijsbxyc", bits.to(torch.float64) * inputV, G)
        elif P.ir_mode == "converged":
            I = _solve_converged(bits, G, inputV, P)
        elif P.ir_mode == "reference":
            I = _solve_reference(bits, G, inputV, P, float_type)
        else:
            raise ValueError(f"unknown ir_mode {P.ir_mode}")
        # ADC per tile (splitArrayArithmetic.py:831-836), sum over row tiles, rescale (:959-968)
        I = I.to(float_type)
        adc = I.div(P.adcUseI).mul(2 ** adcBits - 1).round()
        out = adc.sum(4).reshape(inputBits, nslice, 2, bsize, Cp)[..., :outsize]
        out = out.mul(R * (2 ** dacBits - 1) * (2 ** P.cellBits - 1) / (2 ** adcBits - 1))
        out = out[:, :, 0] - out[:, :, 1]  # [inputBits][nslice][bsize][outsize]
        output = w.new_zeros((bsize, outsize))
        for i in range(inputBits):
            for jj in range(nslice):
                term = out[i, jj].to(w.dtype).mul(2.0 ** (i + P.cellBits * jj + s1 + s2))
                output = output.sub(term) if i == inputBits - 1 else output.add(term)
        return output


def _solve_converged(bits, G, inputV, P):
    """All tile solves of the layer in batched calls of fadesim_solve_batch."""
    from .solver import solve_batch
    nbit, bsize, numX, R = bits.shape
    nslice, _, _, numY, _, C = G.shape
    dev = G.device
    slots = G.reshape(-1, R, C).to(torch.float32).contiguous()  # slot id = ((j*2+s)*numX + x)*numY + y
    I = torch.zeros((nbit, nslice, 2, bsize, numX, numY, C), dtype=torch.float64, device=dev)
    nz = bits.view(nbit, bsize, numX, R).any(-1)  # non-zero input vectors
    ii, bb, xx = torch.nonzero(nz, as_tuple=True)
    if ii.numel() == 0:
        return I
    # instances: every non-zero input vector x every (slice, sign, column tile)
    j = torch.arange(nslice, device=dev)
    sgn = torch.arange(2, device=dev)
    y = torch.arange(numY, device=dev)
    J, S, Y = torch.meshgrid(j, sgn, y, indexing="ij")
    J, S, Y = J.reshape(-1), S.reshape(-1), Y.reshape(-1)
    n_in, n_t = ii.numel(), J.numel()
    inst_i = ii.repeat_interleave(n_t)
    inst_b = bb.repeat_interleave(n_t)
    inst_x = xx.repeat_interleave(n_t)
    inst_j, inst_s, inst_y = J.repeat(n_in), S.repeat(n_in), Y.repeat(n_in)
    gidx = (((inst_j * 2 + inst_s) * numX + inst_x) * numY + inst_y).to(torch.int32)
    total = gidx.numel()
    r = torch.full((1,), float(P.r_wire), dtype=torch.float64, device=dev)
    for c0 in range(0, total, P.max_batch):
        c1 = min(total, c0 + P.max_batch)
        vin = bits[inst_i[c0:c1], inst_b[c0:c1], inst_x[c0:c1]].to(torch.float64) * inputV
        gi = gidx[c0:c1]
        warm = None
        if P.warm_start is not None:
            warm = P.warm_start.predict(slots.index_select(0, gi.long()), r.expand(c1 - c0), vin)
        res = solve_batch(slots, r.expand(c1 - c0), vin, warm=warm, gidx=gi, tol=P.tol)
        if int((res.status != 0).sum()) != 0:
            raise RuntimeError("crossbar solve did not converge")
        I[inst_i[c0:c1], inst_j[c0:c1], inst_s[c0:c1], inst_b[c0:c1], inst_x[c0:c1], inst_y[c0:c1]] = res.iout
    return I


def _solve_reference(bits, G, inputV, P, float_type):
    """The reference's iteration: one hook call per (input bit, weight slice,
    sign), as splitArrayArithmetic.py:488-511 does."""
    from .hook import ir_drop_gs
    nbit, bsize, numX, R = bits.shape
    nslice, _, _, numY, _, C = G.shape
    dev = G.device
    I = torch.zeros((nbit, nslice, 2, bsize, numX, numY, C), dtype=torch.float64, device=dev)
    for i in range(nbit):
        vin = (bits[i].to(float_type) * inputV).contiguous()  # [bsize][numX][R]
        for jj in range(nslice):
            for s in range(2):
                Gh = G[jj, s].permute(0, 2, 1, 3).reshape(numX, R, numY * C).to(float_type).contiguous()
                out = ir_drop_gs(vin, Gh, bsize, R, C, numX, numY, P.ir_drop_iter_time, 1.0 / P.r_wire,
                                 P.ir_drop_GS_method_beta_scale, P.ir_drop_less_than_thresholdBreak_enable,
                                 P.ir_drop_less_than_threshold, P.nonlinear_cell_enable, P.nonlinear_params)
                I[i, jj, s] = out.view(bsize, numX, numY, C).to(torch.float64)
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
