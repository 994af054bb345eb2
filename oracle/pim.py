"""CPU restatement of PimTorch's bit-sliced crossbar MVM (numpy).

TEST INFRASTRUCTURE ONLY: the checker of fadesim_amd.pim.  The crossbar
currents come from a caller-supplied function (the exact network solve of
oracle.exact_slots, or the ideal crossbar), so this module restates only the
integer/fixed-point pipeline around the solver:

* ``quantize``       -- fixedPointArithmetic.py:19-27 (_get_fixed_point_position)
                        and :76-96 (quantization_tensor): s = ceil(log2(max/q)),
                        int = round(x / 2^s), round half to even.
* ``split_matmul``   -- splitArrayArithmetic.py:1161-1260 (splitArr_matmul_nn,
                        PN arrays, complement-code inputs, 1-bit DAC) with
                        matmul :442-1018 for the whole-array PN branch:
                        pad rows/cols to the tile size (:716-735), inputs
                        bit * inputV (:750-751), g = slice * dG + Gmin in the
                        float type (:753-754), ADC round(I / adcUseI * (2^adc-1))
                        per tile (:833-835), sum over row tiles, [:outsize],
                        * rows (2^dac-1)(2^cell-1)/(2^adc-1) (:958-968),
                        out_pos - out_neg, then +-out * 2^(i+j+s1+s2) with the
                        input MSB subtracted (:1224-1227).  OU mode
                        (``ou_rows``): currents per OU row block, ADC full
                        scale ou_rows (Gmax-Gmin) vmax (commonConst.py:242,
                        matmul :838-840), rescale by ou_rows (:968-976).
* ``ideal_currents_ou`` -- the ideal per-OU product (no_ir_drop_cal_ou,
                        ir_drop_solve.py:21-30).

Every array op is done in float32 where the reference's is (fp_type float32).
"""
from __future__ import annotations

import math

import numpy as np


def fixed_point_position(max_abs: float, bits: int, kind: str) -> int:
    """fixedPointArithmetic.py:19-27."""
    if math.isclose(max_abs, 0.0):
        max_abs = 1e-12
    if kind == "normal":
        return math.ceil(math.log2(max_abs / ((1 << (bits - 1)) - 1)))
    if kind == "pn":
        return math.ceil(math.log2(max_abs / ((1 << bits) - 1)))
    raise ValueError(kind)


def quantize(x: np.ndarray, bits: int, kind: str):
    """fixedPointArithmetic.py:76-96 on float32 data: (int64 array, s)."""
    x = np.asarray(x, dtype=np.float32)
    s = fixed_point_position(float(np.abs(x).max()), bits, kind)
    q = np.round(x / np.float32(2.0 ** s))  # division by a power of two is exact
    return q.astype(np.int64), s


def ideal_currents(vin: np.ndarray, G: np.ndarray) -> np.ndarray:
    """no_ir_drop_cal (ir_drop_solve.py:12-15): vin [bsize, numX, R] x
    G [numX, R, numY*C] -> [bsize, numX, numY*C], FP64."""
    return np.einsum("bxr,xrc->bxc", vin.astype(np.float64), G.astype(np.float64))


def ideal_currents_ou(ou_rows: int):
    """vin [bsize, numX, R] x G [numX, R, numY*C] -> per OU row block
    [bsize, numX * R/ou_rows, numY*C], FP64."""
    def cur(vin, G):
        bsize, numX, R = vin.shape
        k = R // ou_rows
        v = vin.astype(np.float64).reshape(bsize, numX, k, ou_rows)
        g = G.astype(np.float64).reshape(numX, k, ou_rows, G.shape[2])
        return np.einsum("bxkr,xkrc->bxkc", v, g).reshape(bsize, numX * k, G.shape[2])
    return cur


def split_matmul(x, w, *, input_bits=8, dac_bits=1, adc_bits=10, weight_bits=8, cell_bits=1,
                 rows=64, cols=64, vmax=1.0, ron=1e3, roff=1e6, currents=ideal_currents, ou_rows=None):
    """output [bsize, outsize] = x [bsize, insize] @ w [insize, outsize] through
    PN crossbars.  ``currents(vin [bsize,numX,R] f32, G [numX,R,numY*C] f32)``
    returns the tile currents [bsize, numX, numY*C].  Returns (output f32,
    dict of intermediates: codes per (bit, slice, sign), ADC inputs)."""
    if dac_bits != 1:
        raise ValueError("complement-code inputs need a 1-bit DAC")
    if weight_bits % cell_bits:
        raise ValueError("cell_bits must divide weight_bits")
    x = np.asarray(x, np.float32)
    w = np.asarray(w, np.float32)
    bsize, insize = x.shape
    outsize = w.shape[1]
    qin, s1 = quantize(x, input_bits, "normal")
    qw, s2 = quantize(w, weight_bits, "pn")
    wpos = np.where(qw > 0, qw, 0)
    wneg = np.abs(np.where(qw < 0, qw, 0))
    numX, numY = -(-insize // rows), -(-outsize // cols)
    Rp, Cp = numX * rows, numY * cols
    gmin, gmax = 1.0 / roff, 1.0 / ron
    dG = (gmax - gmin) / (2 ** cell_bits - 1)
    per_read = ou_rows if ou_rows else rows  # rows behind one ADC reading
    adc_use_i = per_read * (gmax - gmin) * vmax
    input_v = vmax / (2 ** dac_bits - 1)
    mask_cell = (1 << cell_bits) - 1
    lsb = np.float32(2 ** adc_bits - 1)
    rescale = np.float32(per_read * (2 ** dac_bits - 1) * (2 ** cell_bits - 1) / (2 ** adc_bits - 1))
    output = np.zeros((bsize, outsize), np.float32)
    codes, adc_in = {}, {}
    for i in range(0, input_bits, dac_bits):
        in_b = (qin >> i) & 1
        in_b = np.pad(in_b, ((0, 0), (0, Rp - insize)))
        vin = (in_b.astype(np.float32) * np.float32(input_v)).reshape(bsize, numX, rows)
        for j in range(0, weight_bits, cell_bits):
            tot = []
            for sign, wm in (("pos", wpos), ("neg", wneg)):
                sl = (wm >> j) & mask_cell
                sl = np.pad(sl, ((0, Rp - insize), (0, Cp - outsize)))
                G = (sl.astype(np.float32) * np.float32(dG) + np.float32(gmin)).reshape(numX, rows, Cp)
                I = np.asarray(currents(vin, G)).astype(np.float32)
                a = I / np.float32(adc_use_i) * lsb
                code = np.round(a)
                codes[(i, j, sign)] = code
                adc_in[(i, j, sign)] = a
                tot.append(code.sum(1, dtype=np.float32)[:, :outsize] * rescale)
            out = tot[0] - tot[1]
            term = out * np.float32(2.0 ** (i + j + s1 + s2))
            output = output - term if i == input_bits - 1 else output + term
    return output, {"codes": codes, "adc_in": adc_in, "s1": s1, "s2": s2}
