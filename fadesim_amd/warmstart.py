"""Learned warm start for the batched IR-drop solver.

Pipeline (all on the GPU, every kernel from include/fadesim_nn.h):

1. physics features of each crossbar: first- and second-order
   superposition drop fields (exact prefix-sum evaluations of the wire
   Green's functions), conductance, inputs, wire resistance -- 8 channels;
2. a BF16 4-level convolutional encoder-decoder (U-Net with additive skips,
   3x3 convolutions as MFMA implicit GEMMs, FP32 accumulation, ~340k
   parameters) predicting the remaining correction of the wordline and
   bitline drop fields, in units of the second-order term;
3. warm = second-order field + scale * correction, handed to
   ``solve_batch(..., warm=...)``; the solver's convergence test judges it
   like any other initial guess and refines it to the same tolerance.

Training data are solutions of ``solve_batch`` only (field output at a tight
tolerance); the network is trained by the project's own forward and backward
kernels with Adam on FP32 master weights.  Weights are stored as raw
little-endian FP32 plus a JSON manifest in fadesim_amd/weights/.

There is no counterpart in the reference; its related learned component is
the SCN effective-conductance predictor (pimtorch/pimfixedpoint/SCN.py:127-153,
trained on solver output in SCN_data_gen.py:134-179).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import torch

from . import _native as nat

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_NAME = "warmstart_128"
N_FEAT = 8
OUT_C = 16  # only channels 0 (wordline) and 1 (bitline) are used


def kpad(cin: int) -> int:
    """Padded GEMM depth: 9*Cin weights, one bias column, up to a multiple of 32."""
    return ((9 * cin + 1 + 31) // 32) * 32


@dataclass
class Op:
    name: str
    kind: str        # conv | pool | up_add
    src: str
    level: int       # resolution level of the output (0 = full)
    cin: int = 0
    cout: int = 0
    relu: bool = True
    skip: str = ""   # up_add: tensor added after upsampling


def _unet16() -> List[Op]:
    C = [16, 32, 64, 96]
    return [
        Op("e0a", "conv", "x", 0, N_FEAT, C[0]),
        Op("e0", "conv", "e0a", 0, C[0], C[0]),
        Op("p1", "pool", "e0", 1, C[0], C[0]),
        Op("e1a", "conv", "p1", 1, C[0], C[1]),
        Op("e1", "conv", "e1a", 1, C[1], C[1]),
        Op("p2", "pool", "e1", 2, C[1], C[1]),
        Op("e2a", "conv", "p2", 2, C[1], C[2]),
        Op("e2", "conv", "e2a", 2, C[2], C[2]),
        Op("p3", "pool", "e2", 3, C[2], C[2]),
        Op("e3a", "conv", "p3", 3, C[2], C[3]),
        Op("e3", "conv", "e3a", 3, C[3], C[3]),
        Op("t3", "conv", "e3", 3, C[3], C[2]),
        Op("u2", "up_add", "t3", 2, C[2], C[2], skip="e2"),
        Op("d2", "conv", "u2", 2, C[2], C[2]),
        Op("t2", "conv", "d2", 2, C[2], C[1]),
        Op("u1", "up_add", "t2", 1, C[1], C[1], skip="e1"),
        Op("d1", "conv", "u1", 1, C[1], C[1]),
        Op("t1", "conv", "d1", 1, C[1], C[0]),
        Op("u0", "up_add", "t1", 0, C[0], C[0], skip="e0"),
        Op("d0", "conv", "u0", 0, C[0], C[0]),
        Op("out", "conv", "d0", 0, C[0], OUT_C, relu=False),
    ]


def _slim() -> List[Op]:
    """Two full-resolution convolutions (input and output), 16/32/64/64
    channels: about half the multiply-adds of unet16."""
    return [
        Op("e0", "conv", "x", 0, N_FEAT, 16),
        Op("p1", "pool", "e0", 1, 16, 16),
        Op("e1a", "conv", "p1", 1, 16, 32),
        Op("e1", "conv", "e1a", 1, 32, 32),
        Op("p2", "pool", "e1", 2, 32, 32),
        Op("e2a", "conv", "p2", 2, 32, 64),
        Op("e2", "conv", "e2a", 2, 64, 64),
        Op("p3", "pool", "e2", 3, 64, 64),
        Op("e3", "conv", "p3", 3, 64, 64),
        Op("t3", "conv", "e3", 3, 64, 64),
        Op("u2", "up_add", "t3", 2, 64, 64, skip="e2"),
        Op("t2", "conv", "u2", 2, 64, 32),
        Op("u1", "up_add", "t2", 1, 32, 32, skip="e1"),
        Op("t1", "conv", "u1", 1, 32, 16),
        Op("u0", "up_add", "t1", 0, 16, 16, skip="e0"),
        Op("out", "conv", "u0", 0, 16, OUT_C, relu=False),
    ]


def _coarse32() -> List[Op]:
    """The encoder-decoder moved to a 4x coarser grid: the features are
    average-pooled to 32x32 and the network (32 / 64 channels at 32x32 and
    16x16) predicts the correction there; predict() upsamples it bilinearly
    (fadesim_nn_combine_up).  The correction is smooth on the scale of a few
    cells, and a wire-solve iteration on the warm field restores the per-cell
    detail: no full-resolution convolution, ~1/8 of unet16's multiply-adds."""
    return [
        Op("p1", "pool", "x", 1, N_FEAT, N_FEAT),
        Op("p2", "pool", "p1", 2, N_FEAT, N_FEAT),
        Op("e2a", "conv", "p2", 2, N_FEAT, 32),
        Op("e2", "conv", "e2a", 2, 32, 32),
        Op("p3", "pool", "e2", 3, 32, 32),
        Op("e3a", "conv", "p3", 3, 32, 64),
        Op("e3", "conv", "e3a", 3, 64, 64),
        Op("t3", "conv", "e3", 3, 64, 32),
        Op("u2", "up_add", "t3", 2, 32, 32, skip="e2"),
        Op("d2", "conv", "u2", 2, 32, 32),
        Op("out", "conv", "d2", 2, 32, OUT_C, relu=False),
    ]


def _coarse32f() -> List[Op]:
    """coarse32 fed directly with features averaged over 4x4 cells
    (fadesim_nn_features_pooled computes them in registers, with no
    full-resolution feature planes and no pooling passes).  Same convolutions,
    so coarse32 weights load into it."""
    return [op if op.name != "e2a" else Op("e2a", "conv", "x", 2, N_FEAT, 32)
            for op in _coarse32() if op.kind != "pool" or op.name == "p3"]


def _coarse16f() -> List[Op]:
    """coarse32f at half the width (16 / 32 channels): a quarter of its
    multiply-adds."""
    return [
        Op("e2a", "conv", "x", 2, N_FEAT, 16),
        Op("e2", "conv", "e2a", 2, 16, 16),
        Op("p3", "pool", "e2", 3, 16, 16),
        Op("e3a", "conv", "p3", 3, 16, 32),
        Op("e3", "conv", "e3a", 3, 32, 32),
        Op("t3", "conv", "e3", 3, 32, 16),
        Op("u2", "up_add", "t3", 2, 16, 16, skip="e2"),
        Op("d2", "conv", "u2", 2, 16, 16),
        Op("out", "conv", "d2", 2, 16, OUT_C, relu=False),
    ]


ARCHITECTURES = {"unet16": _unet16, "slim": _slim, "coarse32": _coarse32, "coarse32f": _coarse32f,
                 "coarse16f": _coarse16f}


def architecture(name: str = "unet16") -> List[Op]:
    return ARCHITECTURES[name]()


def _s(dev) -> ctypes.c_void_p:
    return ctypes.c_void_p(nat.stream_handle(dev))


def _chk(rc, what):
    nat.check(rc, what)


class WarmStartNet:
    """Parameters, forward, backward and Adam of the warm-start network."""

    def __init__(self, device, M: int = 128, N: int = 128, seed: int = 0, arch: str = "unet16"):
        self.dev = torch.device(device)
        self.M, self.N = M, N
        self.arch = arch
        self.ops = architecture(arch)
        self.offsets: Dict[str, int] = {}
        n = 0
        for op in self.ops:
            if op.kind == "conv":
                self.offsets[op.name] = n
                n += op.cout * kpad(op.cin)
        self.nparams_padded = n
        self.w = torch.zeros(n, dtype=torch.float32, device=self.dev)
        g = torch.Generator(device="cpu").manual_seed(seed)
        for op in self.ops:
            if op.kind != "conv":
                continue
            fan_in = 9 * op.cin
            std = math.sqrt(2.0 / fan_in) if op.relu else 0.1 * math.sqrt(1.0 / fan_in)
            wl = torch.zeros(op.cout, kpad(op.cin))
            wl[:, :9 * op.cin] = torch.randn(op.cout, 9 * op.cin, generator=g) * std
            self._view(self.w, op).copy_(wl.to(self.dev))
        self.wb = self.w.to(torch.bfloat16)
        self._wver = 0  # bumped whenever the bf16 weights change (caches derived layouts)
        self.info: dict = {}
        # tensors whose gradient the backward pass needs: outputs of ops that
        # depend on a convolution (pools of the input features do not)
        self._grad_tensors = set()
        for op in self.ops:
            if op.kind == "conv" or op.src in self._grad_tensors or op.skip in self._grad_tensors:
                self._grad_tensors.add(op.name)

    # ------------------------------------------------------------------ params
    def _view(self, flat, op):
        o = self.offsets[op.name]
        return flat[o:o + op.cout * kpad(op.cin)].view(op.cout, kpad(op.cin))

    @property
    def in_level(self) -> int:
        """Resolution level of the input features (0 = the crossbar's)."""
        return min(op.level for op in self.ops if op.src == "x" and op.kind == "conv") if not any(
            op.src == "x" and op.kind == "pool" for op in self.ops) else 0

    @property
    def out_level(self) -> int:
        """Resolution level of the network output (0 = the crossbar's)."""
        return next(op.level for op in self.ops if op.name == "out")

    def n_params(self) -> int:
        """Trainable parameters (weights and biases, padding excluded)."""
        return sum(op.cout * (9 * op.cin + 1) for op in self.ops if op.kind == "conv")

    def describe(self) -> dict:
        return {"arch": self.arch, "params": self.n_params(), "layers": sum(op.kind == "conv" for op in self.ops),
                "dtype": "bf16 (fp32 accumulate)", **{k: v for k, v in self.info.items() if k != "loss_curve"}}

    def save(self, name: str = DEFAULT_NAME, directory: str = WEIGHTS_DIR):
        os.makedirs(directory, exist_ok=True)
        self.w.detach().cpu().numpy().astype("<f4").tofile(os.path.join(directory, name + ".bin"))
        man = {"format": "raw little-endian float32, layers in order, each [Cout][Kp] with k = tap*Cin + ci, "
                         "bias at k = 9*Cin",
               "arch": self.arch, "rows": self.M, "cols": self.N, "features": N_FEAT, "out_channels": OUT_C,
               "layers": [{"name": op.name, "kind": op.kind, "src": op.src, "level": op.level, "cin": op.cin,
                           "cout": op.cout, "relu": op.relu, "skip": op.skip,
                           "offset": self.offsets.get(op.name), "kp": kpad(op.cin) if op.kind == "conv" else None}
                          for op in self.ops],
               "padded_floats": self.nparams_padded, "params": self.n_params(), "info": self.info}
        with open(os.path.join(directory, name + ".json"), "w") as f:
            json.dump(man, f, indent=1)

    @classmethod
    def load(cls, device, name: str = DEFAULT_NAME, directory: str = WEIGHTS_DIR) -> "WarmStartNet":
        with open(os.path.join(directory, name + ".json")) as f:
            man = json.load(f)
        net = cls(device, man["rows"], man["cols"], arch=man.get("arch", "unet16"))
        import numpy as np
        w = np.fromfile(os.path.join(directory, name + ".bin"), dtype="<f4")
        if w.size != net.nparams_padded:
            raise RuntimeError(f"weight file has {w.size} floats, architecture needs {net.nparams_padded}")
        net.w.copy_(torch.from_numpy(w).to(net.dev))
        net.wb = net.w.to(torch.bfloat16)
        net._wver += 1
        net.info = man.get("info", {})
        return net

    # ------------------------------------------------------------------ features
    def features(self, g, r, vin, base=None, scale=None, want_base=True):
        """Input features at the network's input level, the second-order base
        field [B][2][M][N] and the per-instance output scale (written into
        ``base`` / ``scale`` when given).  want_base=False (pooled features
        only): no base field is written and None is returned in its place."""
        B = g.shape[0]
        lv = self.in_level
        H, W = self._hw(lv)
        feat = torch.empty((B, H, W, N_FEAT), dtype=torch.bfloat16, device=self.dev)
        if not want_base and lv == 0:
            raise ValueError("the full-resolution features kernel always writes the base field")
        if base is None and want_base:
            base = torch.empty((B, 2, self.M, self.N), dtype=torch.float32, device=self.dev)
        if scale is None:
            scale = torch.empty(B, dtype=torch.float32, device=self.dev)
        ptrs = [ctypes.c_void_p(x.data_ptr() if x is not None else None) for x in (g, r, vin, feat, base, scale)]
        if lv == 0:
            _chk(nat.lib().fadesim_nn_features(B, self.M, self.N, *ptrs, _s(self.dev)), "nn_features")
        else:
            _chk(nat.lib().fadesim_nn_features_pooled(B, self.M, self.N, lv, *ptrs, _s(self.dev)),
                 "nn_features_pooled")
        return feat, base, scale

    # ------------------------------------------------------------------ forward
    def _hw(self, level):
        return self.M >> level, self.N >> level

    def forward(self, x, keep=False):
        """x: features [B][H][W][8] bf16 at the input level -> out [B][H'][W'][16]
        bf16 at the output level (and all activations)."""
        B = x.shape[0]
        L = nat.lib()
        acts = {"x": x}
        s = _s(self.dev)
        for op in self.ops:
            H, W = self._hw(op.level)
            if op.kind == "conv":
                y = torch.empty((B, H, W, op.cout), dtype=torch.bfloat16, device=self.dev)
                wv = self._view(self.wb, op)
                _chk(L.fadesim_nn_conv3x3_fwd(B, H, W, op.cin, op.cout, kpad(op.cin),
                                              ctypes.c_void_p(acts[op.src].data_ptr()), ctypes.c_void_p(wv.data_ptr()),
                                              ctypes.c_void_p(y.data_ptr()), int(op.relu), s), op.name)
            elif op.kind == "pool":
                y = torch.empty((B, H, W, op.cout), dtype=torch.bfloat16, device=self.dev)
                _chk(L.fadesim_nn_pool2(B, 2 * H, 2 * W, op.cout, ctypes.c_void_p(acts[op.src].data_ptr()),
                                        ctypes.c_void_p(y.data_ptr()), s), op.name)
            else:  # up_add
                y = torch.empty((B, H, W, op.cout), dtype=torch.bfloat16, device=self.dev)
                _chk(L.fadesim_nn_up2_add(B, H // 2, W // 2, op.cout, ctypes.c_void_p(acts[op.src].data_ptr()),
                                          ctypes.c_void_p(acts[op.skip].data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                          s), op.name)
            acts[op.name] = y
            if not keep:
                for k in list(acts):
                    if k not in ("x",) and k != op.name and not self._needed_later(k, op.name):
                        del acts[k]
        return acts if keep else acts["out"]

    FUSED_LAYERS = ("e2a", "e2", "e3a", "e3", "t3", "d2", "out")

    @property
    def fused(self) -> bool:
        """coarse32f and coarse16f at 128x128 run as one kernel
        (fadesim_nn_coarse_forward)."""
        return self.arch in ("coarse32f", "coarse16f") and (self.M, self.N) == (128, 128)

    def fused_weights(self):
        """The fused network's weights as the kernels read them: 7 layer
        pointers (ctypes array) to fragment-major [Kp/32][Cout][32] bf16 copies
        (a wave's fragment is 1 KB contiguous) and their depths Kp, refreshed
        when the weights change."""
        if getattr(self, "_frag_ver", None) != self._wver:
            byname = {op.name: op for op in self.ops}
            self._frag = [self._view(self.wb, byname[n]).view(byname[n].cout, -1, 32).transpose(0, 1).contiguous()
                          for n in self.FUSED_LAYERS]
            self._frag_ptrs = (ctypes.c_void_p * 7)(*[f.data_ptr() for f in self._frag])
            self._frag_kp = (ctypes.c_int * 7)(*[kpad(byname[n].cin) for n in self.FUSED_LAYERS])
            self._frag_ver = self._wver
        return self._frag_ptrs, self._frag_kp

    @property
    def solves_in_kernel(self) -> bool:
        """coarse16f at 128x128: solve_batch(..., net=self) forms the warm
        start inside the solve (fadesim_nn_solve_batch)."""
        return self.arch == "coarse16f" and (self.M, self.N) == (128, 128)

    def forward_fused(self, x, out=None):
        """x [B][32][32][8] bf16 -> out [B][32][32][16] bf16, the whole network
        in one launch (every activation of a crossbar in LDS)."""
        B = x.shape[0]
        if out is None:
            out = torch.empty((B, 32, 32, OUT_C), dtype=torch.bfloat16, device=self.dev)
        ptrs, kps = self.fused_weights()
        _chk(nat.lib().fadesim_nn_coarse_forward(B, ctypes.c_void_p(x.data_ptr()), ptrs, kps,
                                                 ctypes.c_void_p(out.data_ptr()), _s(self.dev)), "coarse_forward")
        return out

    def _needed_later(self, tensor, current):
        seen = False
        for op in self.ops:
            if seen and (op.src == tensor or op.skip == tensor):
                return True
            if op.name == current:
                seen = True
        return False

    def predict(self, g, r, vin, chunk: int = 2048):
        """Warm-start drops [B][2][M][N] FP32 for solve_batch(..., warm=...)."""
        B = g.shape[0]
        warm = torch.empty((B, 2, self.M, self.N), dtype=torch.float32, device=self.dev)
        g = g.to(torch.float32).contiguous()
        r = r.to(torch.float64).reshape(-1).expand(B).contiguous()
        vin = vin.to(torch.float64).contiguous()
        for b0 in range(0, B, chunk):
            b1 = min(B, b0 + chunk)
            feat, base, scale = self.features(g[b0:b1], r[b0:b1], vin[b0:b1])
            out = self.forward_fused(feat) if self.fused else self.forward(feat)
            lv = self.out_level
            if lv == 0:
                _chk(nat.lib().fadesim_nn_combine(b1 - b0, self.M, self.N, ctypes.c_void_p(base.data_ptr()),
                                                  ctypes.c_void_p(scale.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                                  OUT_C, ctypes.c_void_p(warm[b0:b1].data_ptr()), _s(self.dev)),
                     "nn_combine")
            else:
                _chk(nat.lib().fadesim_nn_combine_up(b1 - b0, self.M, self.N, ctypes.c_void_p(base.data_ptr()),
                                                     ctypes.c_void_p(scale.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), OUT_C, lv,
                                                     ctypes.c_void_p(warm[b0:b1].data_ptr()), _s(self.dev)),
                     "nn_combine_up")
        return warm

    def predict_parts(self, g, r, vin, chunk: int = 2048, with_base: bool = True):
        """The warm start before its final combination: base field [B][2][M][N]
        FP32, network output [B][M >> level][N >> level][16] BF16, scale [B]
        FP32 and the output level (>= 1), for solve_batch(warm=base,
        warm_coarse=out, warm_scale=scale, warm_level=level), which forms
        base + scale x bilinear(out) while loading -- the same field predict()
        returns, without writing it.  with_base=False (pooled features): base
        is None and solve_batch(warm_coarse=out, ...) forms the base itself,
        bit for bit the same."""
        lv = self.out_level
        if lv < 1:
            raise ValueError("predict_parts needs a network whose output is coarser than the crossbar")
        B = g.shape[0]
        g = g.to(torch.float32).contiguous()
        r = r.to(torch.float64).reshape(-1).expand(B).contiguous()
        vin = vin.to(torch.float64).contiguous()
        with_base = with_base or self.in_level == 0
        base = torch.empty((B, 2, self.M, self.N), dtype=torch.float32, device=self.dev) if with_base else None
        scale = torch.empty(B, dtype=torch.float32, device=self.dev)
        H, W = self._hw(lv)
        out = torch.empty((B, H, W, OUT_C), dtype=torch.bfloat16, device=self.dev)
        for b0 in range(0, B, chunk):
            b1 = min(B, b0 + chunk)
            feat, _, _ = self.features(g[b0:b1], r[b0:b1], vin[b0:b1], base[b0:b1] if with_base else None,
                                       scale[b0:b1], want_base=with_base)
            if self.fused:
                self.forward_fused(feat, out=out[b0:b1])
            else:
                out[b0:b1] = self.forward(feat)
        return base, out, scale, lv

    def warm_kwargs(self, g, r, vin, base_in_kernel: bool = True):
        """solve_batch keyword arguments carrying this network's warm start:
        the parts (combined while loading) when the output is coarser than the
        crossbar, else the combined field.  With pooled features and
        base_in_kernel (the default) the solve forms the base field itself
        (two undamped IR2S iterations from zero, bit for bit the features
        kernel's base for grounded bitlines), so it is neither written nor
        read; base_in_kernel=False passes it from the features kernel."""
        if self.out_level >= 1:
            base, out, scale, lv = self.predict_parts(g, r, vin, with_base=not base_in_kernel)
            kw = {"warm_coarse": out, "warm_scale": scale, "warm_level": lv}
            if base is not None:
                kw["warm"] = base
            return kw
        return {"warm": self.predict(g, r, vin)}

    # ------------------------------------------------------------------ backward
    def backward(self, acts, dout, grad):
        """dout: d loss / d out (bf16).  Accumulates weight gradients into grad (FP32 flat)."""
        L = nat.lib()
        s = _s(self.dev)
        B = dout.shape[0]
        grads = {"out": dout}
        for op in reversed(self.ops):
            H, W = self._hw(op.level)
            if op.name not in self._grad_tensors:
                continue
            dy = grads.pop(op.name)
            if op.kind == "conv":
                y = acts[op.name]
                if op.relu:
                    _chk(L.fadesim_nn_relu_mask(dy.numel(), ctypes.c_void_p(dy.data_ptr()), None,
                                                ctypes.c_void_p(y.data_ptr()), s), "relu_mask")
                gv = self._view(grad, op)
                _chk(L.fadesim_nn_conv3x3_wgrad(B, H, W, op.cin, op.cout, kpad(op.cin),
                                                ctypes.c_void_p(acts[op.src].data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                                ctypes.c_void_p(gv.data_ptr()), s), "wgrad " + op.name)
                if op.src in self._grad_tensors:
                    kpt = kpad(op.cout)
                    wT = torch.empty((op.cin, kpt), dtype=torch.bfloat16, device=self.dev)
                    _chk(L.fadesim_nn_flip_weights(op.cin, op.cout, kpad(op.cin), kpt,
                                                   ctypes.c_void_p(self._view(self.wb, op).data_ptr()),
                                                   ctypes.c_void_p(wT.data_ptr()), s), "flip")
                    dx = torch.empty((B, H, W, op.cin), dtype=torch.bfloat16, device=self.dev)
                    _chk(L.fadesim_nn_conv3x3_fwd(B, H, W, op.cout, op.cin, kpt, ctypes.c_void_p(dy.data_ptr()),
                                                  ctypes.c_void_p(wT.data_ptr()), ctypes.c_void_p(dx.data_ptr()), 0, s),
                         "dgrad " + op.name)
                    self._accum(grads, op.src, dx)
            elif op.kind == "pool":
                src = op.src
                if src in grads:
                    _chk(L.fadesim_nn_pool2_bwd(B, 2 * H, 2 * W, op.cout, ctypes.c_void_p(dy.data_ptr()), 1,
                                                ctypes.c_void_p(grads[src].data_ptr()), s), "pool_bwd")
                else:
                    dx = torch.empty((B, 2 * H, 2 * W, op.cout), dtype=torch.bfloat16, device=self.dev)
                    _chk(L.fadesim_nn_pool2_bwd(B, 2 * H, 2 * W, op.cout, ctypes.c_void_p(dy.data_ptr()), 0,
                                                ctypes.c_void_p(dx.data_ptr()), s), "pool_bwd")
                    grads[src] = dx
            else:  # up_add
                dt = torch.empty((B, H // 2, W // 2, op.cout), dtype=torch.bfloat16, device=self.dev)
                _chk(L.fadesim_nn_up2_bwd(B, H // 2, W // 2, op.cout, ctypes.c_void_p(dy.data_ptr()),
                                          ctypes.c_void_p(dt.data_ptr()), s), "up_bwd")
                self._accum(grads, op.src, dt)
                self._accum(grads, op.skip, dy)

    def _accum(self, grads, name, g):
        if name in grads:
            _chk(nat.lib().fadesim_nn_add(g.numel(), ctypes.c_void_p(grads[name].data_ptr()),
                                          ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(grads[name].data_ptr()),
                                          _s(self.dev)), "grad add")
        else:
            grads[name] = g

    # ------------------------------------------------------------------ training
    def train(self, data, epochs=12, batch=64, lr=2e-3, wd=0.0, log=print, seed=0, max_seconds=None):
        """data: dict with feat [P][M][N][8] bf16, target [P][2][M'][N'] FP32 (units of scale) at the
        network's output resolution M' = M >> out_level."""
        L = nat.lib()
        s = _s(self.dev)
        P = data["feat"].shape[0]
        n = self.nparams_padded
        grad = torch.zeros(n, dtype=torch.float32, device=self.dev)
        m = torch.zeros_like(grad)
        v = torch.zeros_like(grad)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        steps_per_epoch = P // batch
        total = epochs * steps_per_epoch
        step = 0
        curve = []
        t0 = time.time()
        for ep in range(epochs):
            perm = torch.randperm(P, generator=gen).to(self.dev)
            ep_loss = 0.0
            for it in range(steps_per_epoch):
                idx = perm[it * batch:(it + 1) * batch]
                x = data["feat"].index_select(0, idx)
                tgt = data["target"].index_select(0, idx)
                acts = self.forward(x, keep=True)
                out = acts["out"]
                dout = torch.empty_like(out)
                loss.zero_()
                _chk(L.fadesim_nn_mse(batch, self.M >> self.out_level, self.N >> self.out_level,
                                      ctypes.c_void_p(out.data_ptr()), OUT_C,
                                      ctypes.c_void_p(tgt.data_ptr()), ctypes.c_void_p(dout.data_ptr()),
                                      ctypes.c_void_p(loss.data_ptr()), s), "mse")
                grad.zero_()
                self.backward(acts, dout, grad)
                step += 1
                lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
                _chk(L.fadesim_nn_adam(n, ctypes.c_void_p(self.w.data_ptr()), ctypes.c_void_p(m.data_ptr()),
                                       ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                       ctypes.c_void_p(self.wb.data_ptr()), ctypes.c_float(lr_t),
                                       ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), step,
                                       ctypes.c_float(wd), ctypes.c_float(1.0), s), "adam")
                self._wver += 1
                if it % 50 == 0:
                    lv = float(loss.item())
                    ep_loss = lv
                    curve.append((step, lv))
                    log(f"epoch {ep} step {it}/{steps_per_epoch} loss {lv:.4e} lr {lr_t:.2e} "
                        f"({time.time() - t0:.0f}s)")
                if max_seconds and time.time() - t0 > max_seconds:
                    break
            if max_seconds and time.time() - t0 > max_seconds:
                log(f"stopping at the time budget after epoch {ep}")
                break
        self.info.update({"epochs_run": ep + 1, "batch": batch, "lr": lr, "train_seconds": time.time() - t0,
                          "final_loss": ep_loss, "loss_curve": curve})
        return curve


def make_instances(B, seed, device, M=128, N=128):
    """Config-2 distribution on the GPU: 4 cell states, r ~ U[1,5] ohm,
    vin in {0, 0.2} V (p = 0.5).  Seeded per call."""
    gen = torch.Generator(device=device).manual_seed(int(seed))
    states = torch.tensor([2e6, 1.6e6, 1.3e6, 1e6], dtype=torch.float64, device=device)
    st = torch.randint(0, 4, (B, M, N), generator=gen, device=device)
    g = (1.0 / states[st]).to(torch.float32)
    r = 1.0 + 4.0 * torch.rand(B, generator=gen, device=device, dtype=torch.float64)
    vin = torch.where(torch.rand((B, M), generator=gen, device=device) < 0.5, 0.2, 0.0).to(torch.float64)
    return g, r, vin


def generate_dataset(net: WarmStartNet, n_instances: int, seed0: int, chunk: int = 4096, tol: float = 1e-8):
    """Features and targets from the project's own solver."""
    from .solver import solve_batch
    feats, targets = [], []
    for c0 in range(0, n_instances, chunk):
        B = min(chunk, n_instances - c0)
        g, r, vin = make_instances(B, seed0 + c0, net.dev)
        res = solve_batch(g, r, vin, tol=tol, want_field=True)
        if int((res.status != 0).sum()) != 0:
            raise RuntimeError("training-data solve did not converge")
        feat, base, scale = net.features(g, r, vin)
        tgt = (res.field - base) / scale.view(B, 1, 1, 1)
        if net.out_level > 0:  # the network predicts block means on its coarse grid
            f = 1 << net.out_level
            tgt = torch.nn.functional.avg_pool2d(tgt, f)
        feats.append(feat)
        targets.append(tgt)
    return {"feat": torch.cat(feats), "target": torch.cat(targets)}


def load_default(device) -> Optional[WarmStartNet]:
    """The committed trained network, or None when no weights are present."""
    if not os.path.exists(os.path.join(WEIGHTS_DIR, DEFAULT_NAME + ".bin")):
        return None
    return WarmStartNet.load(device)
