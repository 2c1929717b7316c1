"""Learned preconditioner: a 3-D U-Net on the library's BF16 matrix-core
kernels (include/efwi_nn.h), trained and run on the GPU.

It maps the raw FWI gradient plus the source illumination (2 input channels,
padded to 16) to a model update (1 output channel). No reference
counterpart: EFWI-3D has no learned part.

Architecture (3 levels, 16 base channels, 3x3x3 convolutions with ReLU,
2x2x2 average pooling, nearest upsampling, skip concatenation):

    enc1a 16->16, enc1b 16->16 = e1
      pool -> enc2a 16->32, enc2b 32->32 = e2
        pool -> bot_a 32->64, bot_b 64->64 = b3
                  dec2a (up(b3) | e2) 96->32, dec2b 32->32 = d2b
                            dec1a (up(d2b) | e1) 48->16, dec1b 16->16 -> head 16->1 (FP32)
    update = g exp(clamp(head, -4, 6))   (g = input channel 0: a learned,
                            spatially varying scaling of the gradient, as a
                            preconditioner is)

Weights and activations are BF16 with FP32 accumulation; FP32 master weights
and Adam state; activation gradients BF16, weight gradients FP32. The
skip concatenations are never formed: the decoder convolutions read the coarse
tensor through its nearest upsampling and the skip activation as one input,
and write their input gradients as two outputs; the ReLU derivative of each
activation is
applied as a mask inside the input-gradient, pooling-gradient and
upsampling-gradient kernels.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _native as N

# (name, cin, cout, level)
LAYERS = [
    ("enc1a", 16, 16, 1), ("enc1b", 16, 16, 1),
    ("enc2a", 16, 32, 2), ("enc2b", 32, 32, 2),
    ("bot_a", 32, 64, 3), ("bot_b", 64, 64, 3),
    ("dec2a", 96, 32, 2), ("dec2b", 32, 32, 2),
    ("dec1a", 48, 16, 1), ("dec1b", 16, 16, 1),
]
IN_CHANNELS = 2       # raw gradient, log illumination
WGRAD_WORKGROUPS = 768   # weight gradient: about three resident workgroups per CU


def nn_check(rc: int, what: str):
    if rc != 0:
        raise N.EfwiError(f"{what} failed ({rc}: {'invalid argument' if rc == 1 else 'kernel launch'})")


def _p(t, offset_elems=0):
    return C.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def init_params(seed: int = 0):
    """Initial FP32 parameters (CPU tensors) from a seeded generator: He-normal
    3x3x3 weights [27][cout][cin] (the padded input channels of enc1a zero),
    zero biases, a small head (the initial scaling exp(o) is close to 1, the
    plain gradient). Returns (w, b, wh, bh)."""
    g = torch.Generator().manual_seed(seed)
    w, b = {}, {}
    for name, cin, cout, _ in LAYERS:
        fan_in = 27 * (IN_CHANNELS if name == "enc1a" else cin)
        wt = torch.randn(27, cout, cin, generator=g) * math.sqrt(2.0 / fan_in)
        if name == "enc1a":
            wt[:, :, IN_CHANNELS:] = 0.0
        w[name] = wt.contiguous()
        b[name] = torch.zeros(cout)
    wh = torch.randn(16, generator=g) * 0.01
    return w, b, wh, torch.zeros(1)


class UNet3D:
    def __init__(self, device: int = 0, seed: int = 0):
        self.dev = torch.device(f"cuda:{device}")
        self.lib = N.lib()
        w, b, wh, bh = init_params(seed)
        init = []
        for name, *_ in LAYERS:
            init += [(name + ".w", w[name]), (name + ".b", b[name])]
        init += [("head.w", wh), ("head.b", bh)]
        # One flat buffer per role (FP32 weights, gradients, Adam moments, the
        # BF16 copy): a single Adam launch updates every parameter. The
        # per-parameter tensors are views; their offsets stay multiples of 8
        # elements (16-byte aligned BF16 weights), as every weight tensor holds
        # 27 cout cin values, every bias a multiple of 16 and the head is last.
        n = sum(v.numel() for _, v in init)
        self._flat = torch.empty(n, device=self.dev)
        self._gflat = torch.zeros(n, device=self.dev)
        self._m1flat = torch.zeros(n, device=self.dev)
        self._m2flat = torch.zeros(n, device=self.dev)
        self._flat16 = torch.empty(n, dtype=torch.bfloat16, device=self.dev)
        self._offsets = {}
        off = 0
        for k, v in init:
            self._offsets[k] = (off, v.shape)
            off += v.numel()
        view = lambda flat, k: flat[self._offsets[k][0]:self._offsets[k][0] + self._offsets[k][1].numel()].view(
            self._offsets[k][1])
        self._flat.copy_(torch.cat([v.reshape(-1) for _, v in init]).to(self.dev))
        self.w = {name: view(self._flat, name + ".w") for name, *_ in LAYERS}
        self.b = {name: view(self._flat, name + ".b") for name, *_ in LAYERS}
        self.wh = view(self._flat, "head.w")
        self.bh = view(self._flat, "head.b")
        self.grads = {k: view(self._gflat, k) for k, _ in init}
        self.m1 = {k: view(self._m1flat, k) for k, _ in init}
        self.m2 = {k: view(self._m2flat, k) for k, _ in init}
        self.w16 = {name: view(self._flat16, name + ".w") for name, *_ in LAYERS}
        self.wt16 = {name: torch.empty_like(self.w16[name]) for name, *_ in LAYERS}
        self._refresh_bf16()
        self.steps = 0
        self._bufs = {}
        self._scratch = None

    # ---- parameters
    def _params(self):
        out = {}
        for name, *_ in LAYERS:
            out[name + ".w"] = self.w[name]
            out[name + ".b"] = self.b[name]
        out["head.w"] = self.wh
        out["head.b"] = self.bh
        return out

    def _refresh_bf16(self):
        self._flat16.copy_(self._flat)
        self._flip()

    def _flip(self):
        """The input-gradient weights (taps mirrored, channels transposed)
        from the BF16 weights, for the mode-0 convolution of the backward."""
        n = len(LAYERS)
        w = (C.c_void_p * n)(*[self.w16[name].data_ptr() for name, *_ in LAYERS])
        wt = (C.c_void_p * n)(*[self.wt16[name].data_ptr() for name, *_ in LAYERS])
        kc = (C.c_int * n)(*[cin for _, cin, _, _ in LAYERS])
        nc = (C.c_int * n)(*[cout for _, _, cout, _ in LAYERS])
        nn_check(self.lib.efwi_nn_flip_weights_n(n, w, kc, nc, wt, self._stream()), "flip")

    def state_dict(self):
        return {k: v.detach().cpu().clone() for k, v in self._params().items()}

    def load_state_dict(self, sd):
        for k, v in self._params().items():
            v.copy_(sd[k].to(self.dev))
        self._refresh_bf16()

    def nparams(self):
        return sum(v.numel() for v in self._params().values())

    # ---- buffers per volume shape
    def _buffers(self, dims):
        if dims in self._bufs:
            return self._bufs[dims]
        nz, nx, ny = dims
        if nz % 4 or nx % 4 or ny % 4:
            raise ValueError(f"U-Net volume {dims}: every size must be a multiple of 4")
        v1 = nz * nx * ny
        v2, v3 = v1 // 8, v1 // 64
        bf = lambda n, c: torch.zeros(n, c, dtype=torch.bfloat16, device=self.dev)
        B = {"a1": bf(v1, 16), "e1": bf(v1, 16), "p1": bf(v2, 16), "a2": bf(v2, 32),
             "e2": bf(v2, 32), "p2": bf(v3, 32), "a3": bf(v3, 64), "b3": bf(v3, 64),
             "d2a": bf(v2, 32), "d2b": bf(v2, 32), "d1a": bf(v1, 16), "d1b": bf(v1, 16),
             "o": torch.zeros(v1, device=self.dev), "out": torch.zeros(v1, device=self.dev),
             "do": torch.zeros(v1, device=self.dev),
             # gradients
             "g1": bf(v1, 16), "g1b": bf(v1, 16), "gu1": bf(v1, 32), "ge1": bf(v1, 16), "gu2": bf(v2, 64),
             "ge2": bf(v2, 32), "g2": bf(v2, 32), "g2b": bf(v2, 32), "gp1": bf(v2, 16), "gp2": bf(v3, 32),
             "g3": bf(v3, 64), "g3b": bf(v3, 64),
             "dout": torch.zeros(v1, device=self.dev), "loss": torch.zeros(1, dtype=torch.float64, device=self.dev)}
        self._bufs[dims] = B
        need = 0
        for name, cin, cout, lev in LAYERS:
            d = tuple(s >> (lev - 1) for s in dims)
            need = max(need, int(self.lib.efwi_nn_wgrad_scratch(cout, cin, *d, self._wgrad_per(cout, cin, d))))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = torch.empty(need, device=self.dev)
        return B

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def _conv(self, name, src, lds, dst, ldd, dims, relu=True, src2=None):
        """Forward convolution of layer `name`; with src2, its input is the
        nearest upsampling of the half-size src (lds channels) followed by the
        channels of src2."""
        _, cin, cout, _ = next(l for l in LAYERS if l[0] == name)
        if src2 is None:
            rc = self.lib.efwi_nn_conv3(_p(src), lds, cin, _p(self.w16[name]), _p(self.b[name]), _p(dst), ldd, cout,
                                        None, 0, *dims, 0, int(relu), self._stream())
        else:
            k2 = src2.shape[1]
            rc = self.lib.efwi_nn_conv3_cat(_p(src), lds, cin - k2, 1, _p(src2), k2, k2, _p(self.w16[name]),
                                            _p(self.b[name]), _p(dst), ldd, cout, None, 0, 0, None, 0, *dims, 0,
                                            int(relu), self._stream())
        nn_check(rc, f"conv {name}")

    def _dgrad(self, name, gout, ldg, gin, ldi, mask, ldm, dims, gin2=None):
        """Input gradient of layer `name` (a forward convolution with the
        flipped weights); with gin2, its first ldi channels go to gin and the
        rest to gin2."""
        _, cin, cout, _ = next(l for l in LAYERS if l[0] == name)
        mk = None if mask is None else _p(mask)
        if gin2 is None:
            rc = self.lib.efwi_nn_conv3(_p(gout), ldg, cout, _p(self.wt16[name]), None, _p(gin), ldi, cin, mk, ldm,
                                        *dims, 0, 0, self._stream())
        else:
            n2 = gin2.shape[1]
            rc = self.lib.efwi_nn_conv3_cat(_p(gout), ldg, cout, 0, None, 0, 0, _p(self.wt16[name]), None, _p(gin),
                                            ldi, cin - n2, _p(gin2), n2, n2, mk, ldm, *dims, 0, 0, self._stream())
        nn_check(rc, f"dgrad {name}")

    def _wgrad_per(self, cout, cin, dims):
        """Bricks per weight-gradient workgroup: enough workgroups to fill the
        GPU (about WGRAD_WORKGROUPS over the (m, n) channel blocks), between 2
        and 24 bricks each (fewer bricks per workgroup multiply the FP32
        partials that are written and summed, more keep more halo planes in
        the workgroup's z ring; measured on MI355X at the U-Net's shapes:
        48->16 at 96x128x128 per 8 205 us, per 24 172 us; 96->32 at 48x64x64
        per 8 112 us, per 12 96 us)."""
        bricks = int(self.lib.efwi_nn_wgrad_scratch(cout, cin, *dims, 1)) // (27 * cout * cin + cout)
        blocks = (cout // 16) * (cin // 16)
        return int(min(24, max(2, bricks * blocks // WGRAD_WORKGROUPS)))

    def _wgrad(self, name, gpre, ldg, src, lds, dims, src2=None):
        """Weight and bias gradients of layer `name`; with src2, its input is
        the nearest upsampling of the half-size src (lds channels) followed by
        the channels of src2."""
        _, cin, cout, _ = next(l for l in LAYERS if l[0] == name)
        s = self._stream()
        per = self._wgrad_per(cout, cin, dims)
        gw, gb = _p(self.grads[name + ".w"]), _p(self.grads[name + ".b"])
        if src2 is None:
            rc = self.lib.efwi_nn_wgrad(_p(gpre), ldg, cout, _p(src), lds, cin, _p(self._scratch),
                                        self._scratch.numel(), gw, gb, *dims, per, s)
        else:
            k2 = src2.shape[1]
            rc = self.lib.efwi_nn_wgrad_cat(_p(gpre), ldg, cout, _p(src), lds, cin - k2, 1, _p(src2), k2, k2,
                                            _p(self._scratch), self._scratch.numel(), gw, gb, *dims, per, s)
        nn_check(rc, f"wgrad {name}")

    # ---- forward / backward
    def forward(self, feat, dims):
        """feat: [V, 16] bfloat16 (channels-last, V = nz*nx*ny); returns the FP32 update [V]."""
        B = self._buffers(dims)
        s = self._stream()
        d1, d2 = dims, tuple(x // 2 for x in dims)
        d3 = tuple(x // 4 for x in dims)
        v1 = d1[0] * d1[1] * d1[2]
        self._conv("enc1a", feat, 16, B["a1"], 16, d1)
        self._conv("enc1b", B["a1"], 16, B["e1"], 16, d1)
        nn_check(self.lib.efwi_nn_pool2(_p(B["e1"]), 16, _p(B["p1"]), 16, 16, *d1, s), "pool1")
        self._conv("enc2a", B["p1"], 16, B["a2"], 32, d2)
        self._conv("enc2b", B["a2"], 32, B["e2"], 32, d2)
        nn_check(self.lib.efwi_nn_pool2(_p(B["e2"]), 32, _p(B["p2"]), 32, 32, *d2, s), "pool2")
        self._conv("bot_a", B["p2"], 32, B["a3"], 64, d3)
        self._conv("bot_b", B["a3"], 64, B["b3"], 64, d3)
        self._conv("dec2a", B["b3"], 64, B["d2a"], 32, d2, src2=B["e2"])
        self._conv("dec2b", B["d2a"], 32, B["d2b"], 32, d2)
        self._conv("dec1a", B["d2b"], 32, B["d1a"], 16, d1, src2=B["e1"])
        self._conv("dec1b", B["d1a"], 16, B["d1b"], 16, d1)
        nn_check(self.lib.efwi_nn_head(_p(B["d1b"]), 16, 16, _p(self.wh), _p(self.bh), _p(B["o"]), v1, s), "head")
        nn_check(self.lib.efwi_nn_gate(_p(feat), 16, _p(B["o"]), _p(B["out"]), v1, s), "gate")
        self._feat, self._dims = feat, dims
        return B["out"]

    def scaling_log(self, dims):
        """The head output o of the last forward (FP32 [nz, nx, ny]): the
        network's log-scaling of the gradient before the clamp to [-4, 6]."""
        return self._buffers(dims)["o"].reshape(dims)

    def backward(self, dout, through_gate=True):
        """Gradients of every parameter (self.grads) from dL/dupdate [V] (FP32),
        or from dL/do of the head output when through_gate is False."""
        B = self._buffers(self._dims)
        s = self._stream()
        d1 = self._dims
        v1 = d1[0] * d1[1] * d1[2]
        if through_gate:
            nn_check(self.lib.efwi_nn_gate_bwd(_p(B["o"]), _p(B["out"]), _p(dout), _p(B["do"]), v1, s), "gate bwd")
            dout = B["do"]
        nn_check(self.lib.efwi_nn_head_bwd(_p(B["d1b"]), 16, 16, _p(self.wh), _p(dout), _p(B["g1b"]), 16,
                                           _p(self.grads["head.w"]), _p(self.grads["head.b"]), v1, s), "head bwd")
        return self._backward_body()

    def _backward_body(self):
        """Gradients of every convolution from g1b, the gradient of the last
        activation."""
        B = self._buffers(self._dims)
        s = self._stream()
        d1 = self._dims
        d2, d3 = tuple(x // 2 for x in d1), tuple(x // 4 for x in d1)
        # level 1 decoder
        self._wgrad("dec1b", B["g1b"], 16, B["d1a"], 16, d1)
        self._dgrad("dec1b", B["g1b"], 16, B["g1"], 16, B["d1a"], 16, d1)
        self._wgrad("dec1a", B["g1"], 16, B["d2b"], 32, d1, src2=B["e1"])
        # the ReLU derivative of up(d2b) | e1 is applied where each part is
        # consumed, not here: the upsampled part's (that of d2b, shared by a
        # parent's 8 children) inside the upsampling gradient, e1's after the
        # pooled gradient is added to it; either way a masked value is exactly
        # zero and a kept one the same sum
        self._dgrad("dec1a", B["g1"], 16, B["gu1"], 32, None, 0, d1, gin2=B["ge1"])
        nn_check(self.lib.efwi_nn_up2_bwd(_p(B["gu1"]), 32, _p(B["g2b"]), 32, 32, *d1, _p(B["d2b"]), 32, s),
                 "up1 bwd")
        # level 2 decoder
        self._wgrad("dec2b", B["g2b"], 32, B["d2a"], 32, d2)
        self._dgrad("dec2b", B["g2b"], 32, B["g2"], 32, B["d2a"], 32, d2)
        self._wgrad("dec2a", B["g2"], 32, B["b3"], 64, d2, src2=B["e2"])
        self._dgrad("dec2a", B["g2"], 32, B["gu2"], 64, None, 0, d2, gin2=B["ge2"])
        nn_check(self.lib.efwi_nn_up2_bwd(_p(B["gu2"]), 64, _p(B["g3b"]), 64, 64, *d2, _p(B["b3"]), 64, s),
                 "up2 bwd")
        # bottleneck
        self._wgrad("bot_b", B["g3b"], 64, B["a3"], 64, d3)
        self._dgrad("bot_b", B["g3b"], 64, B["g3"], 64, B["a3"], 64, d3)
        self._wgrad("bot_a", B["g3"], 64, B["p2"], 32, d3)
        self._dgrad("bot_a", B["g3"], 64, B["gp2"], 32, None, 0, d3)
        # level 2 encoder: skip gradient + pooled gradient, then the ReLU mask
        nn_check(self.lib.efwi_nn_pool2_bwd(_p(B["gp2"]), 32, _p(B["ge2"]), 32, 32, *d2, 1, _p(B["e2"]), 32, s),
                 "pool2 bwd")
        self._wgrad("enc2b", B["ge2"], 32, B["a2"], 32, d2)
        self._dgrad("enc2b", B["ge2"], 32, B["g2"], 32, B["a2"], 32, d2)
        self._wgrad("enc2a", B["g2"], 32, B["p1"], 16, d2)
        self._dgrad("enc2a", B["g2"], 32, B["gp1"], 16, None, 0, d2)
        # level 1 encoder
        nn_check(self.lib.efwi_nn_pool2_bwd(_p(B["gp1"]), 16, _p(B["ge1"]), 16, 16, *d1, 1, _p(B["e1"]), 16, s),
                 "pool1 bwd")
        self._wgrad("enc1b", B["ge1"], 16, B["a1"], 16, d1)
        self._dgrad("enc1b", B["ge1"], 16, B["g1"], 16, B["a1"], 16, d1)
        self._wgrad("enc1a", B["g1"], 16, self._feat, 16, d1)
        return self.grads

    def loss(self, out, target, mask, count):
        B = self._buffers(self._dims)
        nn_check(self.lib.efwi_nn_mse(_p(out), _p(target), _p(mask), out.numel(), float(count), _p(B["dout"]),
                                      _p(B["loss"]), self._stream()), "mse")
        return B["loss"], B["dout"]

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.steps += 1
        nn_check(self.lib.efwi_nn_adam(_p(self._flat), _p(self._gflat), _p(self._m1flat), _p(self._m2flat),
                                       self._flat.numel(), lr, b1, b2, eps, self.steps, _p(self._flat16),
                                       self._stream()), "adam")
        self._flip()

    def train_step(self, feat, target, mask, count, dims, lr=1e-3, log_scaling=False):
        """One Adam step on the mean-square error of the update (target: the
        update), or of the head output (log_scaling: target = the log of the
        per-voxel scaling update / gradient). The loss, its gradient through
        the gate and the head's gradient run as one kernel; the result is that
        of loss() then backward()."""
        out = self.forward(feat, dims)
        B = self._buffers(dims)
        nn_check(self.lib.efwi_nn_loss_head_bwd(_p(B["d1b"]), 16, _p(self.wh), _p(B["o"]),
                                                None if log_scaling else _p(out), _p(target), _p(mask),
                                                out.numel(), float(count), _p(B["g1b"]), 16,
                                                _p(self.grads["head.w"]), _p(self.grads["head.b"]), _p(B["loss"]),
                                                self._stream()), "loss and head gradient")
        self._backward_body()
        self.adam(lr)
        return B["loss"]


def pack_input(g, illum, npml, device=0):
    """Network input [V, 16] bf16 from padded FWI-grid gradient and illumination
    (CUDA float32 tensors), with the scales of channel 0 and 1 taken from the
    volumes' absolute maxima (over the whole padded grid)."""
    lib = N.lib()
    s = C.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)
    mx = torch.zeros(2, device=g.device)
    nn_check(lib.efwi_nn_absmax(_p(g), g.numel(), _p(mx, 0), s), "absmax g")
    nn_check(lib.efwi_nn_absmax(_p(illum), illum.numel(), _p(mx, 1), s), "absmax illum")
    gmax, imax = (float(v) for v in mx.cpu())
    nz, nx, ny = g.shape
    dims = (nz - 2 * npml, nx - 2 * npml, ny - 2 * npml)
    feat = torch.empty(dims[0] * dims[1] * dims[2], 16, dtype=torch.bfloat16, device=g.device)
    nn_check(lib.efwi_nn_pack(_p(g), _p(illum), nz, nx, ny, npml, 1.0 / max(gmax, 1e-30), 1.0 / max(imax, 1e-30),
                              _p(feat), s), "pack")
    return feat, dims


def training_flops(dims) -> float:
    """Matrix-core work of one training step at `dims` voxels: 2 x 27 x cin x
    cout multiply-adds per voxel for the forward convolution, the weight
    gradient of every layer and the input gradient of every layer but the
    first (whose input needs none); channels as computed (the input's 16)."""
    v1 = dims[0] * dims[1] * dims[2]
    total = 0.0
    for name, cin, cout, lev in LAYERS:
        per = 2.0 * 27 * cin * cout * v1 / 8 ** (lev - 1)
        total += per * (2 if name == "enc1a" else 3)
    return total


def _time_training(dims, steps, seed=0):
    """Seconds per training step of the U-Net on a random 2-channel volume of
    `dims` voxels (after 3 untimed steps)."""
    import time
    rng = np.random.default_rng(seed)
    v = dims[0] * dims[1] * dims[2]
    net = UNet3D(0, seed=1)
    f0 = np.zeros((v, 16), np.float32)
    f0[:, :IN_CHANNELS] = rng.standard_normal((v, IN_CHANNELS))
    feat = torch.from_numpy(f0).to(net.dev).to(torch.bfloat16)
    target = torch.from_numpy(rng.standard_normal(v).astype(np.float32)).to(net.dev)
    mask = torch.ones(v, device=net.dev)
    for _ in range(3):
        net.train_step(feat, target, mask, v, dims)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.train_step(feat, target, mask, v, dims)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


if __name__ == "__main__":
    # python -m efwi.unet [nz nx ny [steps]]: time the training step
    import sys
    a = [int(x) for x in sys.argv[1:]]
    dims = tuple(a[:3]) if len(a) >= 3 else (96, 128, 128)
    steps = a[3] if len(a) >= 4 else 10
    t = _time_training(dims, steps)
    tf = training_flops(dims) / t / 1e12
    print(f"U-Net training step at {dims}: {t * 1e3:.2f} ms, {tf:.0f} TFLOP/s "
          f"({100 * tf / 2500:.1f} % of the 2.5 PFLOP/s dense BF16 peak)", flush=True)
