"""CPU oracle of the learned preconditioner's network operations (numpy).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import it, as the checker of the GPU kernels in
efwi-3d_amd/csrc/efwi_unet.hip; no product code path calls it. The network has
no counterpart in the reference (EFWI-3D has no learned part), so there is no
reference file to cite: each function restates the textbook operation, and
tests/test_nn_oracle.py pins them (scipy.ndimage.correlate for the
convolution, adjoint identities for its input and weight gradients and for
pooling/upsampling, finite differences for the loss).

Layout as on the GPU: volumes are channels-last arrays [nz, nx, ny, c];
3x3x3 weights are [27, cout, cin] with tap t = 9 (dz+1) + 3 (dx+1) + (dy+1).
Arithmetic in float64 on the given (typically bf16-rounded) values.
"""
from __future__ import annotations

import numpy as np

TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    """float32 -> bfloat16 bit patterns (uint16), ties to even."""
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _shift(xp, d, shape):
    """xp padded by 1 on the three spatial axes: the view x[v + d]."""
    dz, dx, dy = d
    nz, nx, ny = shape
    return xp[1 + dz:1 + dz + nz, 1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny]


def conv3(x, w, bias=None, relu=False):
    """out[v][n] = sum_t sum_k x[v + d_t][k] w[t][n][k] (+ bias[n]), zero padding."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    shape = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros(shape + (w.shape[1],))
    for t, d in enumerate(TAPS):
        out += _shift(xp, d, shape) @ w[t].T
    if bias is not None:
        out += np.asarray(bias, np.float64)
    if relu:
        out = np.maximum(out, 0.0)
    return out


def conv3_dgrad(dout, w):
    """din[v][k] = sum_t sum_n dout[v - d_t][n] w[t][n][k] (the transpose of conv3)."""
    dout = np.asarray(dout, np.float64)
    w = np.asarray(w, np.float64)
    shape = dout.shape[:3]
    dp = np.pad(dout, ((1, 1), (1, 1), (1, 1), (0, 0)))
    din = np.zeros(shape + (w.shape[2],))
    for t, (dz, dx, dy) in enumerate(TAPS):
        din += _shift(dp, (-dz, -dx, -dy), shape) @ w[t]
    return din


def conv3_wgrad(dpre, x):
    """dw[t][n][k] = sum_v dpre[v][n] x[v + d_t][k]."""
    dpre = np.asarray(dpre, np.float64)
    x = np.asarray(x, np.float64)
    shape = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    g = dpre.reshape(-1, dpre.shape[3])
    return np.stack([g.T @ _shift(xp, d, shape).reshape(-1, x.shape[3]) for d in TAPS])


def pool2(x):
    x = np.asarray(x, np.float64)
    nz, nx, ny, c = x.shape
    return x.reshape(nz // 2, 2, nx // 2, 2, ny // 2, 2, c).mean(axis=(1, 3, 5))


def pool2_bwd(dc):
    dc = np.asarray(dc, np.float64)
    return np.repeat(np.repeat(np.repeat(dc, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    x = np.asarray(x, np.float64)
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_bwd(df):
    df = np.asarray(df, np.float64)
    nz, nx, ny, c = df.shape
    return df.reshape(nz // 2, 2, nx // 2, 2, ny // 2, 2, c).sum(axis=(1, 3, 5))


def head(act, w, b):
    return np.asarray(act, np.float64) @ np.asarray(w, np.float64) + float(b)


def head_bwd(act, w, dout):
    act = np.asarray(act, np.float64)
    dout = np.asarray(dout, np.float64)
    dact = np.where(act > 0, dout[..., None] * np.asarray(w, np.float64), 0.0)
    dw = (act * dout[..., None]).reshape(-1, act.shape[-1]).sum(0)
    return dact, dw, dout.sum()


GATE_LO, GATE_HI = -4.0, 6.0


def gate(feat, o):
    """u = g exp(clamp(o, -4, 6)), g = channel 0 of the network input."""
    return np.asarray(feat, np.float64)[..., 0] * np.exp(np.clip(np.asarray(o, np.float64), GATE_LO, GATE_HI))


def gate_bwd(o, u, du):
    o = np.asarray(o, np.float64)
    return np.where((o > GATE_LO) & (o < GATE_HI), np.asarray(du, np.float64) * np.asarray(u, np.float64), 0.0)


def mse(out, target, mask, count):
    d = (np.asarray(out, np.float64) - target) * mask
    return float((d * d).sum() / count), 2.0 * d / count


def adam(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    return w - lr * (m1 * c1) / (np.sqrt(m2 * c2) + eps), m1, m2


def pack(g, illum, npml, inv_g, inv_i):
    """Network input (16 channels) from padded FWI-grid volumes."""
    p = npml
    gi = np.asarray(g, np.float64)[p:-p, p:-p, p:-p]
    ii = np.asarray(illum, np.float64)[p:-p, p:-p, p:-p]
    f = np.zeros(gi.shape + (16,))
    f[..., 0] = gi * inv_g
    f[..., 1] = np.log10(ii * inv_i + 1e-6) / 6.0 + 1.0
    return f
