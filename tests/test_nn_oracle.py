"""Pins of the network oracle (oracle/nn_oracle.py) on the CPU.

The convolution is checked against scipy.ndimage.correlate (a library
routine: zero padding, 3x3x3 window, weight tap t = 9 (dz+1) + 3 (dx+1) +
(dy+1) aligned with the input offset); its input gradient and weight gradient
against the adjoint identities <conv(x, w), y> = <x, dgrad(y, w)> =
<w, wgrad(y, x)>, which fail for a transposed, flipped or shifted operand;
pooling and upsampling against their adjoint identities and closed forms; the
loss gradient against centred finite differences; bf16 rounding against
known bit patterns (ties to even)."""
import numpy as np
import pytest
from scipy import ndimage

import nn_oracle as O


def _rand(rng, *shape):
    return rng.standard_normal(shape)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.140625, 1e-40], np.float32)
    bits = O.bf16_bits(x)
    # 1 + 2^-8 is a tie between 1.0 (even) and 1.0078125: to 1.0; 1 + 3*2^-8 ties up to 1.015625
    assert list(bits[:5]) == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x4049]
    assert O.from_bits(bits)[0] == 1.0


def test_conv3_equals_scipy_correlate():
    rng = np.random.default_rng(0)
    x = _rand(rng, 5, 6, 7, 3)
    w = _rand(rng, 27, 2, 3)
    b = _rand(rng, 2)
    out = O.conv3(x, w, bias=b)
    for n in range(2):
        ref = np.full(x.shape[:3], b[n])
        for k in range(3):
            kern = w[:, n, k].reshape(3, 3, 3)  # [dz+1][dx+1][dy+1]
            ref += ndimage.correlate(x[..., k], kern, mode="constant", cval=0.0)
        assert np.allclose(out[..., n], ref, rtol=1e-12, atol=1e-12)
    assert np.array_equal(O.conv3(x, w, bias=b, relu=True), np.maximum(out, 0))


def test_conv3_gradients_are_adjoint():
    rng = np.random.default_rng(1)
    x = _rand(rng, 4, 5, 6, 3)
    w = _rand(rng, 27, 4, 3)
    y = _rand(rng, 4, 5, 6, 4)
    lhs = float(np.sum(O.conv3(x, w) * y))
    assert np.isclose(lhs, float(np.sum(x * O.conv3_dgrad(y, w))), rtol=1e-12)
    assert np.isclose(lhs, float(np.sum(w * O.conv3_wgrad(y, x))), rtol=1e-12)
    # a flipped tap order breaks the identity
    wf = w[::-1].copy()
    assert not np.isclose(lhs, float(np.sum(x * O.conv3_dgrad(y, wf))), rtol=1e-6)


def test_pool_and_upsample():
    rng = np.random.default_rng(2)
    x = _rand(rng, 4, 6, 8, 2)
    y = _rand(rng, 2, 3, 4, 2)
    assert np.isclose(np.sum(O.pool2(x) * y), np.sum(x * O.pool2_bwd(y)))
    assert np.isclose(np.sum(O.up2(y) * x), np.sum(y * O.up2_bwd(x)))
    assert np.isclose(O.pool2(x)[1, 2, 3, 1], x[2:4, 4:6, 6:8, 1].mean())
    assert np.array_equal(O.up2(y)[3, 5, 7], y[1, 2, 3])


def test_head_and_loss_gradients():
    rng = np.random.default_rng(3)
    act = np.maximum(_rand(rng, 3, 4, 5, 16), 0)
    w = _rand(rng, 16)
    t = _rand(rng, 3, 4, 5)
    m = (rng.uniform(size=(3, 4, 5)) > 0.3).astype(float)
    count = m.sum()
    out = O.head(act, w, 0.3)
    loss, dout = O.mse(out, t, m, count)
    # finite differences of the loss in the head weights and a few activations
    dact, dw, db = O.head_bwd(act, w, dout)
    h = 1e-6
    for c in (0, 7, 15):
        wp, wm = w.copy(), w.copy()
        wp[c] += h
        wm[c] -= h
        fd = (O.mse(O.head(act, wp, 0.3), t, m, count)[0] - O.mse(O.head(act, wm, 0.3), t, m, count)[0]) / (2 * h)
        assert np.isclose(dw[c], fd, rtol=1e-6, atol=1e-9)
    fd = (O.mse(O.head(act, w, 0.3 + h), t, m, count)[0] - O.mse(O.head(act, w, 0.3 - h), t, m, count)[0]) / (2 * h)
    assert np.isclose(db, fd, rtol=1e-6)
    idx = tuple(np.argwhere(act[..., 5] > 0)[0]) + (5,)
    ap, am = act.copy(), act.copy()
    ap[idx] += h
    am[idx] -= h
    fd = (O.mse(O.head(ap, w, 0.3), t, m, count)[0] - O.mse(O.head(am, w, 0.3), t, m, count)[0]) / (2 * h)
    assert np.isclose(dact[idx], fd, rtol=1e-5, atol=1e-9)


def test_gate_gradient():
    rng = np.random.default_rng(6)
    feat = _rand(rng, 3, 4, 5, 16)
    o = 0.3 * _rand(rng, 3, 4, 5)
    t = _rand(rng, 3, 4, 5)
    m = np.ones((3, 4, 5))
    u = O.gate(feat, o)
    assert np.allclose(u, feat[..., 0] * np.exp(o))
    _, du = O.mse(u, t, m, m.sum())
    do = O.gate_bwd(o, u, du)
    h = 1e-6
    for idx in ((0, 0, 0), (2, 3, 4), (1, 2, 1)):
        op, om = o.copy(), o.copy()
        op[idx] += h
        om[idx] -= h
        fd = (O.mse(O.gate(feat, op), t, m, m.sum())[0] - O.mse(O.gate(feat, om), t, m, m.sum())[0]) / (2 * h)
        assert np.isclose(do[idx], fd, rtol=1e-6, atol=1e-10)
    # outside the clamp the output is constant in o
    big = np.full_like(o, 9.0)
    assert np.allclose(O.gate(feat, big), feat[..., 0] * np.exp(6.0))
    assert not O.gate_bwd(big, O.gate(feat, big), du).any()


def test_adam_first_step_is_sign_times_lr():
    g = np.array([0.5, -2.0, 1e-3])
    w, m1, m2 = O.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01, eps=0.0)
    assert np.allclose(w, -0.01 * np.sign(g))


def test_pack_channels():
    rng = np.random.default_rng(4)
    g = _rand(rng, 8, 9, 10)
    il = rng.uniform(0.1, 1.0, size=(8, 9, 10))
    f = O.pack(g, il, 2, 0.5, 1.0)
    assert f.shape == (4, 5, 6, 16)
    assert np.allclose(f[..., 0], g[2:-2, 2:-2, 2:-2] * 0.5)
    assert np.allclose(f[..., 1], np.log10(il[2:-2, 2:-2, 2:-2] + 1e-6) / 6 + 1)
    assert not f[..., 2:].any()
