"""The learned preconditioner's U-Net on the CPU oracle's operations
(oracle/nn_oracle.py), at the GPU's BF16 rounding points: the reference of
the network-level GPU tests (test infrastructure)."""
import numpy as np

import nn_oracle as O


def oracle_network(w, b, wh, bh, x, dims, target, mask, count):
    """The U-Net's forward and weight gradients with the oracle's operations,
    rounding to BF16 where the GPU stores BF16. w, b: per-layer FP32 weights
    [27][cout][cin] and biases; wh, bh: the head; x: the BF16 input
    [nz][nx][ny][16] as float32."""
    r = O.bf16_round
    W = {k: r(np.asarray(v, np.float32)) for k, v in w.items()}
    Bs = {k: np.asarray(v, np.float32) for k, v in b.items()}
    f = lambda a: r(a.astype(np.float32))
    x = np.asarray(x, np.float32).reshape(dims + (16,))
    conv = lambda n, a: f(O.conv3(a, W[n], Bs[n], relu=True))
    a1 = conv("enc1a", x)
    e1 = conv("enc1b", a1)
    p1 = f(O.pool2(e1))
    a2 = conv("enc2a", p1)
    e2 = conv("enc2b", a2)
    p2 = f(O.pool2(e2))
    a3 = conv("bot_a", p2)
    b3 = conv("bot_b", a3)
    cat2 = np.concatenate([O.up2(b3), e2], axis=3)
    d2a = conv("dec2a", cat2)
    d2b = conv("dec2b", d2a)
    cat1 = np.concatenate([O.up2(d2b), e1], axis=3)
    d1a = conv("dec1a", cat1)
    d1b = conv("dec1b", d1a)
    o = O.head(d1b, wh, bh).astype(np.float32)
    out = O.gate(x, o).astype(np.float32)
    _, dout = O.mse(out, target.reshape(dims), mask.reshape(dims), count)
    do = O.gate_bwd(o, out, dout.astype(np.float32)).astype(np.float32)
    G = {}
    g1b, G["head.w"], _ = O.head_bwd(d1b, wh, do)
    g1b = f(g1b)

    def layer(n, gpre, src, mask_src):
        G[n + ".w"] = O.conv3_wgrad(gpre, src)
        G[n + ".b"] = gpre.reshape(-1, gpre.shape[3]).sum(0)
        if mask_src is None:
            return f(O.conv3_dgrad(gpre, W[n]))
        return f(O.conv3_dgrad(gpre, W[n]) * (mask_src > 0))
    g = layer("dec1b", g1b, d1a, d1a)
    gcat1 = layer("dec1a", g, cat1, cat1)
    g = f(O.up2_bwd(gcat1[..., :32]))
    g = layer("dec2b", g, d2a, d2a)
    gcat2 = layer("dec2a", g, cat2, cat2)
    g = f(O.up2_bwd(gcat2[..., :64]))
    g = layer("bot_b", g, a3, a3)
    gp2 = layer("bot_a", g, p2, None)
    ge2 = f(f(gcat2[..., 64:] + O.pool2_bwd(gp2)) * (e2 > 0))
    g = layer("enc2b", ge2, a2, a2)
    gp1 = layer("enc2a", g, p1, None)
    ge1 = f(f(gcat1[..., 32:] + O.pool2_bwd(gp1)) * (e1 > 0))
    g = layer("enc1b", ge1, a1, a1)
    layer("enc1a", g, x, None)
    return out, G
