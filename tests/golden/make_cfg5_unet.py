"""Writes tests/golden/cfg5_unet_step.npz (run from the repository root; a few
minutes of CPU time, ~12 GB of memory):

    python tests/golden/make_cfg5_unet.py

One training step of the learned preconditioner's U-Net at the training
volume size 96x128x128 (3 levels, 16 base channels, 359,761 parameters) on the
CPU oracle's operations at the GPU's BF16 rounding points
(tests/unet_reference.py): seeded parameters (efwi.unet.init_params(1) with
seeded biases and head), a seeded 2-channel input and target, the masked
mean-squared loss; keeps every layer's weight and bias gradient and a summary
of the network output."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "efwi-3d_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"),
                os.path.dirname(os.path.abspath(__file__))]

DIMS = (96, 128, 128)
SEED = 9
OUT_LATTICE = slice(None, None, 97)


def inputs():
    """Parameters and data of the step, in the order the GPU test draws them."""
    import nn_oracle as O
    from efwi.unet import init_params
    rng = np.random.default_rng(SEED)
    V = DIMS[0] * DIMS[1] * DIMS[2]
    w, b, wh, bh = init_params(1)
    w = {k: v.numpy() for k, v in w.items()}
    b = {k: rng.uniform(-0.05, 0.1, v.numel()).astype(np.float32) for k, v in b.items()}
    wh = rng.standard_normal(16).astype(np.float32) * 0.25
    f0 = np.zeros((V, 16), np.float32)
    f0[:, :2] = rng.standard_normal((V, 2))
    target = rng.standard_normal(V).astype(np.float32)
    return w, b, wh, float(bh[0]), O.bf16_round(f0), target


def main():
    from unet_reference import oracle_network
    w, b, wh, bh, feat, target = inputs()
    V = feat.shape[0]
    out, G = oracle_network(w, b, wh, bh, feat, DIMS, target, np.ones(V, np.float32), V)
    out = np.asarray(out, np.float64).ravel()
    res = {f"grad_{k}": np.asarray(v, np.float32) for k, v in G.items()}
    res["out_sum"] = np.float64(out.sum())
    res["out_sq"] = np.float64((out ** 2).sum())
    res["out_lattice"] = out[OUT_LATTICE].astype(np.float32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg5_unet_step.npz")
    np.savez_compressed(path, **res)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
