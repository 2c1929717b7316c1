"""The benchmark's workloads at their own sizes on the GPU:

  * the production grid (200x200x120 + 20 C-PML cells) for all 3000 steps
    against the oracle's summary of the same run (tests/golden/
    cfg2_full_summary.npz, the oracle in its checkpoint mode);
  * the 16 batched shots (84x180x180 padded, apertures 30-140 cells): every
    shot's batched gradient and misfit bit-identical to its single-shot run,
    two of them against the oracle's summaries (cfg3_summary.npz);
  * the checkpointed 320x320x160 grid (4000 steps, 4 shots): finite gradients
    that vanish in the C-PML, the device memory against its budget, and
    checkpointing bit-identical to the full ring where that fits;
  * one U-Net training step at 96x128x128 against the oracle's network
    (cfg5_unet_step.npz)."""
import dataclasses
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / n if n > 0 else np.linalg.norm(a.ravel())


def _gradient_run(P, shot, obs):
    P.zero_gradient()
    j, syn, res = P.run_shot(shot, 1, obs, want_syn=True, want_res=True)
    g = [x.cpu().numpy() for x in P.gradient()]
    return j, syn.cpu().numpy(), res.cpu().numpy(), g


def _check_summary(got, ref, prefix=""):
    keys = [k[len(prefix):] for k in ref.files if k.startswith(prefix)]
    assert abs(got["misfit64"] - ref[prefix + "misfit64"]) <= 1e-9 * abs(ref[prefix + "misfit64"])
    for k in keys:
        e = _rel(got[k], ref[prefix + k])
        print(f"{prefix}{k}: rel-L2 {e:.3e}")
        assert e <= TOL, k
    assert np.array_equal(got["traces"], ref[prefix + "traces"])
    for k in keys:
        if k.endswith("_lattice"):
            assert np.array_equal(got[k], ref[prefix + k]), k


def test_cfg2_all_3000_steps_match_oracle():
    from efwi import synth
    from efwi.propagator import Propagator
    sys.path.insert(0, GOLDEN)
    from make_cfg2_full_summary import summarize
    path = os.path.join(GOLDEN, "cfg2_full_summary.npz")
    w = synth.cfg2()
    print(f"[gpu] cfg2 grid={w.grid.shape} nt={w.grid.nt} nrec={w.shots[0].nrec}", flush=True)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, obs, _ = P.run_shot(w.shots[0], 2, None, want_syn=True)
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        j, syn, res, g = _gradient_run(P, w.shots[0], obs)
        obs = obs.cpu().numpy()
        torch.cuda.synchronize()
    got = summarize(syn, res, j, *g)
    got["obs_trace_energy"] = np.sum(np.asarray(obs, np.float64) ** 2, axis=2)
    for x in g:
        assert np.isfinite(x).all()
    _check_summary(got, np.load(path))
    p64 = os.path.join(GOLDEN, "cfg2_full_summary_fp64.npz")
    if os.path.exists(p64):
        r64 = np.load(p64)
        for k in ("trace_energy", "traces", "res_energy", "glam_lattice", "gmu_lattice", "grho_lattice"):
            print(f"{k}: FP32 (GPU) against FP64 (oracle) rel-L2 {_rel(got[k], r64[k]):.3e}")


def test_cfg3_batched_shots_match_single_shots_and_oracle():
    from efwi import synth
    from efwi.propagator import Propagator
    sys.path.insert(0, GOLDEN)
    from make_cfg3_summary import SHOTS, summarize
    w = synth.cfg3()
    print(f"[gpu] cfg3 grid={w.grid.shape} nt={w.grid.nt} shots={len(w.shots)} "
          f"nrec={min(s.nrec for s in w.shots)}..{max(s.nrec for s in w.shots)}", flush=True)
    ref = np.load(os.path.join(GOLDEN, "cfg3_summary.npz"))
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, obs = P.run_batch(w.shots, 2, want_syn=True)
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        mis, _ = P.run_batch(w.shots, 1, obs)
        gb = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(w.shots))]
        print("[gpu] cfg3 batch done", flush=True)
        for k, s in enumerate(w.shots):
            j, syn, res, g = _gradient_run(P, s, obs[k])
            assert j == mis[k], k
            for a, b in zip(g, gb[k]):
                assert np.array_equal(a, b), k
            if k in SHOTS:
                _check_summary(summarize(syn, res, j, *g), ref, prefix=f"shot{k}_")
        torch.cuda.synchronize()


def test_cfg4_checkpointed_full_grid():
    from efwi import synth
    from efwi.propagator import Propagator
    w = synth.cfg4()
    g = w.grid
    assert g.memory_mode == 1
    print(f"[gpu] cfg4 grid={g.shape} nt={g.nt} shots={len(w.shots)}", flush=True)
    # the checkpoints, the resident part of the ring (as many of the last
    # segments as 85 % of the free memory holds) and the state: within the
    # device and below the whole ring alone
    budget = torch.cuda.get_device_properties(0).total_memory
    with Propagator(g) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        obs = [P.run_shot(s, 2, None, want_syn=True)[1] for s in w.shots]
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        P.zero_gradient()
        total = 0.0
        for s, o in zip(w.shots, obs):
            j, _, _ = P.run_shot(s, 1, o)
            total += j
        held = P.device_bytes
        res, nseg = P.resident_segments()
        grads = [x.cpu().numpy() for x in P.gradient()]
        torch.cuda.synchronize()
    full_ring = 9 * 4 * (g.nt - 1) * _band_cells(g)
    print(f"[gpu] cfg4 misfit {total:.6e}, device memory {held / 1e9:.1f} GB, {res} of {nseg} segments resident "
          f"(full ring alone {full_ring / 1e9:.0f} GB)", flush=True)
    assert 1 <= res < nseg
    assert held < budget and held < full_ring
    p = g.npml
    for x in grads:
        assert np.isfinite(x).all()
        assert np.abs(x[p:-p, p:-p, p:-p]).max() > 0
        inner = np.zeros(x.shape, bool)
        inner[p:-p, p:-p, p:-p] = True
        assert not np.any(x[~inner])


def _band_cells(g):
    lo = [g.npml - 2] * 3
    hi = [n - g.npml + 1 for n in g.shape]
    e = [h - l + 1 for l, h in zip(lo, hi)]
    return 10 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]) - 100 * (e[0] + e[1] + e[2]) + 1000


def test_cfg4_grid_checkpointed_equals_full_ring():
    """On the same 320x320x160 grid, 160 steps (where the full ring fits):
    memory mode 1 with segments of 13 steps, the last 3 of 13 resident and the
    others replayed, gives the full ring's gradients and misfits bit for bit,
    over all 4 shots."""
    from efwi import synth
    from efwi.propagator import Propagator
    w = synth.cfg4(nt=160)
    outs = []
    for mode, every in ((0, 0), (1, 13)):
        g = dataclasses.replace(w.grid, memory_mode=mode, checkpoint_every=every)
        print(f"[gpu] cfg4 grid nt=160 memory_mode={mode}", flush=True)
        with Propagator(g) as P:
            if mode == 1:
                P.set_resident_segments(3)
                assert P.resident_segments() == (3, 13)
            P.set_model(*(torch.from_numpy(a) for a in w.true))
            obs = [P.run_shot(s, 2, None, want_syn=True)[1] for s in w.shots]
            P.set_model(*(torch.from_numpy(a) for a in w.model))
            P.zero_gradient()
            js = [P.run_shot(s, 1, o)[0] for s, o in zip(w.shots, obs)]
            outs.append((js, [x.cpu().numpy() for x in P.gradient()]))
            torch.cuda.synchronize()
    (ja, ga), (jb, gb) = outs
    assert ja == jb
    assert np.abs(ga[0]).max() > 0
    for a, b in zip(ga, gb):
        assert np.array_equal(a, b)


def test_cfg5_unet_training_step_matches_oracle():
    from efwi.unet import UNet3D
    sys.path.insert(0, GOLDEN)
    from make_cfg5_unet import DIMS, OUT_LATTICE, inputs
    ref = np.load(os.path.join(GOLDEN, "cfg5_unet_step.npz"))
    w, b, wh, bh, feat, target = inputs()
    V = feat.shape[0]
    net = UNet3D(0, seed=1)
    for k in net.w:
        assert np.array_equal(net.w[k].cpu().numpy(), w[k]), k
    for k in net.b:
        net.b[k].copy_(torch.from_numpy(b[k]))
    net.wh.copy_(torch.from_numpy(wh))
    print(f"[gpu] U-Net training step at {DIMS}, {net.nparams()} parameters", flush=True)
    x = torch.from_numpy(feat).to("cuda:0").to(torch.bfloat16)
    out = net.forward(x, DIMS)
    loss, dout = net.loss(out, torch.from_numpy(target).to("cuda:0"), torch.ones(V, device="cuda:0"), V)
    net.backward(dout)
    torch.cuda.synchronize()
    o = out.cpu().numpy().astype(np.float64).ravel()
    assert abs(o.sum() - ref["out_sum"]) <= 5e-3 * np.sqrt(ref["out_sq"] * V)
    assert _rel(o[OUT_LATTICE], ref["out_lattice"]) <= 5e-3
    worst = {}
    for name in ref.files:
        if not name.startswith("grad_"):
            continue
        k, r = name[5:], ref[name]
        worst[k] = _rel(net.grads[k].cpu().numpy().reshape(r.shape), r)
    assert len(worst) == 21
    print("[gpu] relative L2 of the weight gradients:", {k: round(v, 5) for k, v in worst.items()}, flush=True)
    for k, e in worst.items():
        assert e < 5e-3, (k, e)
