"""The oracle (CPU) and the GPU library against the committed golden fixtures
of tests/golden/ (written by tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["small_reference", "small_consistent"]


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(orc, name):
    import make_golden
    w, out = make_golden.compute(name)
    ref = _load(name)
    assert str(ref["digest"]) == make_golden.input_digest(w), "inputs changed"
    for k in ("obs", "syn", "res", "glam", "gmu", "grho"):
        assert np.array_equal(out[k], ref[k]), k
    assert out["misfit64"] == ref["misfit64"] and out["misfit32"] == ref["misfit32"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_golden(name):
    import torch
    import make_golden
    from efwi import synth
    from efwi.propagator import Propagator
    from util import rel_l2
    ref = _load(name)
    w = synth.small(**make_golden.CASES[name])
    assert str(ref["digest"]) == make_golden.input_digest(w), "inputs changed"
    print(f"[gpu] golden {name}", flush=True)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, obs, _ = P.run_shot(w.shots[0], 2)
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        P.zero_gradient()
        j, syn, res = P.run_shot(w.shots[0], 1, torch.from_numpy(ref["obs"]), want_syn=True, want_res=True)
        grads = [g.cpu().numpy() for g in P.gradient()]
    p = w.grid.npml
    inner = (slice(p, -p),) * 3
    got = dict(obs=obs.cpu().numpy(), syn=syn.cpu().numpy(), res=res.cpu().numpy(),
               glam=grads[0][inner], gmu=grads[1][inner], grho=grads[2][inner])
    for k, v in got.items():
        e = rel_l2(v, ref[k])
        print(f"  {k}: rel-L2 {e:.2e} bit-exact={np.array_equal(v, ref[k])}")
        assert e <= 1e-5, k
    assert abs(j - float(ref["misfit64"])) <= 1e-12 * float(ref["misfit64"])
