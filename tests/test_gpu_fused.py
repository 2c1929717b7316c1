"""The fused forward step (stress + velocity of a time step in one z-marching
launch, efwi_set_fusion) against the two-kernel step and the oracle:
seismograms, residuals and the three gradients (whose backward pass starts
from the fused forward's final state and its boundary ring) are bit-identical,
for every workgroup shape, both formulations, a ragged y extent, a thin C-PML
(ring saved by a pass of its own), sources next to and outside the stencil box
and the checkpointed memory mode."""
import dataclasses

import numpy as np
import pytest
import torch

from util import oracle_run

pytestmark = pytest.mark.gpu


def _run(w, model, obs, fusion=None, calc_id=1):
    from efwi.propagator import Propagator
    print(f"[gpu] fused: grid={w.grid.shape} nt={w.grid.nt} form={w.grid.formulation} "
          f"mode={w.grid.memory_mode} fusion={fusion}", flush=True)
    with Propagator(w.grid) as P:
        if fusion:
            P.set_fusion(*fusion)
        P.set_model(*(torch.from_numpy(a) for a in model))
        P.zero_gradient()
        o = None if obs is None else torch.from_numpy(np.asarray(obs[0], np.float32))
        j, syn, res = P.run_shot(w.shots[0], calc_id, o, want_syn=True, want_res=calc_id != 2)
        out = dict(j=j, syn=syn.cpu().numpy(), res=None if res is None else res.cpu().numpy())
        if calc_id == 1:
            out["g"] = [g.cpu().numpy() for g in P.gradient()]
        torch.cuda.synchronize()
    return out


def _same(a, b):
    assert a["j"] == b["j"]
    assert np.array_equal(a["syn"], b["syn"])
    if a["res"] is not None:
        assert np.array_equal(a["res"], b["res"])
    for x, y in zip(a.get("g", []), b.get("g", [])):
        assert np.array_equal(x, y)


@pytest.fixture(scope="module")
def case(orc):
    from efwi import synth
    w = synth.small(n=32, npml=6, nt=140)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    return w, obs


@pytest.mark.parametrize("rows,zc", [(4, 16), (8, 7), (12, 5), (12, 64), (4, 1)])
def test_fused_forward_is_bit_identical(orc, case, rows, zc):
    w, obs = case
    ref = _run(w, w.model, obs)
    got = _run(w, w.model, obs, fusion=(rows, zc))
    _same(got, ref)
    o = oracle_run(orc, w, obs=obs, calc_id=1)
    assert np.array_equal(got["syn"], o.syn[0].astype(np.float32))
    assert np.array_equal(got["g"][0], o.glam.astype(np.float32))


@pytest.mark.parametrize("formulation", [0, 1])
def test_fused_forward_both_formulations(orc, formulation):
    from efwi import synth
    w = synth.small(n=28, npml=5, nt=120, formulation=formulation)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    _same(_run(w, w.model, obs, fusion=(8, 12)), _run(w, w.model, obs))


@pytest.mark.parametrize("n,npml,ny", [(24, 3, None), (26, 6, 71), (20, 6, 200)])
def test_fused_forward_thin_pml_ragged_and_long_rows(orc, n, npml, ny):
    from efwi import synth
    w = synth.small(n=n, npml=npml, nt=110, ny=ny)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    _same(_run(w, w.model, obs, fusion=(12, 9)), _run(w, w.model, obs))


@pytest.mark.parametrize("src_z", [1, 2, 5])
def test_fused_forward_sources_near_the_edge(orc, case, src_z):
    w, obs = case
    s0 = w.shots[0]
    w = dataclasses.replace(w, shots=[dataclasses.replace(s0, src=(src_z, s0.src[1], s0.src[2]))])
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    _same(_run(w, w.model, obs, fusion=(4, 16)), _run(w, w.model, obs))


def test_fused_forward_checkpointed(orc, case):
    w, obs = case
    wc = dataclasses.replace(w, grid=dataclasses.replace(w.grid, memory_mode=1, checkpoint_every=9))
    _same(_run(wc, w.model, obs, fusion=(8, 16)), _run(w, w.model, obs))


def test_fused_forward_repeated_shots_and_synthetic_only(case):
    """Two gradient shots in a row on one context (the state sets alternate
    across shots) and a synthetic-only run give the same results as the
    two-kernel step."""
    from efwi.propagator import Propagator
    w, obs = case
    o = torch.from_numpy(obs[0])
    outs = []
    for fusion in (None, (12, 16)):
        with Propagator(w.grid) as P:
            if fusion:
                P.set_fusion(*fusion)
            P.set_model(*(torch.from_numpy(a) for a in w.model))
            P.zero_gradient()
            j1, _, _ = P.run_shot(w.shots[0], 1, o)
            _, syn, _ = P.run_shot(w.shots[0], 2, None, want_syn=True)
            j2, _, _ = P.run_shot(w.shots[0], 1, o)
            outs.append((j1, j2, syn.cpu().numpy(), [g.cpu().numpy() for g in P.gradient()]))
    a, b = outs
    assert a[0] == b[0] and a[1] == b[1] and a[0] == a[1]
    assert np.array_equal(a[2], b[2])
    for x, y in zip(a[3], b[3]):
        assert np.array_equal(x, y)
