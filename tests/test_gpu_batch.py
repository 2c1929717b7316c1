"""Batched shots (one persistent launch per stencil step over every (shot,
tile) item, each time step replayed from a captured hipGraph) against the
same shots run one by one: per-shot misfits, seismograms and gradients must
be bit-identical."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def three_shots():
    from efwi import synth
    from efwi.propagator import ShotSpec
    w = synth.small(n=32, npml=6, nt=140)
    base = w.shots[0]
    rng = np.random.default_rng(7)
    shots = [base]
    for k, side in enumerate((5, 11)):
        sx, sy = (int(v) for v in rng.integers(10, 34, size=2))
        xs = np.arange(sx - side // 2, sx - side // 2 + side)
        gx, gy = np.meshgrid(xs, xs + (sy - sx), indexing="ij")
        rec = np.stack([np.full(gx.size, 8), gx.ravel(), gy.ravel()], axis=1).astype(np.int32)
        shots.append(ShotSpec((10 + k, sx, sy), base.stf, rec))
    return w, shots


def _single(w, shots, model, obs=None, calc=1):
    from efwi.propagator import Propagator
    out = []
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in model))
        for k, s in enumerate(shots):
            P.zero_gradient()
            j, syn, _ = P.run_shot(s, calc, None if obs is None else obs[k], want_syn=True)
            g = [x.cpu().numpy() for x in P.gradient()] if calc == 1 else None
            out.append((j, syn.cpu().numpy(), g))
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single_shots(three_shots, use_graph):
    from efwi.propagator import Propagator
    w, shots = three_shots
    print(f"[gpu] batch of {len(shots)} shots, nrec {[s.nrec for s in shots]}, graph={use_graph}", flush=True)
    obs = [o[1] for o in _single(w, shots, w.true, calc=2)]
    ref = _single(w, shots, w.model, obs=obs)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        mis, syns = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs], want_syn=True, use_graph=use_graph)
        grads = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(shots))]
        # a second run reuses the captured graphs and must give the same
        mis2, _ = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs], use_graph=use_graph)
        torch.cuda.synchronize()
    for k, (j, syn, g) in enumerate(ref):
        assert mis[k] == j and mis2[k] == j, k
        assert np.array_equal(syns[k].cpu().numpy(), syn), k
        for a, b in zip(grads[k], g):
            assert np.abs(b).max() > 0
            assert np.array_equal(a, b), k


def test_batch_thin_pml_equals_single_shots():
    """A 3-cell C-PML: the boundary band leaves the stencil box, so the batched
    ring saves and restores are passes of their own (not folded into the
    batched kernels); still bit-identical to the shots run one by one."""
    from efwi import synth
    from efwi.propagator import Propagator, ShotSpec
    w = synth.small(n=24, npml=3, nt=100)
    base = w.shots[0]
    rec = np.array([[5, x, y] for x in range(6, 24, 3) for y in range(6, 24, 4)], np.int32)
    shots = [base, ShotSpec((9, 12, 15), base.stf, rec)]
    obs = [o[1] for o in _single(w, shots, w.true, calc=2)]
    ref = _single(w, shots, w.model, obs=obs)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        mis, syns = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs], want_syn=True)
        grads = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(shots))]
        torch.cuda.synchronize()
    for k, (j, syn, g) in enumerate(ref):
        assert mis[k] == j, k
        assert np.array_equal(syns[k].cpu().numpy(), syn), k
        for a, b in zip(grads[k], g):
            assert np.abs(b).max() > 0
            assert np.array_equal(a, b), k


def test_batch_forward_only(three_shots):
    from efwi.propagator import Propagator
    w, shots = three_shots
    ref = _single(w, shots, w.true, calc=2)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, syns = P.run_batch(shots, 2)
    for k in range(len(shots)):
        assert np.array_equal(syns[k].cpu().numpy(), ref[k][1])


@pytest.mark.parametrize("take,static_pct,tile", [(1, 0, (1, 1)), (3, 50, (2, 2)), (4, 100, (4, 1)), (64, 90, (1, 2))])
def test_batch_work_split(three_shots, take, static_pct, tile):
    """Every split of the queues between the static share and the atomic
    claims, with one- and multi-wave workgroups, covers each tile once."""
    from efwi.propagator import Propagator
    w, shots = three_shots
    print(f"[gpu] batch split take={take} static={static_pct}% tile={tile}", flush=True)
    obs = [o[1] for o in _single(w, shots, w.true, calc=2)]
    ref = _single(w, shots, w.model, obs=obs)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        for k in ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd"):
            P.set_tile(k, *tile)
        P.batch_config(take, static_pct)
        mis, syns = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs], want_syn=True)
        grads = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(shots))]
    for k, (j, syn, g) in enumerate(ref):
        assert mis[k] == j, k
        assert np.array_equal(syns[k].cpu().numpy(), syn), k
        for a, b in zip(grads[k], g):
            assert np.array_equal(a, b), k


@pytest.mark.parametrize("group", [1, 2, 3, 0, -1])
def test_batch_groups_equal_single_shots(three_shots, group):
    """Groups of 1, 2 or 3 shots per persistent launch (run one group after
    another, sharing the boundary rings of a group), the automatic choice and
    the single-shot kernels (-1, each shot on its own state): per-shot
    misfits, seismograms and gradients bit-identical to single shots, also on
    a second call with another mode."""
    from efwi.propagator import Propagator
    w, shots = three_shots
    print(f"[gpu] batch groups of {group}", flush=True)
    obs = [o[1] for o in _single(w, shots, w.true, calc=2)]
    ref = _single(w, shots, w.model, obs=obs)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        P.batch_group(group)
        mis, syns = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs], want_syn=True)
        grads = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(shots))]
        P.batch_group(3 - group if group in (1, 2) else (-1 if group >= 0 else 2))
        mis2, _ = P.run_batch(shots, 1, [torch.from_numpy(o) for o in obs])
        grads2 = [[x.cpu().numpy() for x in P.batch_gradient(k)] for k in range(len(shots))]
        torch.cuda.synchronize()
    for k, (j, syn, g) in enumerate(ref):
        assert mis[k] == j and mis2[k] == j, k
        assert np.array_equal(syns[k].cpu().numpy(), syn), k
        for a, b, c in zip(grads[k], grads2[k], g):
            assert np.array_equal(a, c) and np.array_equal(b, c), k
