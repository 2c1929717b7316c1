"""GPU (libefwi.so) against the CPU oracle on the same seeded inputs.

The GPU kernels evaluate the reference's float expressions in the same order
with IEEE division and no fused multiply-add, so seismograms, adjoint sources
and gradients are expected to agree with the oracle's fp32 mode bit for bit;
the asserted bar is the north-star tolerance, relative L2 <= 1e-5, and the
bit-exact agreement is reported. The misfit is an FP64 sum whose order
differs from the oracle's, so it is compared to 1e-12.
"""
import numpy as np
import pytest
import torch

from util import oracle_run, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _gpu_run(w, model, calc_id, obs=None, tiles=None, resident=None):
    from efwi.propagator import Propagator
    print(f"[gpu] {w.name} grid={w.grid.shape} nt={w.grid.nt} form={w.grid.formulation} calc={calc_id} "
          f"tiles={tiles} resident={resident}", flush=True)
    with Propagator(w.grid) as P:
        for k, (r, z) in (tiles or {}).items():
            P.set_tile(k, r, z)
        if resident is not None:
            P.set_resident_segments(resident)
        P.set_model(*(torch.from_numpy(a) for a in model))
        P.zero_gradient()
        total = 0.0
        syns, ress = [], []
        for k, s in enumerate(w.shots):
            o = None if obs is None else torch.from_numpy(np.asarray(obs[k], np.float32))
            j, syn, res = P.run_shot(s, calc_id, o, want_syn=True, want_res=calc_id != 2)
            syns.append(syn.cpu().numpy())
            ress.append(None if res is None else res.cpu().numpy())
            if j is not None:
                total += j
        grads = [g.cpu().numpy() for g in P.gradient()] if calc_id == 1 else None
        torch.cuda.synchronize()
    return total, syns, ress, grads


@pytest.fixture(scope="module")
def small_case(orc):
    from efwi import synth
    w = synth.small(n=32, npml=6, nt=140)
    ref_true = oracle_run(orc, w, model=w.true, calc_id=2)
    obs = [s.astype(np.float32) for s in ref_true.syn]
    return w, obs


def test_forward_matches_oracle(orc, small_case):
    w, _ = small_case
    _, syn, _, _ = _gpu_run(w, w.true, 2)
    ref = oracle_run(orc, w, model=w.true, calc_id=2)
    err = rel_l2(syn[0], ref.syn[0])
    print(f"forward rel-L2 {err:.3e}, bit-exact={np.array_equal(syn[0], ref.syn[0].astype(np.float32))}")
    assert np.abs(syn[0]).max() > 0
    assert err <= TOL


@pytest.mark.parametrize("formulation", [0, 1])
def test_gradient_matches_oracle(orc, small_case, formulation):
    from efwi import synth
    w0, obs = small_case
    w = synth.small(n=32, npml=6, nt=140, formulation=formulation)
    j, syn, res, grads = _gpu_run(w, w.model, 1, obs)
    ref = oracle_run(orc, w, model=w.model, obs=obs, calc_id=1)
    assert abs(j - ref.misfit64) <= 1e-12 * abs(ref.misfit64)
    assert rel_l2(syn[0], ref.syn[0]) <= TOL
    assert rel_l2(res[0], ref.res[0]) <= TOL
    for name, a, b in zip(("lambda", "mu", "rho"), grads, (ref.glam, ref.gmu, ref.grho)):
        e = rel_l2(a, b)
        print(f"grad {name}: rel-L2 {e:.3e} bit-exact={np.array_equal(a, b.astype(np.float32))}")
        assert np.abs(b).max() > 0
        assert e <= TOL, name


@pytest.mark.parametrize("rows,zdepth", [(1, 1), (4, 2)])
def test_long_y_axis_is_bit_identical(orc, rows, zdepth):
    """A grid 212 cells long in y (200 + 2 x 6 C-PML): four 64-cell y
    segments per row, two of them wholly in the y interior (the wavefronts
    that take the no-C-PML code path in every kernel), with the source in
    the second segment, so waves cross segment ends both ways. Seismograms,
    residuals and all three gradients must equal the oracle's bit for bit."""
    from efwi import synth
    w = synth.small(n=20, npml=6, nt=150, ny=200)
    assert w.grid.shape[2] == 212
    ref_true = oracle_run(orc, w, model=w.true, calc_id=2)
    obs = [s.astype(np.float32) for s in ref_true.syn]
    t = {k: (rows, zdepth) for k in ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd")}
    j, syn, res, grads = _gpu_run(w, w.model, 1, obs, tiles=t)
    ref = oracle_run(orc, w, model=w.model, obs=obs, calc_id=1)
    assert abs(j - ref.misfit64) <= 1e-12 * abs(ref.misfit64)
    assert np.array_equal(syn[0], ref.syn[0].astype(np.float32))
    assert np.array_equal(res[0], ref.res[0].astype(np.float32))
    for name, a, b in zip(("lambda", "mu", "rho"), grads, (ref.glam, ref.gmu, ref.grho)):
        b = b.astype(np.float32)
        print(f"grad {name}: max |diff| {np.abs(a - b).max():.3e} of max {np.abs(b).max():.3e}")
        assert np.abs(b[:, :, 64:192]).max() > 0
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("n,npml,ny", [(24, 3, None), (26, 6, 71)])
def test_thin_pml_and_ragged_rows_are_bit_identical(orc, n, npml, ny):
    """A 3-cell C-PML (the boundary band reaches the outer two cells, outside
    the stencil kernels' box, so every ring save and restore is a pass of its
    own) and a y extent of 83 padded cells (rows padded to 96, the last
    64-cell segment partly past the row's end, the band's far y slab in it).
    Seismograms, residuals and gradients must equal the oracle's bit for bit."""
    from efwi import synth
    w = synth.small(n=n, npml=npml, nt=120, ny=ny)
    ref_true = oracle_run(orc, w, model=w.true, calc_id=2)
    obs = [s.astype(np.float32) for s in ref_true.syn]
    j, syn, res, grads = _gpu_run(w, w.model, 1, obs)
    ref = oracle_run(orc, w, model=w.model, obs=obs, calc_id=1)
    assert abs(j - ref.misfit64) <= 1e-12 * abs(ref.misfit64)
    assert np.array_equal(syn[0], ref.syn[0].astype(np.float32))
    assert np.array_equal(res[0], ref.res[0].astype(np.float32))
    for name, a, b in zip(("lambda", "mu", "rho"), grads, (ref.glam, ref.gmu, ref.grho)):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b.astype(np.float32)), name


@pytest.mark.parametrize("src_z", [1, 5])
def test_source_near_the_edge_is_bit_identical(orc, src_z):
    """Sources in the C-PML (padded z 5 of 6 C-PML cells: injected by the
    stencil kernel going forward, by the separate source kernel going back)
    and outside the stencil box (padded z 1: the separate source kernel both
    ways, so the stress fields' boundary save is a pass of its own after it).
    Seismograms, residuals and gradients must equal the oracle's bit for bit."""
    import dataclasses
    from efwi import synth
    w = synth.small(n=32, npml=6, nt=140)
    s0 = w.shots[0]
    w = dataclasses.replace(w, shots=[dataclasses.replace(s0, src=(src_z, s0.src[1], s0.src[2]))])
    ref_true = oracle_run(orc, w, model=w.true, calc_id=2)
    obs = [s.astype(np.float32) for s in ref_true.syn]
    j, syn, res, grads = _gpu_run(w, w.model, 1, obs)
    ref = oracle_run(orc, w, model=w.model, obs=obs, calc_id=1)
    assert abs(j - ref.misfit64) <= 1e-12 * abs(ref.misfit64)
    assert np.abs(ref.syn[0]).max() > 0
    assert np.array_equal(syn[0], ref.syn[0].astype(np.float32))
    assert np.array_equal(res[0], ref.res[0].astype(np.float32))
    for name, a, b in zip(("lambda", "mu", "rho"), grads, (ref.glam, ref.gmu, ref.grho)):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b.astype(np.float32)), name


@pytest.mark.parametrize("every", [7, 0])
def test_checkpointed_ring_is_bit_identical(orc, small_case, every):
    """memory_mode 1 (checkpoints every `every` steps, every segment but the
    last replayed) gives the same gradient as the full ring, bit for bit."""
    import dataclasses
    from efwi import synth
    w0, obs = small_case
    w = synth.small(n=32, npml=6, nt=140)
    _, _, _, g_ring = _gpu_run(dataclasses.replace(w, grid=dataclasses.replace(w.grid, memory_mode=0)), w.model, 1, obs)
    wc = dataclasses.replace(w, grid=dataclasses.replace(w.grid, memory_mode=1, checkpoint_every=every))
    j, _, _, g_ck = _gpu_run(wc, w.model, 1, obs, resident=1)
    for a, b in zip(g_ck, g_ring):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("resident", [2, 11, 19, 20, 0])
def test_resident_segments_are_bit_identical(orc, small_case, resident):
    """memory_mode 1 with the last `resident` of 20 segments (139 steps,
    segments of 7) keeping the forward pass's ring and the earlier ones
    replayed into the first segment's slots; 0 = as many as the device holds
    (all of them on this grid): the full ring's misfit, seismograms and
    gradients, bit for bit."""
    import dataclasses
    from efwi import synth
    from efwi.propagator import Propagator
    w0, obs = small_case
    w = synth.small(n=32, npml=6, nt=140)
    ref = _gpu_run(dataclasses.replace(w, grid=dataclasses.replace(w.grid, memory_mode=0)), w.model, 1, obs)
    wc = dataclasses.replace(w, grid=dataclasses.replace(w.grid, memory_mode=1, checkpoint_every=7))
    got = _gpu_run(wc, w.model, 1, obs, resident=resident)
    with Propagator(wc.grid) as P:
        P.set_resident_segments(resident)
        r, n = P.resident_segments()
    assert n == 20 and r == (resident or 20)
    assert got[0] == ref[0]
    for a, b in zip(got[1], ref[1]):
        assert np.array_equal(a, b)
    for a, b in zip(got[3], ref[3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("rows,zdepth", [(2, 3), (4, 8), (1, 64), (4, 16), (8, 16), (8, 40)])
def test_tile_shapes_are_bit_identical(orc, small_case, rows, zdepth):
    """Launch shapes only remap cells to threads (zdepth >= 16: the z-marching
    kernels with register queues and LDS tiles): gradients and seismograms are
    bit-identical to the default shape."""
    from efwi import synth
    w0, obs = small_case
    w = synth.small(n=32, npml=6, nt=140)
    ref = _gpu_run(w, w.model, 1, obs)
    t = {k: (rows, zdepth) for k in ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd")}
    got = _gpu_run(w, w.model, 1, obs, tiles=t)
    assert got[0] == ref[0]
    assert np.array_equal(got[1][0], ref[1][0])
    for a, b in zip(got[3], ref[3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("rows,zdepth", [(4, 16), (8, 16), (2, 40)])
def test_marching_shapes_on_several_y_segments(orc, rows, zdepth):
    """The z-marching kernels on a grid 212 cells long in y (four 64-cell
    segments: every segment edge needs its neighbour segment's cells from the
    LDS tile's side halo), forward and gradient against the oracle bit for
    bit."""
    from efwi import synth
    w = synth.small(n=20, npml=6, nt=120, ny=200)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    ref = oracle_run(orc, w, obs=obs, calc_id=1)
    t = {k: (rows, zdepth) for k in ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd")}
    print(f"[gpu] marching shapes {rows}x{zdepth} on a 212-cell y axis", flush=True)
    j, syn, res, grads = _gpu_run(w, w.model, 1, obs, tiles=t)
    assert np.array_equal(syn[0], ref.syn[0].astype(np.float32))
    assert abs(j - ref.misfit64) <= 1e-12 * abs(ref.misfit64)
    for g, c in zip(grads, (ref.glam, ref.gmu, ref.grho)):
        assert np.abs(c).max() > 0
        assert np.array_equal(g, c.astype(np.float32))


def test_autotuner_picks_a_candidate(small_case):
    import torch
    from efwi import synth
    from efwi.propagator import Propagator
    w = synth.small(n=32, npml=6, nt=20)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        rep = P.autotune(reps=2)
        tiles = P.tiles()
    for k, times in rep.items():
        assert len(times) == 12 and all(t > 0 for t in times.values())
        best = min(times, key=times.get)
        assert tiles[k] == best, (k, tiles[k], best)


@pytest.mark.parametrize("formulation,tile", [(0, (1, 1)), (1, (2, 2)), (0, (4, 3))])
def test_backward_split_is_bit_identical(orc, formulation, tile):
    """The interior/rim split of the backward kernels (efwi_set_split): a grid
    212 cells long in y (segments 1 and 2 wholly in the y interior, the source
    among them) with the split on, against the oracle and the unsplit run,
    seismograms, residuals and gradients bit for bit, at flat shapes that
    march several planes per thread."""
    from efwi import synth
    from efwi.propagator import Propagator
    w = synth.small(n=20, npml=6, nt=120, ny=200, formulation=formulation)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    outs = []
    for split in (False, True):
        print(f"[gpu] backward split={split} form={formulation} tile={tile}", flush=True)
        with Propagator(w.grid) as P:
            for k in ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd"):
                P.set_tile(k, *tile)
            P.set_split(split)
            assert P.split() == split
            P.set_model(*(torch.from_numpy(a) for a in w.model))
            P.zero_gradient()
            j, syn, res = P.run_shot(w.shots[0], 1, torch.from_numpy(obs[0]), want_syn=True, want_res=True)
            outs.append((j, res.cpu().numpy(), [g.cpu().numpy() for g in P.gradient()]))
            torch.cuda.synchronize()
    (ja, ra, ga), (jb, rb, gb) = outs
    ref = oracle_run(orc, w, obs=obs, calc_id=1)
    assert ja == jb
    assert np.array_equal(ra, rb)
    for a, b, c in zip(ga, gb, (ref.glam, ref.gmu, ref.grho)):
        assert np.abs(c[:, :, 64:192]).max() > 0
        assert np.array_equal(a, b)
        assert np.array_equal(b, c.astype(np.float32))
