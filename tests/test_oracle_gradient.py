"""The oracle's whole gradient pipeline against centred finite differences.

In the consistent formulation (reference quirks corrected) with the C-PML
switched off, the reference's continuous-adjoint kernels are the exact
transposes of the forward operator in a homogeneous background, so the
gradient must match the finite-difference derivative of the misfit along any
direction supported in the interior. This pins the adjoint source, the
adjoint kernels, the reverse reconstruction and the imaging terms together,
including their time-level alignment. The measured constant is 1/2: the
reference injects r (not dJ/dsyn = -2 r) as adjoint source, so its gradient
is half the derivative of J = sum r^2 (Src/libCUFD.cu:420-441,
Src/utilities.cu:497-507).

With the reference formulation and the C-PML on, the gradient is not the
derivative of the discrete misfit (dsxy/dx typo, continuous-adjoint C-PML);
its finite-difference ratio is reported and only loosely bounded.
"""
import numpy as np
import pytest

N, NPML, NT = 36, 6, 160


def _setup(orc):
    from efwi import synth
    g = orc.Grid(N, N, N, NPML, NT, 10.0, 10.0, 10.0, 1e-3, 15.0)
    vp, vs, rho = 2500.0, 1450.0, 2.2
    shp = (N, N, N)
    base = {"lam": np.full(shp, (vp ** 2 - 2 * vs ** 2) * rho, np.float32),
            "mu": np.full(shp, vs ** 2 * rho, np.float32),
            "rho": np.full(shp, rho, np.float32)}
    stf = synth.ricker_integrated(15.0, NT, 1e-3).astype(np.float32)
    xs = np.arange(NPML, N - NPML, 3)
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    rec = np.stack([np.full(gx.size, NPML + 2), gx.ravel(), gy.ravel()], 1)
    shots = orc.Shots(src=np.array([[NPML + 3, N // 2, N // 2]]), stf=stf[None], rec=[rec])
    return g, base, shots


def _direction(rng):
    from scipy.ndimage import gaussian_filter
    d = gaussian_filter(rng.standard_normal((N, N, N)), 2.0)
    m = np.zeros((N, N, N))
    m[NPML:N - NPML, NPML:N - NPML, NPML:N - NPML] = 1
    d = d * m
    return d / np.abs(d).max()


def _fd_ratios(orc, formulation, cpml_off, seed=0):
    g, base, shots = _setup(orc)
    rng = np.random.default_rng(seed)
    true = dict(base)
    true["lam"] = (base["lam"] * (1 + 0.05 * _direction(rng))).astype(np.float32)
    obs = orc.run(g, shots, true["lam"], true["mu"], true["rho"], calc_id=2, precision="fp64",
                  formulation=formulation, cpml_off=cpml_off).syn
    obs = [o.astype(np.float32) for o in obs]
    r = orc.run(g, shots, base["lam"], base["mu"], base["rho"], obs=obs, calc_id=1, precision="fp64",
                formulation=formulation, cpml_off=cpml_off)
    out = {}
    for name, grad in (("lam", r.glam), ("mu", r.gmu), ("rho", r.grho)):
        d = _direction(rng)
        h = 1e-3

        def J(sign):
            m = dict(base)
            m[name] = base[name] * (1 + sign * h * d)
            return orc.run(g, shots, m["lam"], m["mu"], m["rho"], obs=obs, calc_id=0, precision="fp64",
                           formulation=formulation, cpml_off=cpml_off).misfit64
        fd = (J(+1) - J(-1)) / (2 * h)
        out[name] = float(np.sum(grad * base[name] * d)) / fd
    return out


def test_gradient_is_half_the_misfit_derivative(orc):
    # The model enters the oracle as float32, which quantises a 1e-3 relative
    # perturbation to ~1e-4; the bar reflects that.
    ratios = _fd_ratios(orc, "consistent", cpml_off=True)
    print("consistent, no C-PML: <g, dm> / FD =", ratios)
    for k, v in ratios.items():
        assert abs(v - 0.5) < 2e-3, (k, v)


def test_reference_formulation_fd_discrepancy_reported(orc):
    ratios = _fd_ratios(orc, "reference", cpml_off=False)
    print("reference formulation with C-PML: <g, dm> / FD =", ratios)
    for k, v in ratios.items():
        assert 0.3 < v < 0.7, (k, v)


def test_precision_modes_agree(orc):
    """fp32 (what the GPU computes), fp32 with the reference's FP64
    promotions, and fp64 runs of the same stencil on a heterogeneous case.

    With zero observed data the residual is the synthetic data itself, and
    the comparison measures the float error of the stencil pipeline alone
    (asserted below 1e-5). With observed data from a nearby model, the
    residual is a small difference of two large seismograms and the same
    absolute float error becomes a larger relative error of the residual and
    of the gradient (reported)."""
    from efwi import synth
    from util import oracle_run, rel_l2
    w = synth.small(n=32, npml=6, nt=140)
    obs_true = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    obs_zero = [np.zeros_like(o) for o in obs_true]
    for label, obs, bar in (("obs=0", obs_zero, 1e-5), ("obs=true model", obs_true, None)):
        res = {p: oracle_run(orc, w, obs=obs, calc_id=1, precision=p) for p in ("fp32", "fp32_literal", "fp64")}
        for p in ("fp32", "fp32_literal"):
            e = {k: rel_l2(getattr(res[p], k), getattr(res["fp64"], k)) for k in ("glam", "gmu", "grho")}
            e["syn"] = rel_l2(res[p].syn[0], res["fp64"].syn[0])
            e["res"] = rel_l2(res[p].res[0], res["fp64"].res[0])
            e["misfit"] = abs(res[p].misfit64 - res["fp64"].misfit64) / res["fp64"].misfit64
            print(f"{label}: {p} vs fp64:", {k: f"{v:.2e}" for k, v in e.items()})
            if bar is not None:
                assert max(e.values()) < bar, e
            else:
                assert e["syn"] < 1e-5, e
        # fp32 (the GPU's arithmetic) against the reference's literal
        # arithmetic, whose 2.0 * mu literals promote the normal-stress
        # updates and several adjoint terms to FP64 (Src/stress.cu:57-59,
        # 112-114,132, Src/stress_adj.cu:131-140, Src/velocity_adj.cu:72,100,129)
        lit = {k: rel_l2(getattr(res["fp32"], k), getattr(res["fp32_literal"], k)) for k in ("glam", "gmu", "grho")}
        lit["syn"] = rel_l2(res["fp32"].syn[0], res["fp32_literal"].syn[0])
        lit["res"] = rel_l2(res["fp32"].res[0], res["fp32_literal"].res[0])
        print(f"{label}: fp32 vs fp32_literal:", {k: f"{v:.2e}" for k, v in lit.items()})
        assert lit["syn"] < 1e-5, lit
        if bar is not None:
            # the pipeline's own float error: within the 1e-5 bar everywhere
            assert max(lit.values()) < bar, lit
        else:
            # a residual from a nearby model is a small difference of large
            # seismograms (|r| / |syn| ~ 1e-4 here): each arithmetic's rounding
            # is amplified alike, and the two float arithmetics differ from each
            # other by no more than they differ from the FP64 run
            for k in ("glam", "gmu", "grho", "res"):
                e32 = rel_l2(getattr(res["fp32"], k) if k != "res" else res["fp32"].res[0],
                             getattr(res["fp64"], k) if k != "res" else res["fp64"].res[0])
                elit = rel_l2(getattr(res["fp32_literal"], k) if k != "res" else res["fp32_literal"].res[0],
                              getattr(res["fp64"], k) if k != "res" else res["fp64"].res[0])
                assert lit[k] <= 1.5 * (e32 + elit), (k, lit[k], e32, elit)
                assert e32 <= 4.0 * elit, (k, e32, elit)


@pytest.mark.parametrize("every", [1, 7, 50])
def test_checkpointed_oracle_equals_full_ring(orc, every):
    """The oracle's checkpoint mode (segments replayed from saved forward
    states to rebuild their part of the boundary ring) gives the full ring's
    seismograms, residuals and gradients bit for bit; it lets the oracle run
    the 3000-step production grid, whose full ring exceeds host memory."""
    from efwi import synth
    from util import oracle_run
    w = synth.small(n=24, npml=5, nt=90)
    obs = [s.astype(np.float32) for s in oracle_run(orc, w, model=w.true, calc_id=2).syn]
    ring = oracle_run(orc, w, obs=obs, calc_id=1)
    g = w.grid
    og = orc.Grid(g.nz, g.nx, g.ny, g.npml, g.nt, g.dz, g.dx, g.dy, g.dt, g.f0)
    from util import oracle_shots
    ck = orc.run(og, oracle_shots(orc, w.shots), *w.model, obs=obs, calc_id=1, checkpoint_every=every)
    assert np.abs(ring.glam).max() > 0
    assert ck.misfit64 == ring.misfit64
    assert np.array_equal(ck.syn[0], ring.syn[0]) and np.array_equal(ck.res[0], ring.res[0])
    for a, b in ((ck.glam, ring.glam), (ck.gmu, ring.gmu), (ck.grho, ring.grho)):
        assert np.array_equal(a, b)
