"""Full-size plain-method workload (64^3 physical + 10 PML cells, 500 steps,
4096 receivers) on the GPU against the oracle's summary of the same run
(tests/golden/cfg1_summary.npz, written by make_cfg1_summary.py), plus
size-independent properties at that size, and the FP64 dot product that the
L-BFGS driver uses."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    n = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / n if n > 0 else np.linalg.norm(a.ravel())


@pytest.fixture(scope="module")
def cfg1_gpu():
    from efwi import synth
    from efwi.propagator import Propagator
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_cfg1_summary import summarize
    w = synth.cfg1()
    print(f"[gpu] cfg1 grid={w.grid.shape} nt={w.grid.nt} nrec={len(w.shots[0].rec)}", flush=True)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, obs, _ = P.run_shot(w.shots[0], 2, None, want_syn=True)
        obs = obs.cpu().numpy()
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        P.zero_gradient()
        j, syn, res = P.run_shot(w.shots[0], 1, torch.from_numpy(obs), want_syn=True, want_res=True)
        g = [x.cpu().numpy() for x in P.gradient()]
        torch.cuda.synchronize()
    out = summarize(syn.cpu().numpy(), res.cpu().numpy(), j, *g)
    out["obs_trace_energy"] = np.sum(np.asarray(obs, np.float64) ** 2, axis=2)
    return out, g, dict(obs=obs, syn=syn.cpu().numpy(), res=res.cpu().numpy(), misfit=j)


def test_cfg1_matches_oracle_summary(cfg1_gpu):
    got, _, _ = cfg1_gpu
    ref = np.load(os.path.join(HERE, "golden", "cfg1_summary.npz"))
    assert abs(got["misfit64"] - ref["misfit64"]) <= 1e-9 * abs(ref["misfit64"])
    for k in ref.files:
        e = _rel(got[k], ref[k])
        print(f"{k}: rel-L2 {e:.3e}")
        assert e <= TOL, k
    assert np.array_equal(got["traces"], ref["traces"])


def test_cfg1_gradient_is_finite_and_localised(cfg1_gpu):
    """The anomaly sits at the grid centre: the lambda gradient is finite and
    zero in the PML (reverse kernels cover the interior only)."""
    _, (glam, gmu, grho), _ = cfg1_gpu
    for g in (glam, gmu, grho):
        assert np.all(np.isfinite(g))
    p = 10
    assert np.all(glam[:p - 0] == 0) and np.all(glam[-p:] == 0)
    assert np.abs(glam[p:-p, p:-p, p:-p]).max() > 0


def test_cfg1_against_literal_reference_arithmetic(orc, cfg1_gpu):
    """The GPU (fp32 arithmetic) at the full 64^3 workload against the oracle
    run with the reference's literal arithmetic, in which the 2.0 * mu
    literals promote the normal-stress updates and several adjoint terms to
    FP64 (Src/stress.cu:57-59,112-114,132, Src/stress_adj.cu:131-140,
    Src/velocity_adj.cu:72,100,129), and against an FP64 run, with the
    observed data of the anomaly-free model. Seismograms agree within 1e-5.
    The residual is a small difference of large seismograms (|r| / |syn| ~
    3e-3 here), so every float arithmetic's rounding is amplified alike in it
    and in the gradients: the GPU and the literal arithmetic differ from each
    other by about as much as each differs from FP64 (printed, and bounded
    by that)."""
    from efwi import synth
    from util import oracle_run
    _, g, d = cfg1_gpu
    w = synth.cfg1()
    obs = [d["obs"].astype(np.float32)]
    print("[gpu] cfg1 oracle runs: fp32_literal, fp64", flush=True)
    runs = {p: oracle_run(orc, w, obs=obs, calc_id=1, precision=p) for p in ("fp32_literal", "fp64")}
    gpu = dict(syn=d["syn"], res=d["res"], glam=g[0], gmu=g[1], grho=g[2])

    def fields(r):
        return dict(syn=r.syn[0], res=r.res[0], glam=r.glam, gmu=r.gmu, grho=r.grho)

    lit, f64 = fields(runs["fp32_literal"]), fields(runs["fp64"])
    e_lit = {k: _rel(gpu[k], lit[k]) for k in gpu}
    e_gpu64 = {k: _rel(gpu[k], f64[k]) for k in gpu}
    e_lit64 = {k: _rel(lit[k], f64[k]) for k in gpu}
    for k in gpu:
        print(f"{k}: gpu vs literal {e_lit[k]:.2e}  gpu vs fp64 {e_gpu64[k]:.2e}  literal vs fp64 {e_lit64[k]:.2e}")
    print(f"misfit: gpu {d['misfit']:.9e} literal {runs['fp32_literal'].misfit64:.9e} fp64 {runs['fp64'].misfit64:.9e}")
    assert e_lit["syn"] <= TOL and e_gpu64["syn"] <= TOL
    for k in ("res", "glam", "gmu", "grho"):
        assert e_lit[k] <= 1.5 * (e_gpu64[k] + e_lit64[k]), k
        assert e_gpu64[k] <= 4.0 * e_lit64[k], k
        assert e_lit[k] <= 2e-2, k


@pytest.mark.parametrize("n", [1, 255, 256 * 1024 + 7, 5 * 10 ** 6])
def test_dot64_matches_numpy(n):
    from efwi._native import dot64
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    got = dot64(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    ref = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    # each product of two floats is exact in FP64; only the summation order differs
    assert abs(got - ref) <= 1e-12 * max(1.0, float(np.sum(np.abs(a.astype(np.float64) * b))))


def test_cfg2_first_300_steps_match_oracle_summary():
    """The production grid (200x200x120 physical + 20 C-PML cells, padded
    160x240x240, 10,000 receivers) for its first 300 steps against the fp32
    oracle's summary (tests/golden/cfg2_n300_summary.npz, written by
    make_cfg2_summary.py): seismograms, residuals, misfit and the gradients'
    per-plane sums. The GPU path is bit-identical to the oracle's fp32 mode,
    so the selected traces must be equal and every summary within 1e-5
    relative L2 (in 300 steps the anomalies' reflections have not reached the
    surface: the misfit is small, and the comparison is of the arithmetic)."""
    import sys
    from efwi import synth
    from efwi.propagator import Propagator
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_cfg2_summary import NT, summarize
    w = synth.cfg2(nt=NT)
    print(f"[gpu] cfg2 grid={w.grid.shape} nt={w.grid.nt} nrec={len(w.shots[0].rec)}", flush=True)
    with Propagator(w.grid) as P:
        P.set_model(*(torch.from_numpy(a) for a in w.true))
        _, obs, _ = P.run_shot(w.shots[0], 2, None, want_syn=True)
        P.set_model(*(torch.from_numpy(a) for a in w.model))
        P.zero_gradient()
        j, syn, res = P.run_shot(w.shots[0], 1, obs, want_syn=True, want_res=True)
        g = [x.cpu().numpy() for x in P.gradient()]
        torch.cuda.synchronize()
    got = summarize(syn.cpu().numpy(), res.cpu().numpy(), j, *g)
    got["obs_trace_energy"] = np.sum(np.asarray(obs.cpu().numpy(), np.float64) ** 2, axis=2)
    ref = np.load(os.path.join(HERE, "golden", "cfg2_n300_summary.npz"))
    assert abs(got["misfit64"] - ref["misfit64"]) <= 1e-9 * abs(ref["misfit64"])
    for k in ref.files:
        e = _rel(got[k], ref[k])
        print(f"{k}: rel-L2 {e:.3e}")
        assert e <= TOL, k
    assert np.array_equal(got["traces"], ref["traces"])
