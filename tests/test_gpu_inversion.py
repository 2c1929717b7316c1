"""Device-resident inversion: the autograd misfit against the propagator,
the scaled vp gradient against a centred finite difference, and a few
L-BFGS iterations that must lower the misfit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    from efwi import synth
    from efwi.inversion import Problem
    from efwi.propagator import Propagator
    w = synth.small(n=32, npml=6, nt=160)
    print(f"[gpu] inversion grid={w.grid.shape} nt={w.grid.nt} shots={len(w.shots)}", flush=True)
    P = Propagator(w.grid)
    P.set_model(*(torch.from_numpy(a) for a in w.true))
    obs = [P.run_shot(s, 2, None, want_syn=True)[1].clone() for s in w.shots]
    dev = "cuda"
    vp_t = torch.from_numpy(w.velocities_true[0]).to(dev)
    vp0, vs, rho = (torch.from_numpy(a).to(dev) for a in w.velocities_model)
    p = w.grid.npml
    mask = torch.zeros(w.grid.shape, device=dev)
    mask[p:-p, p:-p, p:-p] = 1.0
    # mute the source and receiver depths, where the gradient has its
    # near-field spike (standard practice; the misfit stays the plain one)
    mask[:w.shots[0].src[0] + 4] = 0.0
    prob = Problem(P, w.shots, obs, vs, rho, mask)
    yield prob, vp0, vp_t
    P.close()


def test_autograd_matches_propagator_gradient(problem):
    from efwi.inversion import shot_misfit, vp_to_lame
    prob, vp0, _ = problem
    lam, mu, rho = (t.detach().requires_grad_(True) for t in vp_to_lame(vp0, prob.vs, prob.rho_kg))
    f = shot_misfit(lam, mu, rho, prob.prop, prob.shots, prob.obs)
    (3.0 * f).backward()
    prob.prop.set_model(lam.detach(), mu.detach(), rho.detach())
    prob.prop.zero_gradient()
    j = sum(prob.prop.run_shot(s, 1, o)[0] for s, o in zip(prob.shots, prob.obs))
    g = prob.prop.gradient()
    assert float(f) == j
    for a, b in zip((lam.grad, mu.grad, rho.grad), g):
        assert torch.equal(a, 3.0 * b)


def test_scaled_vp_gradient_matches_finite_difference(problem):
    """Centred difference of the misfit along a smooth random vp direction
    against <dJ/dvp, d>."""
    prob, vp0, _ = problem
    f0, g = prob.evaluate(vp0)
    rng = np.random.default_rng(0)
    d = torch.from_numpy(rng.standard_normal(vp0.shape).astype(np.float32)).cuda() * prob.mask
    d = torch.nn.functional.avg_pool3d(d[None, None], 5, 1, 2)[0, 0] * prob.mask
    h = 1e-3 * float(vp0.abs().mean()) / float(d.abs().max())
    fp, _ = prob.evaluate(vp0 + h * d)
    fm, _ = prob.evaluate(vp0 - h * d)
    fd = (fp - fm) / (2 * h)
    an = float(torch.dot(g.double().ravel(), d.double().ravel()))
    print(f"fd {fd:.6e} analytic {an:.6e} ratio {an / fd:.5f}")
    # the reference formulation's backward is not the exact adjoint of its
    # forward (its C-PML and dsxy/dx terms differ), so the bound is loose; the
    # point is the factor 2 that makes (f, g) consistent: the raw kernel
    # gradient gives ratios near 0.5
    assert 0.7 < an / fd < 1.3


def test_lbfgs_lowers_the_misfit(problem):
    from efwi.inversion import lbfgs_invert
    prob, vp0, vp_t = problem
    vp, hist = lbfgs_invert(prob, vp0, niter=4)
    print("misfit", hist.f, "nfev", hist.nfev)
    assert len(hist.f) == 5
    assert all(b < a for a, b in zip(hist.f, hist.f[1:]))
    assert hist.f[-1] < 0.8 * hist.f[0]
