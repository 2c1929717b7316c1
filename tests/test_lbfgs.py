"""L-BFGS two-loop recursion and the projected inversion loop on analytic
problems (CPU, float64 torch dot products)."""
import numpy as np
import pytest
import torch

from efwi.lbfgs import LBFGS, armijo


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return torch.tensor(q @ np.diag(np.linspace(1.0, 20.0, n)) @ q.T)


def test_secant_equation_holds_for_latest_pair():
    """-direction(y_k) = H y_k must equal s_k (the BFGS update enforces the
    secant equation for the newest pair)."""
    A = _spd(12, 0)
    H = LBFGS(history=5)
    rng = np.random.default_rng(1)
    for _ in range(4):
        s = torch.tensor(rng.standard_normal(12))
        H.update(s, A @ s)
    s = torch.tensor(rng.standard_normal(12))
    y = A @ s
    H.update(s, y)
    assert torch.allclose(-H.direction(y), s, rtol=1e-10, atol=1e-10)


def test_full_history_recovers_inverse_on_quadratic():
    """With n mutually A-conjugate pairs of a quadratic (the steps of exact
    line searches) the two-loop H equals A^-1 (hereditary property)."""
    n = 8
    A = _spd(n, 3)
    H = LBFGS(history=n)
    _, vecs = torch.linalg.eigh(A)
    for k in range(n):
        s = vecs[:, k] * (k + 1.0)
        H.update(s, A @ s)
    g = torch.tensor(np.random.default_rng(4).standard_normal(n))
    assert torch.allclose(-H.direction(g), torch.linalg.solve(A, g), rtol=1e-8, atol=1e-10)


def test_empty_history_is_steepest_descent():
    g = torch.tensor([1.0, -2.0, 3.0], dtype=torch.float64)
    assert torch.equal(LBFGS().direction(g), -g)


def test_curvature_check_skips_nonconvex_pairs():
    H = LBFGS()
    assert not H.update(torch.tensor([1.0, 0.0]), torch.tensor([-1.0, 0.0]))
    assert H.s == []


def test_armijo_backtracks_to_sufficient_decrease():
    f = lambda a: ((a - 0.1) ** 2, None)
    step, fx, _, tries = armijo(0.01, -0.2, f, 1.0)
    assert fx <= 0.01 + 1e-4 * step * -0.2 and tries > 1


class _Quadratic:
    def __init__(self, n=30, seed=5):
        self.A = _spd(n, seed)
        self.b = torch.tensor(np.random.default_rng(seed).standard_normal(n))
        self.nfev = 0

    def evaluate(self, x):
        self.nfev += 1
        r = self.A @ x - self.b
        return float(0.5 * x @ self.A @ x - self.b @ x), r

    def project(self, x):
        return x


def test_lbfgs_invert_converges_on_quadratic():
    from efwi.inversion import lbfgs_invert
    p = _Quadratic()
    x, hist = lbfgs_invert(p, torch.zeros(30, dtype=torch.float64), niter=60, first_step=0.5)
    xs = torch.linalg.solve(p.A, p.b)
    assert all(b < a for a, b in zip(hist.f, hist.f[1:]))
    assert torch.linalg.norm(x - xs) <= 1e-6 * torch.linalg.norm(xs)


def test_learned_step_is_kept_only_when_it_lowers_the_misfit():
    from efwi.inversion import lbfgs_invert
    p = _Quadratic()
    xs = torch.linalg.solve(p.A, p.b)
    _, good = lbfgs_invert(p, torch.zeros(30, dtype=torch.float64), niter=1,
                           step_hook=lambda it, x, g, d: xs.clone())
    assert good.kind == ["init", "learned"]
    _, bad = lbfgs_invert(p, torch.zeros(30, dtype=torch.float64), niter=2,
                          step_hook=lambda it, x, g, d: x + 1e3)
    assert bad.kind == ["init", "lbfgs", "lbfgs"]
