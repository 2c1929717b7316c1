"""L-BFGS with FP64 dot products, for the inversion driver.

The reference's driver hands the misfit and its gradient to scipy's L-BFGS-B
(`optimize.minimize(method='L-BFGS-B', maxcor=30, ...)`,
3D_inversion_BFGS.py:187-190) on the host in float64. Here the model, the
gradient and the history stay on the GPU as float32 vectors, and every inner
product of the two-loop recursion is an FP64 reduction (libefwi's k_dot64),
so the direction does not depend on float32 summation order.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable

import torch


def default_dot(a: torch.Tensor, b: torch.Tensor) -> float:
    """FP64 dot product: the GPU kernel for CUDA float32 tensors, torch otherwise."""
    if a.is_cuda and a.dtype == torch.float32:
        from ._native import dot64
        return dot64(a.reshape(-1), b.reshape(-1))
    return float(torch.dot(a.reshape(-1).double(), b.reshape(-1).double()))


@dataclass
class LBFGS:
    """Two-loop recursion (Nocedal & Wright, Algorithm 7.4) with the initial
    scaling gamma = s.y / y.y; curvature pairs with s.y <= eps |s||y| are
    skipped."""
    history: int = 30
    dot: Callable = default_dot
    eps: float = 1e-10
    s: list = field(default_factory=list)
    y: list = field(default_factory=list)
    rho: list = field(default_factory=list)

    def reset(self):
        self.s.clear(); self.y.clear(); self.rho.clear()

    def update(self, s: torch.Tensor, y: torch.Tensor) -> bool:
        sy = self.dot(s, y)
        nrm = (self.dot(s, s) * self.dot(y, y)) ** 0.5
        if not sy > self.eps * nrm:
            return False
        self.s.append(s.clone()); self.y.append(y.clone()); self.rho.append(1.0 / sy)
        if len(self.s) > self.history:
            self.s.pop(0); self.y.pop(0); self.rho.pop(0)
        return True

    def direction(self, g: torch.Tensor, precond=None) -> torch.Tensor:
        """-H g for the current inverse-Hessian approximation H. With
        `precond` (a map q -> H0 q, positive definite in effect), the initial
        approximation is gamma H0 with gamma = s.y / y.H0 y (Nocedal & Wright,
        eq. 7.20 generalised); without it H0 = I."""
        q = g.clone()
        alpha = []
        for s, y, rho in zip(reversed(self.s), reversed(self.y), reversed(self.rho)):
            a = rho * self.dot(s, q)
            q.add_(y, alpha=-a)
            alpha.append(a)
        if precond is not None:
            q = precond(q)
            if self.s:
                yh = self.dot(self.y[-1], precond(self.y[-1]))
                if yh > 0:
                    q.mul_(self.dot(self.s[-1], self.y[-1]) / yh)
        elif self.s:
            gamma = self.dot(self.s[-1], self.y[-1]) / self.dot(self.y[-1], self.y[-1])
            q.mul_(gamma)
        for (s, y, rho), a in zip(zip(self.s, self.y, self.rho), reversed(alpha)):
            b = rho * self.dot(y, q)
            q.add_(s, alpha=a - b)
        return q.neg_()


def armijo(f0: float, slope: float, evaluate: Callable, step: float, c1: float = 1e-4, shrink: float = 0.5,
           max_tries: int = 12):
    """Backtracking line search on f(step) <= f0 + c1 step slope.

    `evaluate(step)` returns (f, payload). Returns (step, f, payload, tries) of
    the first accepted step, or of the best tried step when none is accepted
    (step is then negative in sign convention: the caller checks f < f0)."""
    best = None
    for k in range(max_tries):
        f, payload = evaluate(step)
        if best is None or f < best[1]:
            best = (step, f, payload, k + 1)
        if f <= f0 + c1 * step * slope:
            return step, f, payload, k + 1
        step *= shrink
    return best
