"""Device-resident FWI driver: misfit and gradient of all shots as a torch
autograd function, and an L-BFGS inversion loop over them.

The reference inverts with scipy's L-BFGS-B on the host
(3D_inversion_BFGS.py:152-190, obj_wrapper.py): every evaluation copies the
model to the host, writes files, runs cufd and copies the gradient back.
Here the model, gradient and L-BFGS history stay in HBM; only the misfit
(a float64 scalar) crosses to the host.

Gradient scale: the reference's gradient kernels accumulate half the
derivative of the misfit they report (the misfit is the plain sum of
squared residuals; the imaging condition carries no factor 2). ShotMisfit
returns the reference's gradient unchanged, and the inversion multiplies it
by `grad_scale` = 2 so that line searches see a consistent (f, g) pair.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .lbfgs import LBFGS, armijo, default_dot
from .propagator import Propagator


class ShotMisfit(torch.autograd.Function):
    """misfit(lambda, mu, rho) summed over shots, with the hand-written
    adjoint as backward. forward runs the whole gradient pipeline (calc 1)
    and caches the gradient; backward returns it scaled by grad_output."""

    @staticmethod
    def forward(ctx, lam, mu, rho, prop: Propagator, shots, obs):
        prop.set_model(lam, mu, rho)
        prop.zero_gradient()
        total = 0.0
        for s, o in zip(shots, obs):
            j, _, _ = prop.run_shot(s, 1, o)
            total += j
        ctx.grads = prop.gradient()
        return torch.tensor(total, dtype=torch.float64, device=lam.device)

    @staticmethod
    def backward(ctx, g):
        s = g.to(torch.float32)
        gl, gm, gr = ctx.grads
        return gl * s, gm * s, gr * s, None, None, None


def shot_misfit(lam, mu, rho, prop, shots, obs):
    return ShotMisfit.apply(lam, mu, rho, prop, shots, obs)


def vp_to_lame(vp, vs, rho_kg):
    """(lambda, mu, rho[g/cc]) from velocities and density in kg/m^3, as
    FWI_3d.forward (Ops/FWI/FWI_ops_3d.py:84-90)."""
    rho = rho_kg / 1e3
    return (vp ** 2 - 2.0 * vs ** 2) * rho, vs ** 2 * rho, rho


def stable_vp_max(grid) -> float:
    """The largest P velocity the grid's time step keeps stable: Courant number
    vp dt sqrt(3) (9/8 + 1/24) / min(dz, dx, dy) at most 1 (efwi_set_model
    refuses models above it; the reference's compCourantNumber,
    Src/utilities.cu:186-202, exits there)."""
    h = min(grid.dz, grid.dx, grid.dy)
    return float(h / (float(np.float32(grid.dt)) * np.sqrt(3.0) * (9.0 / 8.0 + 1.0 / 24.0)))


@dataclass
class Problem:
    """Invert vp on the padded grid with vs and rho fixed, as the reference's
    driver does (3D_inversion_BFGS.py:138-152); `mask` is 1 where vp is free.
    Trial models are projected onto [vmin, vmax], with vmax lowered to just
    below the grid's stability limit (the bounds of the reference's L-BFGS-B,
    3D_inversion_BFGS.py:188)."""
    prop: Propagator
    shots: list
    obs: list
    vs: torch.Tensor
    rho_kg: torch.Tensor
    mask: torch.Tensor
    vmin: float = 1000.0
    vmax: float = 6000.0
    grad_scale: float = 2.0
    nfev: int = 0

    def __post_init__(self):
        self.vmax = min(self.vmax, 0.999 * stable_vp_max(self.prop.grid))

    def evaluate(self, vp: torch.Tensor):
        """(misfit, dJ/dvp masked) at vp."""
        x = vp.detach().requires_grad_(True)
        lam, mu, rho = vp_to_lame(x, self.vs, self.rho_kg)
        f = shot_misfit(lam, mu, rho, self.prop, self.shots, self.obs)
        f.backward()
        self.nfev += 1
        return float(f.detach()), (x.grad * self.mask * self.grad_scale).detach()

    def project(self, vp):
        return vp.clamp(self.vmin, self.vmax)


@dataclass
class History:
    f: list = field(default_factory=list)
    t: list = field(default_factory=list)
    nfev: list = field(default_factory=list)
    kind: list = field(default_factory=list)


def lbfgs_invert(problem: Problem, vp0: torch.Tensor, niter: int, first_step: float = 20.0,
                 history: int = 30, dot=default_dot, step_hook=None, log=None):
    """Projected L-BFGS with Armijo backtracking. The first direction is the
    steepest descent scaled to a largest update of `first_step` m/s; later
    iterations try the unit step. `step_hook(it, vp, g, d)` may return an
    alternative trial model (a learned step, say); it is kept only when it
    lowers the misfit, otherwise the L-BFGS step is taken (a hook with a
    `record(accepted)` method is told which).

    Returns (vp, History) with the misfit after each iteration and the
    wall time since the start (synchronised)."""
    sync = torch.cuda.synchronize if vp0.is_cuda else (lambda: None)
    sync()
    t0 = time.perf_counter()
    H = LBFGS(history=history, dot=dot)
    vp = problem.project(vp0.clone())
    f, g = problem.evaluate(vp)
    hist = History()
    hist.f.append(f); hist.t.append(0.0); hist.nfev.append(problem.nfev); hist.kind.append("init")
    for it in range(niter):
        d = H.direction(g)
        if step_hook is not None:
            alt = step_hook(it, vp, g, d)
            if alt is not None:
                alt = problem.project(alt)
                fa, ga = problem.evaluate(alt)
                if hasattr(step_hook, "record"):
                    step_hook.record(fa < f)
                if fa < f:
                    H.update(alt - vp, ga - g)
                    vp, f, g = alt, fa, ga
                    sync()
                    hist.f.append(f); hist.t.append(time.perf_counter() - t0)
                    hist.nfev.append(problem.nfev); hist.kind.append("learned")
                    continue
        slope = dot(g, d)
        if not slope < 0:
            H.reset()
            d = -g
            slope = dot(g, d)
        if not H.s:
            dmax = float(d.abs().max())
            step = first_step / dmax if dmax > 0 else 0.0
        else:
            step = 1.0

        def trial(a):
            x = problem.project(vp + a * d)
            fx, gx = problem.evaluate(x)
            return fx, (x, gx)

        a, fn, (vn, gn), _ = armijo(f, slope, trial, step)
        if not fn < f:
            break
        H.update(vn - vp, gn - g)
        vp, f, g = vn, fn, gn
        sync()
        hist.f.append(f); hist.t.append(time.perf_counter() - t0)
        hist.nfev.append(problem.nfev); hist.kind.append("lbfgs")
        if log:
            log(f"[lbfgs] iteration {it + 1}: misfit {f:.6e}, {problem.nfev} evaluations, {hist.t[-1]:.1f} s")
    return vp, hist
