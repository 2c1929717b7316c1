"""Learned preconditioner for the vp inversion: training data from the plain
method, training on the GPU's own kernels, and the learned step inside the
L-BFGS driver (the plain step is the fallback whenever the learned one does
not lower the misfit; the plain misfit judges both).

Training pairs (the plain method on random layered-plus-anomaly models):
input = (raw vp gradient g at a smoothed starting model, log source
illumination H = sum_t |v|^2 of the same shots); target = the pseudo-Hessian-
scaled gradient (Shin et al. 2001; p = 1/2 and eps = 1e-2 measured best on
small inversions) in the normalisation of the inputs,
    (g / max|g|) (1 + eps) max H^p / (H^p + eps max H^p),
an exact function of the two input channels (the overall scale of a
preconditioner is immaterial: L-BFGS rescales it); the update is its negative.
The network outputs that update as g exp(o); it is trained on the log of the
pairs' per-voxel ratio update / gradient (log-space mean-square error on o),
which weighs every voxel alike (an error on the update itself is weighted by
g^2 and all but ignores the weakly illuminated voxels). The
network works on the physical region (the C-PML cells are cropped); in the
inversion it is the initial inverse Hessian of L-BFGS, so the step length
and the curvature corrections come from the driver.

No reference counterpart: EFWI-3D has no learned part.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import synth
from .inversion import Problem
from .propagator import GridSpec, Propagator, ShotSpec
from .unet import UNet3D, pack_input

ILLUM_EPS = 1e-2
ILLUM_POWER = 0.5


def random_model(rng, nz, nx, ny, nlayers=(4, 6), vp_range=(1800.0, 4000.0), nanom=3, amp=0.08, sigma=6.0):
    """Physical-size vp, vs, rho[kg/m^3] of a seeded layered model with
    Gaussian vp anomalies (the recipe of the benchmark models), and the
    smoothed starting vp."""
    from scipy.ndimage import gaussian_filter
    k = int(rng.integers(nlayers[0], nlayers[1] + 1))
    vps = np.sort(rng.uniform(*vp_range, size=k))
    interfaces = np.sort(rng.choice(np.arange(4, nz - 4), size=k - 1, replace=False))
    vp_b = synth.layered(nz, nx, ny, list(vps), interfaces)
    pert = np.zeros_like(vp_b)
    for _ in range(nanom):
        c = (rng.integers(nz // 4, 3 * nz // 4), rng.integers(nx // 4, 3 * nx // 4), rng.integers(ny // 4, 3 * ny // 4))
        pert += (1.0 if rng.uniform() < 0.5 else -1.0) * amp * synth.gaussian(nz, nx, ny, c, sigma)
    vp = vp_b * (1.0 + pert)
    vs = np.maximum(vp / 1.73, 1100.0)
    rho = 310.0 * vp ** 0.25
    vp0 = gaussian_filter(vp, 4.0)
    return vp, vs, rho, vp0


@dataclass
class Survey:
    grid: GridSpec
    shots: list


def surface_survey(nz, nx, ny, npml, nt, h, dt, f0, nshots, rng, rec_step=2):
    """Shots at z 5 at seeded surface positions, receivers every rec_step cells at z 3."""
    g = GridSpec(nz + 2 * npml, nx + 2 * npml, ny + 2 * npml, npml, nt, h, h, h, dt, f0)
    stf = synth.ricker_integrated(f0, nt, dt).astype(np.float32)
    xs, ys = np.arange(0, nx, rec_step), np.arange(0, ny, rec_step)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    rec = (np.stack([np.full(gx.size, 3), gx.ravel(), gy.ravel()], 1) + npml).astype(np.int32)
    shots = []
    for _ in range(nshots):
        sx, sy = int(rng.integers(nx // 5, 4 * nx // 5)), int(rng.integers(ny // 5, 4 * ny // 5))
        shots.append(ShotSpec((5 + npml, sx + npml, sy + npml), stf, rec))
    return Survey(g, shots)


def _padded(a, p, dev):
    return torch.from_numpy(synth.pad_replicate(a, p).astype(np.float32)).to(dev)


def make_problem(prop, survey, vp, vs, rho, dev):
    """Observed data of the true model (vp, vs, rho physical) and the vp
    Problem (vs, rho fixed; the C-PML cells and the source/receiver depths
    masked, as the reference driver's mask)."""
    p = survey.grid.npml
    vpt, vst, rht = (_padded(a, p, dev) for a in (vp, vs, rho))
    from .inversion import vp_to_lame
    prop.set_model(*vp_to_lame(vpt, vst, rht))
    obs = [prop.run_shot(s, 2)[1] for s in survey.shots]
    mask = torch.zeros(survey.grid.shape, device=dev)
    mask[p + 8:-p, p:-p, p:-p] = 1.0
    return Problem(prop, survey.shots, obs, vst, rht, mask)


def pair(problem, vp0, dev):
    """(network input, target) at the starting model vp0 (padded tensor)."""
    p = problem.prop.grid.npml
    problem.prop.set_illumination(True)
    f, g = problem.evaluate(vp0)
    illum = problem.prop.illumination()
    feat, dims = pack_input(g, illum, p)
    h = illum.clamp_min(0).pow(ILLUM_POWER)
    hmax = h.max()
    scale = (1.0 + ILLUM_EPS) * hmax / (h + ILLUM_EPS * hmax)
    u = (g / g.abs().max().clamp_min(1e-30)) * scale
    u = u[p:-p, p:-p, p:-p].reshape(-1).contiguous()
    logs = torch.log(scale[p:-p, p:-p, p:-p]).reshape(-1).contiguous()
    return feat, u, logs, dims, f


def generate(nmodels, dims=(96, 128, 128), npml=20, nt=600, h=10.0, dt=1e-3, f0=10.0, nshots=1, seed=0,
             device=0, log=None):
    """Training pairs of `nmodels` seeded random models (physical dims z, x, y)."""
    dev = torch.device(f"cuda:{device}")
    rng = np.random.default_rng(seed)
    nz, nx, ny = dims
    data = []
    t0 = time.perf_counter()
    survey = surface_survey(nz, nx, ny, npml, nt, h, dt, f0, nshots, rng)
    prop = Propagator(survey.grid, device)
    try:
        for k in range(nmodels):
            survey = Survey(survey.grid, surface_survey(nz, nx, ny, npml, nt, h, dt, f0, nshots, rng).shots)
            vp, vs, rho, vp0 = random_model(rng, nz, nx, ny)
            prob = make_problem(prop, survey, vp, vs, rho, dev)
            feat, u, logs, fdims, _ = pair(prob, _padded(vp0, npml, dev), dev)
            data.append((feat, u, logs))
            if log and (k + 1) % 16 == 0:
                log(f"[learned] {k + 1}/{nmodels} training pairs, {time.perf_counter() - t0:.1f} s")
    finally:
        prop.close()
    return data, fdims, time.perf_counter() - t0


def train(data, dims, epochs=6, lr=1e-3, seed=0, device=0, log=None, warmup=0.05):
    """Adam on the pairs (log-space loss on the scaling), one volume per step:
    linear learning-rate warm-up over the first `warmup` fraction of the
    steps, then cosine decay."""
    net = UNet3D(device, seed=seed)
    V = dims[0] * dims[1] * dims[2]
    mask = torch.ones(V, device=torch.device(f"cuda:{device}"))
    rng = np.random.default_rng(seed)
    steps = epochs * len(data)
    losses = []
    t0 = time.perf_counter()
    k = 0
    for ep in range(epochs):
        for i in rng.permutation(len(data)):
            w = max(1, int(warmup * steps))
            a = lr * (k + 1) / w if k < w else lr * 0.5 * (1.0 + math.cos(math.pi * (k - w) / max(1, steps - w)))
            feat, _, logs = data[i]
            losses.append(net.train_step(feat, logs, mask, V, dims, lr=a, log_scaling=True).clone())  # buffer reused
            k += 1
        ep_loss = float(torch.cat(losses[-len(data):]).mean())
        if log:
            log(f"[learned] epoch {ep + 1}/{epochs}: loss {ep_loss:.4e}, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    return net, [float(x) for x in torch.cat(losses).cpu()], time.perf_counter() - t0


class Preconditioner:
    """H0 q = s * q with the scaling field s = exp(clamp(o, -4, 6)) that the
    network outputs for the current gradient and illumination (one forward
    per iteration, `refresh`); s lives on the physical region and H0 q is zero
    in the C-PML and outside the inversion mask. Linear and positive
    definite, as an initial inverse Hessian of L-BFGS must be."""

    def __init__(self, net, problem):
        self.net, self.problem = net, problem
        self.s = None

    def refresh(self, g, illum):
        p = self.problem.prop.grid.npml
        feat, dims = pack_input(g, illum, p)
        self.net.forward(feat, dims)
        o = self.net.scaling_log(dims)
        s = torch.zeros_like(g)
        s[p:-p, p:-p, p:-p] = torch.exp(o.clamp(-4.0, 6.0))
        self.s = s.mul_(self.problem.mask)

    def __call__(self, q):
        return q * self.s


def learned_invert(problem, vp0, niter, net, first_step=20.0, history=30, dot=None, log=None):
    """L-BFGS whose initial inverse Hessian is the learned preconditioner:
    every iteration first tries the preconditioned direction (Armijo line
    search, unit step once curvature pairs exist, the first step scaled to a
    largest update of `first_step` m/s); when that does not lower the misfit,
    the plain L-BFGS step of the same history is taken. Both share the
    curvature pairs of the accepted steps. Returns (vp, History, counts)."""
    from .inversion import History
    from .lbfgs import LBFGS, armijo, default_dot
    dot = dot or default_dot
    sync = torch.cuda.synchronize
    problem.prop.set_illumination(True)
    P = Preconditioner(net, problem)
    sync()
    t0 = time.perf_counter()
    H = LBFGS(history=history, dot=dot)
    vp = problem.project(vp0.clone())
    f, g = problem.evaluate(vp)
    P.refresh(g, problem.prop.illumination())
    hist = History()
    hist.f.append(f); hist.t.append(0.0); hist.nfev.append(problem.nfev); hist.kind.append("init")
    counts = {"learned": 0, "lbfgs": 0}
    for it in range(niter):
        took = None
        for kind in ("learned", "lbfgs"):
            d = H.direction(g, precond=P if kind == "learned" else None)
            slope = dot(g, d)
            if not slope < 0:
                continue
            if not H.s:
                dmax = float(d.abs().max())
                step = first_step / dmax if dmax > 0 else 0.0
            else:
                step = 1.0

            def trial(a):
                x = problem.project(vp + a * d)
                fx, gx = problem.evaluate(x)
                return fx, (x, gx, problem.prop.illumination())

            _, fn, (vn, gn, il), _ = armijo(f, slope, trial, step, max_tries=6 if kind == "learned" else 12)
            if fn < f:
                took = (kind, vn, fn, gn, il)
                break
        if took is None:
            break
        kind, vn, fn, gn, il = took
        H.update(vn - vp, gn - g)
        vp, f, g = vn, fn, gn
        P.refresh(g, il)
        counts[kind] += 1
        sync()
        hist.f.append(f); hist.t.append(time.perf_counter() - t0)
        hist.nfev.append(problem.nfev); hist.kind.append(kind)
        if log:
            log(f"[learned] iteration {it + 1} ({kind}): misfit {f:.6e}, {problem.nfev} evaluations, {hist.t[-1]:.1f} s")
    return vp, hist, counts
