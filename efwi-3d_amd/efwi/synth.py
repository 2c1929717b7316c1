"""Synthetic workloads: the layered-plus-anomaly models, sources and receiver
spreads of the benchmark configurations, generated from fixed seeds.

  "cfg1"  64x64x64 physical grid, dx 10 m, dt 1 ms, 500 steps, 10 PML cells;
          4 flat layers (vp 2000/2500/3000/3500 m/s, interfaces drawn from
          seed 0), vs = vp/1.73, rho = 310 vp^0.25 kg/m^3, a +10 % Gaussian vp
          anomaly (sigma 5 cells) at the centre; Ricker f0 10 Hz at
          (z 5, x 32, y 32); 64x64 receivers at z 3. Observed data come from
          the anomaly-free model.
  "cfg2"  120x200x200 physical grid (z, x, y), dx 12.5 m, dt 1 ms, 3000 steps,
          20 PML cells; 6 layers with vp in 1800-4000 m/s (seed 1),
          vs = max(vp/1.73, 1100), Gardner density, 3 Gaussian anomalies of
          +-8 % (sigma 8 cells); Ricker f0 8 Hz at (z 5, x 100, y 100);
          100x100 receivers every 2 cells at z 3.
All coordinates returned by `workload` are on the padded grid.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np

from .propagator import GridSpec, ShotSpec


def ricker_integrated(f, nt, dt):
    """fwi_utils.sourceGene restated in numpy (Ops/FWI/fwi_utils.py:90-105)."""
    e = np.pi * np.pi * f * f
    td = 1.2 / f
    s = np.zeros(nt)
    for it in range(nt):
        t = dt * it - td
        s[it] = (1 - 2 * e * t ** 2) * np.exp(-e * t ** 2)
    for it in range(1, nt):
        s[it] = s[it] + s[it - 1]
    return s * dt


def layered(nz, nx, ny, vps, interfaces):
    vp = np.empty((nz, nx, ny), np.float64)
    edges = [0] + [int(i) for i in interfaces] + [nz]
    for k, v in enumerate(vps):
        vp[edges[k]:edges[k + 1]] = v
    return vp


def gaussian(nz, nx, ny, c, sigma):
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    r2 = (z - c[0]) ** 2 + (x - c[1]) ** 2 + (y - c[2]) ** 2
    return np.exp(-r2 / (2.0 * sigma ** 2))


def pad_replicate(a, p):
    return np.pad(a, p, mode="edge")


def lame(vp, vs, rho_kgm3):
    """(lambda, mu, rho[g/cc]) as FWI_3d forms them (Ops/FWI/FWI_ops_3d.py:91-93),
    in float32 arithmetic."""
    vp = vp.astype(np.float32); vs = vs.astype(np.float32)
    rho = (rho_kgm3.astype(np.float32) / np.float32(1e3)).astype(np.float32)
    lam = ((vp ** 2 - np.float32(2.0) * vs ** 2) * rho).astype(np.float32)
    mu = (vs ** 2 * rho).astype(np.float32)
    return lam, mu, rho


@dataclass
class Workload:
    name: str
    grid: GridSpec
    shots: list           # list[ShotSpec]
    true: tuple           # (lam, mu, rho) padded float32, the model that makes the observed data
    model: tuple          # (lam, mu, rho) padded float32, the model whose misfit/gradient is computed
    vp_model: np.ndarray  # padded vp of `model` (for stability checks)
    velocities_true: tuple = None   # padded float32 (vp, vs, rho[kg/m^3]) of `true`
    velocities_model: tuple = None  # padded float32 (vp, vs, rho[kg/m^3]) of `model`

    def courant(self):
        g = self.grid
        vmax = float(self.vp_model.max())
        return vmax * g.dt * np.sqrt(3.0) * (9.0 / 8.0 + 1.0 / 24.0) / min(g.dz, g.dx, g.dy)


def _finish(name, nzp, nxp, nyp, p, h, dt, nt, f0, vp_t, vs_t, rho_t, vp_m, vs_m, rho_m, src, rec_zxy,
            formulation=0):
    grid = GridSpec(nzp + 2 * p, nxp + 2 * p, nyp + 2 * p, p, nt, h, h, h, dt, f0, formulation)
    pads = lambda a: pad_replicate(a, p)
    true = lame(pads(vp_t), pads(vs_t), pads(rho_t))
    model = lame(pads(vp_m), pads(vs_m), pads(rho_m))
    stf = ricker_integrated(f0, nt, dt).astype(np.float32)
    shots = [ShotSpec((s[0] + p, s[1] + p, s[2] + p), stf, np.asarray(rec_zxy, np.int32) + p) for s in src]
    f32 = lambda a: pads(a).astype(np.float32)
    return Workload(name, grid, shots, true, model, pads(vp_m), (f32(vp_t), f32(vs_t), f32(rho_t)),
                    (f32(vp_m), f32(vs_m), f32(rho_m)))


def cfg1(nt=500, n=64, npml=10):
    """Plain method, one shot (64^3 physical grid at full size)."""
    rng = np.random.default_rng(0)
    nz = nx = ny = n
    interfaces = np.sort(rng.integers(n // 8, n - n // 8, size=3))
    vp_b = layered(nz, nx, ny, [2000.0, 2500.0, 3000.0, 3500.0], interfaces)
    vs_b = vp_b / 1.73
    rho_b = 310.0 * vp_b ** 0.25
    c = n // 2
    vp_a = vp_b * (1.0 + 0.10 * gaussian(nz, nx, ny, (c, c, c), 5.0 * n / 64))
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rec = np.stack([np.full(gx.size, 3), gx.ravel(), gy.ravel()], axis=1)
    return _finish("cfg1", nz, nx, ny, npml, 10.0, 1e-3, nt, 10.0, vp_b, vs_b, rho_b, vp_a, vs_b, rho_b,
                   [(5, c, c)], rec)


def cfg2(nt=3000, scale=1.0):
    """Single-shot production stencil: 120x200x200 physical grid, 20 PML cells."""
    rng = np.random.default_rng(1)
    nz, nx, ny = int(round(120 * scale)), int(round(200 * scale)), int(round(200 * scale))
    npml = max(4, int(round(20 * scale)))
    vps = np.sort(rng.uniform(1800.0, 4000.0, size=6))
    interfaces = np.sort(rng.choice(np.arange(8, nz - 8), size=5, replace=False))
    vp_b = layered(nz, nx, ny, vps, interfaces)
    vs_b = np.maximum(vp_b / 1.73, 1100.0)
    rho_b = 310.0 * vp_b ** 0.25
    pert = np.zeros_like(vp_b)
    for k in range(3):
        cz = rng.integers(nz // 4, 3 * nz // 4)
        cx = rng.integers(nx // 4, 3 * nx // 4)
        cy = rng.integers(ny // 4, 3 * ny // 4)
        sign = 1.0 if rng.uniform() < 0.5 else -1.0
        pert += sign * 0.08 * gaussian(nz, nx, ny, (cz, cx, cy), 8.0 * scale)
    vp_a = vp_b * (1.0 + pert)
    xs = np.arange(0, nx, 2)
    ys = np.arange(0, ny, 2)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    rec = np.stack([np.full(gx.size, 3), gx.ravel(), gy.ravel()], axis=1)
    return _finish("cfg2", nz, nx, ny, npml, 12.5, 1e-3, nt, 8.0, vp_b, vs_b, rho_b, vp_a, vs_b, rho_b,
                   [(5, nx // 2, ny // 2)], rec)


def small(n=32, npml=6, nt=120, nrec_side=8, seed=3, formulation=0, h=10.0, dt=1e-3, f0=15.0, ny=None):
    """A small heterogeneous case for parity tests (seeded): n^3 cells, or
    n x n x ny with a longer y axis."""
    rng = np.random.default_rng(seed)
    nz = nx = n
    ny = n if ny is None else ny
    vps = np.sort(rng.uniform(2000.0, 3200.0, size=3))
    interfaces = np.sort(rng.choice(np.arange(4, nz - 4), size=2, replace=False))
    vp_b = layered(nz, nx, ny, vps, interfaces)
    vp_b = vp_b * (1.0 + 0.02 * rng.standard_normal(vp_b.shape))
    vs_b = vp_b / 1.73
    rho_b = 310.0 * vp_b ** 0.25
    vp_a = vp_b * (1.0 + 0.05 * gaussian(nz, nx, ny, (nz // 2, nx // 2, ny // 2), n / 10.0))
    xs = np.linspace(1, nx - 2, nrec_side).astype(int)
    ys = np.linspace(1, ny - 2, nrec_side).astype(int)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    rec = np.stack([np.full(gx.size, 2), gx.ravel(), gy.ravel()], axis=1)
    return _finish("small", nz, nx, ny, npml, h, dt, nt, f0, vp_b, vs_b, rho_b, vp_a, vs_b, rho_b,
                   [(3, nx // 2, ny // 2)], rec, formulation)


def cfg3(nt=500, nshots=16, nx=160, ny=160, nz=64, npml=10, seed=3):
    """Batched shots on the plain-method recipe over a wider surface: 64x160x160
    physical grid (10 PML cells), dx 10 m, dt 1 ms, 500 steps; 16 seeded shots
    at z 5, each with a square receiver aperture at z 3 centred on its source,
    side drawn from [30, 140] cells at 1-cell spacing (clipped to the grid), so
    900 to 19,600 receivers per shot."""
    rng = np.random.default_rng(seed)
    interfaces = np.sort(rng.integers(nz // 8, nz - nz // 8, size=3))
    vp_b = layered(nz, nx, ny, [2000.0, 2500.0, 3000.0, 3500.0], interfaces)
    vs_b = vp_b / 1.73
    rho_b = 310.0 * vp_b ** 0.25
    vp_a = vp_b * (1.0 + 0.10 * gaussian(nz, nx, ny, (nz // 2, nx // 2, ny // 2), 5.0))
    w = _finish("cfg3", nz, nx, ny, npml, 10.0, 1e-3, nt, 10.0, vp_b, vs_b, rho_b, vp_a, vs_b, rho_b,
                [(5, nx // 2, ny // 2)], np.zeros((1, 3), np.int32) + 3)
    shots = []
    for k in range(nshots):
        sx, sy = (int(v) for v in rng.integers(8, [nx - 8, ny - 8]))
        side = int(rng.integers(30, 141))
        xs = np.arange(sx - side // 2, sx - side // 2 + side)
        ys = np.arange(sy - side // 2, sy - side // 2 + side)
        xs = xs[(xs >= 0) & (xs < nx)]
        ys = ys[(ys >= 0) & (ys < ny)]
        gx, gy = np.meshgrid(xs, ys, indexing="ij")
        rec = np.stack([np.full(gx.size, 3), gx.ravel(), gy.ravel()], axis=1).astype(np.int32) + npml
        shots.append(ShotSpec((5 + npml, sx + npml, sy + npml), w.shots[0].stf, rec))
    w.shots = shots
    return w


def cfg4(nt=4000, nshots=4, scale=1.0, seed=4):
    """Differentiation with checkpointing: 160x320x320 physical grid (20 PML
    cells), dx 10 m, dt 0.8 ms, 4000 steps; the plain-method recipe (4 seeded
    layers, +10 % Gaussian vp anomaly at the centre with sigma 5 cells per 64,
    vs = vp/1.73, rho = 310 vp^0.25) with vp up to 4200 m/s; 4 seeded shots
    (Ricker, f0 10 Hz) at z 5, receivers every 2 cells at z 3. The boundary
    ring of all steps would need ~300 GB, so the gradient runs with sqrt(nt)
    checkpoints (memory_mode 1)."""
    rng = np.random.default_rng(seed)
    nz, nx, ny = int(round(160 * scale)), int(round(320 * scale)), int(round(320 * scale))
    npml = max(4, int(round(20 * scale)))
    interfaces = np.sort(rng.integers(nz // 8, nz - nz // 8, size=3))
    vp_b = layered(nz, nx, ny, [2000.0, 2700.0, 3400.0, 4200.0], interfaces)
    vs_b = vp_b / 1.73
    rho_b = 310.0 * vp_b ** 0.25
    vp_a = vp_b * (1.0 + 0.10 * gaussian(nz, nx, ny, (nz // 2, nx // 2, ny // 2), 5.0 * nx / 64))
    xs = np.arange(0, nx, 2)
    ys = np.arange(0, ny, 2)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    rec = np.stack([np.full(gx.size, 3), gx.ravel(), gy.ravel()], axis=1)
    src = [(5, int(a), int(b)) for a, b in zip(rng.integers(nx // 4, 3 * nx // 4, nshots),
                                               rng.integers(ny // 4, 3 * ny // 4, nshots))]
    w = _finish("cfg4", nz, nx, ny, npml, 10.0, 8e-4, nt, 10.0, vp_b, vs_b, rho_b, vp_a, vs_b, rho_b, src, rec)
    w.grid = dataclasses.replace(w.grid, memory_mode=1)
    return w


WORKLOADS = {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3, "cfg4": cfg4}
