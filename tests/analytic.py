"""Closed-form wavefield of an explosive point source in a homogeneous elastic
medium, for pinning the source and receiver terms of the oracle and of the
GPU path.

The reference adds 1500^2 * s[it] * dt to szz, sxx and syy of the source cell
at step it (Src/utilities.cu:442-479), and records v after step it in slot
it + 1 (Src/utilities.cu:481-495, Src/libCUFD.cu:351). In the continuous
equations this is a stress-rate source dsigma_ii/dt = S(t) delta(x) with
S(t_n) = 1500^2 s[n], t_n = n dt (leapfrog: the stress update of step n spans
[t_n - dt/2, t_n + dt/2] and its source term is the midpoint rule), and a cell
of volume h^3 turns delta(x) into 1/h^3. With the stress glut convention
sigma = C:eps - M delta, that is an explosive moment tensor M_ij = M0 delta_ij
with dM0/dt = -S h^3. Its exact solution (Aki & Richards, eq. 4.32 reduced to
an isotropic source) has only a radial velocity

    v_r(r, t) = -(1500^2 h^3 / (4 pi rho a^2)) (s(tau) / r^2 + s'(tau) / (a r)),

tau = t - r / a, a = sqrt((lambda + 2 mu) / rho) (rho in the op's units,
g/cc, with lambda and mu in the same units: the speeds do not depend on them,
the amplitude scales with 1 / rho). On the staggered grid vx sits half a cell
along +x from the cell's stress point, vy half a cell along +y, vz half a cell
along +z (the stencils of Src/stress.cu:40-68 difference v[i] - v[i-1] into the
stress at i). s(tau) between samples is the band-limited interpolation of the
samples.
"""
import numpy as np

AMP = 1500.0 ** 2
# half-cell offsets (z, x, y) of the recorded components, in the order x, y, z
# of the seismogram files
COMPONENTS = (("x", (0.0, 0.5, 0.0), 1), ("y", (0.0, 0.0, 0.5), 2), ("z", (0.5, 0.0, 0.0), 0))


class Explosion:
    """Analytic traces for a source time function `stf` sampled at t_n = n dt."""

    def __init__(self, stf, dt, rho, vp, h):
        stf = np.asarray(stf, np.float64)
        self.N = 8 * len(stf)
        self.F = np.fft.rfft(stf, self.N)
        self.f = np.fft.rfftfreq(self.N, dt)
        self.w = np.full(self.f.size, 2.0)
        self.w[0] = 1.0
        self.w[-1] = 1.0
        self.dt, self.rho, self.vp, self.h = dt, rho, vp, h

    def _s(self, tau):
        E = np.exp(2j * np.pi * np.outer(tau, self.f))
        s = np.real(E @ (self.F * self.w)) / self.N
        ds = np.real(E @ (2j * np.pi * self.f * self.F * self.w)) / self.N
        return s, ds

    def trace(self, nt, rel_cells, axis, shift=0.0):
        """Velocity component `axis` (0 z, 1 x, 2 y) at offset `rel_cells`
        (z, x, y, in cells) from the source, samples t_k = k dt (+ shift dt)."""
        rel = np.asarray(rel_cells, np.float64) * self.h
        r = np.linalg.norm(rel)
        tau = (np.arange(nt) - shift) * self.dt - r / self.vp
        s, ds = self._s(tau)
        vr = -(AMP * self.h ** 3 / (4 * np.pi * self.rho * self.vp ** 2)) * (s / r ** 2 + ds / (self.vp * r))
        return vr * rel[axis] / r

    def fit(self, tr, nt, rel_cells, axis):
        """Best time shift (samples, to 0.025 within +-1.5) and amplitude
        factor of `tr` against the analytic trace, the relative L2 misfit at
        that fit, and the misfit of the unshifted, unscaled analytic trace."""
        tr = np.asarray(tr, np.float64)
        if not np.isfinite(tr).all():
            return {"shift": np.nan, "amp": np.nan, "err": np.inf, "err0": np.inf}
        best = None

        def probe(s):
            nonlocal best
            an = self.trace(nt, rel_cells, axis, s)
            a = float(np.dot(tr, an) / np.dot(an, an))
            na = np.linalg.norm(a * an)
            e = float(np.linalg.norm(tr - a * an) / na) if na > 0 else np.inf
            if best is None or e < best[0]:
                best = (e, float(s), a)

        for s in np.linspace(-1.5, 1.5, 31):
            probe(s)
        c = best[1]
        for s in np.linspace(c - 0.1, c + 0.1, 9):
            probe(s)
        an0 = self.trace(nt, rel_cells, axis, 0.0)
        e0 = float(np.linalg.norm(tr - an0) / np.linalg.norm(an0))
        return {"shift": best[1], "amp": best[2], "err": best[0], "err0": e0}


def components(offsets):
    """(receiver index, component index, offset in cells of that component,
    axis) of every component whose direction cosine to the source is at least
    0.3 (the others carry a small fraction of the radial velocity)."""
    out = []
    for k, off in enumerate(offsets):
        for ci, (_, d, axis) in enumerate(COMPONENTS):
            rel = np.asarray(off, np.float64) + np.asarray(d)
            if abs(rel[axis]) / np.linalg.norm(rel) >= 0.3:
                out.append((k, ci, rel, axis))
    return out
