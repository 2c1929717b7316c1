"""Shared helpers of the test suite (oracle adapters and error measures)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def oracle_grid(orc, g):
    return orc.Grid(g.nz, g.nx, g.ny, g.npml, g.nt, g.dz, g.dx, g.dy, g.dt, g.f0)


def oracle_shots(orc, shots):
    return orc.Shots(src=np.array([s.src for s in shots]),
                     stf=np.stack([np.asarray(s.stf, np.float32) for s in shots]),
                     rec=[np.asarray(s.rec) for s in shots],
                     win_start=[s.win_start for s in shots if s.win_start is not None],
                     win_end=[s.win_end for s in shots if s.win_end is not None],
                     weights=[s.weights if s.weights is not None else np.ones(len(s.rec), np.float32)
                              for s in shots if s.win_start is not None],
                     src_weight=np.array([s.src_weight for s in shots], np.float32))


def oracle_run(orc, w, model=None, obs=None, calc_id=1, precision="fp32", formulation=None,
               cpml_off=False):
    g = w.grid
    form = ["reference", "consistent"][g.formulation if formulation is None else formulation]
    m = w.model if model is None else model
    return orc.run(oracle_grid(orc, g), oracle_shots(orc, w.shots), *m, obs=obs, calc_id=calc_id,
                   precision=precision, formulation=form, cpml_off=cpml_off)
