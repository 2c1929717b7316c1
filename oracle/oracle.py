"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is the CPU restatement of the reference's elastic propagation path
(guoketing/EFWI-3D, Ops/FWI/Src_3d) in ``efwi_oracle.cpp``. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker; the product never calls it.

Arrays use the torch layout ``[nz][nx][ny]`` of the padded grid. Seismograms
are per shot ``[3][nrec][nt]`` with component order (x, y, z), the order of
the reference's ``Shot_x/Shot_y/Shot_z`` files (Src/libCUFD.cu:704-725).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")

# Field ids of efwi_oracle.cpp (enum Field).
FIELDS = [
    "vz", "vx", "vy", "szz", "sxx", "sxz", "syy", "sxy", "syz",
    "avz", "avx", "avy", "aszz", "asxx", "asxz", "asyy", "asxy", "asyz",
    "m_dvz_dz", "m_dvz_dx", "m_dvz_dy", "m_dvx_dz", "m_dvx_dx", "m_dvx_dy",
    "m_dvy_dz", "m_dvy_dx", "m_dvy_dy",
    "m_dszz_dz", "m_dsxz_dx", "m_dsxz_dz", "m_dsxx_dx", "m_dsxy_dx", "m_dsxy_dy",
    "m_dsyy_dy", "m_dsyz_dy", "m_dsyz_dz",
    "lam", "mu", "byc", "glam", "gmu", "grho",
]
FID = {n: i for i, n in enumerate(FIELDS)}
KERNELS = {"stress_fwd": 0, "velocity_fwd": 1, "stress_rev": 2, "velocity_rev": 3,
           "velocity_adj": 4, "stress_adj": 5}
PRECISION = {"fp32": 0, "fp32_literal": 1, "fp64": 2}
FORMULATION = {"reference": 0, "consistent": 1}


class _Grid(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("npml", C.c_int),
                ("nt", C.c_int), ("dz", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
                ("dt", C.c_float), ("f0", C.c_float)]


class _Survey(C.Structure):
    _fields_ = [("nshots", C.c_int), ("src", C.POINTER(C.c_int)), ("stf", C.POINTER(C.c_float)),
                ("rec_off", C.POINTER(C.c_long)), ("rec", C.POINTER(C.c_int)),
                ("if_win", C.c_int), ("win_start", C.POINTER(C.c_float)),
                ("win_end", C.POINTER(C.c_float)), ("weights", C.POINTER(C.c_float)),
                ("src_weight", C.POINTER(C.c_float))]


def build(force: bool = False) -> str:
    """Compile liborc.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Grid), C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set.argtypes = [vp, C.c_int, dp]
        L.orc_get.argtypes = [vp, C.c_int, dp]
        L.orc_kernel.argtypes = [vp, C.c_int]
        L.orc_cpml_off.argtypes = [vp]
        L.orc_get_cpml.argtypes = [vp, C.c_int, dp]
        L.orc_add_source.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]
        L.orc_band_mask.argtypes = [C.POINTER(_Grid), C.POINTER(C.c_ubyte)]
        L.orc_taper_plain.argtypes = [C.c_int, C.c_float, C.c_float, fp]
        L.orc_fp32_block_sumsq.restype = C.c_float
        L.orc_fp32_block_sumsq.argtypes = [fp, C.c_long]
        L.orc_cpml_profile.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, fp]
        L.orc_set_checkpoint.argtypes = [C.c_int]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(_Grid), C.POINTER(_Survey), C.c_int, C.c_int, C.c_int,
                              C.c_int, fp, fp, fp, fp, dp, dp, dp, fp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class Grid:
    nz: int
    nx: int
    ny: int
    npml: int
    nt: int
    dz: float
    dx: float
    dy: float
    dt: float
    f0: float

    def c(self) -> _Grid:
        return _Grid(self.nz, self.nx, self.ny, self.npml, self.nt, self.dz, self.dx, self.dy,
                     self.dt, self.f0)

    @property
    def shape(self):
        return (self.nz, self.nx, self.ny)


@dataclass
class Shots:
    """Per-shot sources and receivers on the padded grid (z, x, y)."""
    src: np.ndarray            # [nshots][3] int
    stf: np.ndarray            # [nshots][nt] float32 rows (untapered)
    rec: list                  # list of [nrec_s][3] int arrays
    win_start: list = field(default_factory=list)
    win_end: list = field(default_factory=list)
    weights: list = field(default_factory=list)
    src_weight: np.ndarray | None = None

    @property
    def nshots(self):
        return len(self.rec)

    def rec_off(self):
        off = [0]
        for r in self.rec:
            off.append(off[-1] + len(r))
        return np.asarray(off, dtype=np.int64)


def cpml_profile(n, npml, dh, f0, dt):
    """[6][n] float32 rows K, a, b, K_half, a_half, b_half (Src/utilities.cu:204-320)."""
    out = np.zeros((6, n), dtype=np.float32)
    lib().orc_cpml_profile(n, npml, dh, f0, dt, _ptr(out, C.c_float))
    return out


def taper_w2(nt, dt, ratio):
    out = np.zeros(nt, dtype=np.float32)
    lib().orc_taper_plain(nt, dt, ratio, _ptr(out, C.c_float))
    return out


def fp32_block_sumsq(e):
    e = _f32(e).ravel()
    return float(lib().orc_fp32_block_sumsq(_ptr(e, C.c_float), e.size))


def band_mask(g: Grid) -> np.ndarray:
    out = np.zeros(g.shape, dtype=np.uint8)
    lib().orc_band_mask(C.byref(g.c()), _ptr(out, C.c_ubyte))
    return out.astype(bool)


class Stepper:
    """Direct access to the oracle's kernels for unit tests."""

    def __init__(self, g: Grid, precision="fp64", formulation="reference", cpml_off=False):
        self.g = g
        self._c = g.c()
        self.h = lib().orc_create(C.byref(self._c), PRECISION[precision], FORMULATION[formulation])
        if cpml_off:
            lib().orc_cpml_off(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set(self, name, arr):
        a = np.ascontiguousarray(np.broadcast_to(arr, self.g.shape), dtype=np.float64)
        lib().orc_set(self.h, FID[name], _ptr(a, C.c_double))

    def get(self, name):
        a = np.zeros(self.g.shape, dtype=np.float64)
        lib().orc_get(self.h, FID[name], _ptr(a, C.c_double))
        return a

    def kernel(self, name):
        lib().orc_kernel(self.h, KERNELS[name])

    def add_source(self, z, x, y, amp, sign=1):
        lib().orc_add_source(self.h, z, x, y, amp, sign)

    def cpml(self, axis):
        n = self.g.shape[axis]
        a = np.zeros((6, n), dtype=np.float64)
        lib().orc_get_cpml(self.h, axis, _ptr(a, C.c_double))
        return a


@dataclass
class Result:
    syn: list           # per shot [3][nrec][nt] float64
    res: list           # per shot [3][nrec][nt] float64 (adjoint source)
    misfit64: float
    misfit32: float
    glam: np.ndarray | None
    gmu: np.ndarray | None
    grho: np.ndarray | None
    seconds: float


def run(g: Grid, shots: Shots, lam, mu, rho, obs=None, calc_id=1, precision="fp32",
        formulation="reference", cpml_off=False, nthreads=None, checkpoint_every=0) -> Result:
    """Run the cufd pipeline (Src/libCUFD.cu:37-799) for all shots in order.

    ``obs`` is a list of per-shot [3][nrec][nt] arrays (required for calc_id 0/1).
    ``rho`` is in the units the op receives (g/cc for FWIFunction).
    ``checkpoint_every`` > 0 keeps the boundary ring of that many steps only,
    rebuilt per segment from checkpoints of the forward state (same results
    bit for bit; for grids whose full ring exceeds host memory).
    """
    if nthreads is not None:
        os.environ["OMP_NUM_THREADS"] = str(nthreads)
    lam, mu, rho = _f32(lam), _f32(mu), _f32(rho)
    assert lam.shape == g.shape and mu.shape == g.shape and rho.shape == g.shape
    off = shots.rec_off()
    ntot = int(off[-1])
    rec = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.int32).reshape(-1, 3)
                                               for r in shots.rec]), dtype=np.int32)
    src = np.ascontiguousarray(shots.src, dtype=np.int32).reshape(-1, 3)
    stf = _f32(shots.stf).reshape(shots.nshots, g.nt)
    # bounds of every receiver / source (the reference would read out of bounds)
    for a in (rec, src):
        assert (a[:, 0] >= 0).all() and (a[:, 0] < g.nz).all()
        assert (a[:, 1] >= 0).all() and (a[:, 1] < g.nx).all()
        assert (a[:, 2] >= 0).all() and (a[:, 2] < g.ny).all()
    if_win = 1 if shots.win_start else 0
    keep = [rec, src, stf, off]
    if if_win:
        ws = _f32(np.concatenate(shots.win_start)); we = _f32(np.concatenate(shots.win_end))
        wt = _f32(np.concatenate(shots.weights)) if shots.weights else np.ones(ntot, np.float32)
        sw = _f32(shots.src_weight if shots.src_weight is not None else np.ones(shots.nshots))
        keep += [ws, we, wt, sw]
        wptr = [_ptr(ws, C.c_float), _ptr(we, C.c_float), _ptr(wt, C.c_float), _ptr(sw, C.c_float)]
    else:
        wptr = [None, None, None, None]
    sv = _Survey(shots.nshots, _ptr(src, C.c_int), _ptr(stf, C.c_float), _ptr(off, C.c_long),
                 _ptr(rec, C.c_int), if_win, *wptr)
    if calc_id != 2:
        assert obs is not None and len(obs) == shots.nshots
        obs_flat = _f32(np.concatenate([np.asarray(o, dtype=np.float32).ravel() for o in obs]))
        assert obs_flat.size == 3 * ntot * g.nt
    else:
        obs_flat = np.zeros(1, np.float32)
    syn = np.zeros(3 * ntot * g.nt, dtype=np.float64)
    res = np.zeros(3 * ntot * g.nt, dtype=np.float64)
    m64 = C.c_double(0.0)
    m32 = C.c_float(0.0)
    secs = C.c_double(0.0)
    N = g.nz * g.nx * g.ny
    if calc_id == 1:
        gl = np.zeros(N, np.float64); gm = np.zeros(N, np.float64); gr = np.zeros(N, np.float64)
        gp = [_ptr(gl, C.c_double), _ptr(gm, C.c_double), _ptr(gr, C.c_double)]
    else:
        gl = gm = gr = None
        gp = [None, None, None]
    gc = g.c()
    lib().orc_set_checkpoint(int(checkpoint_every))
    rc = lib().orc_run(C.byref(gc), C.byref(sv), PRECISION[precision], FORMULATION[formulation],
                       calc_id, int(cpml_off), _ptr(lam, C.c_float), _ptr(mu, C.c_float),
                       _ptr(rho, C.c_float), _ptr(obs_flat, C.c_float), _ptr(syn, C.c_double),
                       _ptr(res, C.c_double), C.byref(m64), C.byref(m32), *gp, C.byref(secs))
    if rc != 0:
        raise RuntimeError(f"orc_run failed with {rc}")
    syn_l, res_l = [], []
    for s in range(shots.nshots):
        n = int(off[s + 1] - off[s])
        a, b = 3 * int(off[s]) * g.nt, 3 * int(off[s + 1]) * g.nt
        syn_l.append(syn[a:b].reshape(3, n, g.nt))
        res_l.append(res[a:b].reshape(3, n, g.nt))
    shp = g.shape
    return Result(syn_l, res_l, m64.value, m32.value,
                  None if gl is None else gl.reshape(shp), None if gm is None else gm.reshape(shp),
                  None if gr is None else gr.reshape(shp), secs.value)
