"""Device-resident elastic propagator: a thin Python face of the efwi_ctx C API.

Tensors stay on the GPU: the model, observed data, seismograms and gradients
are torch CUDA tensors in the torch layout [nz][nx][ny] (padded grid) and
[3][nrec][nt] (x, y, z) for seismograms. All work is queued on torch's
current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as N


@dataclass(frozen=True)
class GridSpec:
    """Padded grid, time stepping and options (the para JSON of the reference)."""
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
    formulation: int = 0     # 0 reference, 1 consistent
    if_win: bool = False
    memory_mode: int = -1    # -1 auto, 0 boundary ring of all steps, 1 ring rebuilt per checkpointed segment
    checkpoint_every: int = 0  # memory_mode 1 segment length (0 = ceil(sqrt(nt - 1)))

    @property
    def shape(self):
        return (self.nz, self.nx, self.ny)

    @property
    def ncells(self):
        return self.nz * self.nx * self.ny

    def interior_cells(self):
        p = self.npml
        return (self.nz - 2 * p) * (self.nx - 2 * p) * (self.ny - 2 * p)


@dataclass
class ShotSpec:
    """One shot on the padded grid (the survey JSON value plus npml)."""
    src: tuple                     # (z, x, y)
    stf: np.ndarray                # [nt] untapered source time function
    rec: np.ndarray                # [nrec, 3] (z, x, y)
    win_start: np.ndarray | None = None
    win_end: np.ndarray | None = None
    weights: np.ndarray | None = None
    src_weight: float = 1.0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nrec(self):
        return int(len(self.rec))

    def c_struct(self) -> N.Shot:
        stf = np.ascontiguousarray(self.stf, dtype=np.float32)
        rec = np.ascontiguousarray(np.asarray(self.rec, dtype=np.int32).reshape(-1, 3))
        keep = [stf, rec]

        def fp(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float))

        s = N.Shot(int(self.src[0]), int(self.src[1]), int(self.src[2]),
                   stf.ctypes.data_as(C.POINTER(C.c_float)), rec.shape[0],
                   rec.ctypes.data_as(C.POINTER(C.c_int)), fp(self.win_start), fp(self.win_end),
                   fp(self.weights), float(self.src_weight))
        self._keep = keep
        return s


class Propagator:
    """Owns one efwi_ctx (device state of one grid on one GPU)."""

    def __init__(self, grid: GridSpec, device: int | None = None):
        if not torch.cuda.is_available():
            raise N.EfwiError("no HIP device visible: the propagator runs on the GPU only")
        self.grid = grid
        self.device = torch.cuda.current_device() if device is None else int(device)
        cfg = N.Config(grid.nz, grid.nx, grid.ny, grid.npml, grid.nt, grid.dz, grid.dx, grid.dy,
                       grid.dt, grid.f0, int(grid.formulation), int(bool(grid.if_win)), self.device,
                       int(grid.memory_mode), int(grid.checkpoint_every))
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_create(C.byref(cfg), C.byref(h)), "efwi_create")
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            torch.cuda.synchronize(self.device)
            N.lib().efwi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self) -> int:
        return int(N.lib().efwi_device_bytes(self._h))

    def profile(self, enable: bool = True):
        """Time every stencil/ring launch with HIP events on the launch stream."""
        N.check(N.lib().efwi_profile(self._h, int(enable)), "efwi_profile")

    def profile_read(self, reset: bool = True) -> dict:
        """{kernel: (total_ms, launches)} since the last reset."""
        ms = (C.c_double * len(N.KERNELS))()
        n = (C.c_longlong * len(N.KERNELS))()
        N.check(N.lib().efwi_profile_read(self._h, ms, n, int(reset)), "efwi_profile_read")
        return {k: (ms[i], int(n[i])) for i, k in enumerate(N.KERNELS)}

    def set_tile(self, kernel: str, rows: int, zdepth: int):
        """Launch shape of one time-loop kernel (results do not depend on it)."""
        N.check(N.lib().efwi_set_tile(self._h, N.KERNELS.index(kernel), int(rows), int(zdepth)), "efwi_set_tile")

    def set_fusion(self, rows: int, zchunk: int = 16, pass_: int = 0):
        """Fused forward step (stress + velocity in one launch; rows 4, 8 or
        12, 0 = two kernels). Results do not depend on it."""
        N.check(N.lib().efwi_set_fusion(self._h, int(pass_), int(rows), int(zchunk)), "efwi_set_fusion")

    def set_split(self, on: bool):
        """Interior/rim split of the backward kernels (results do not depend on it)."""
        N.check(N.lib().efwi_set_split(self._h, int(bool(on))), "efwi_set_split")

    def split(self) -> bool:
        v = C.c_int()
        N.check(N.lib().efwi_get_split(self._h, C.byref(v)), "efwi_get_split")
        return bool(v.value)

    def set_resident_segments(self, segments: int = 0):
        """memory_mode 1: the time loop's last `segments` segments keep the
        forward pass's boundary ring instead of being replayed from
        checkpoints (0: as many as the free device memory holds; results do
        not depend on it)."""
        N.check(N.lib().efwi_set_resident_segments(self._h, int(segments)), "efwi_set_resident_segments")

    def resident_segments(self) -> tuple:
        """(resident segments, segments of the time loop)."""
        r, n = C.c_int(), C.c_int()
        N.check(N.lib().efwi_get_resident_segments(self._h, C.byref(r), C.byref(n)), "efwi_get_resident_segments")
        return r.value, n.value

    def tiles(self) -> dict:
        out = {}
        for k in N.TUNED_KERNELS:
            r, z = C.c_int(), C.c_int()
            N.check(N.lib().efwi_get_tile(self._h, N.KERNELS.index(k), C.byref(r), C.byref(z)), "efwi_get_tile")
            out[k] = (r.value, z.value)
        return out

    def autotune(self, reps: int = 5) -> dict:
        """Time every candidate launch shape of the time-loop kernels on this
        grid and keep the fastest; returns {kernel: {(rows, zdepth): ms}}."""
        nc = len(N.TILE_CANDIDATES)
        rep = (C.c_double * (4 * nc))()
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_autotune(self._h, int(reps), rep, self._stream()), "efwi_autotune")
        kinds = ("stress_fwd", "velocity_fwd", "velocity_bwd", "stress_bwd")
        return {k: {N.TILE_CANDIDATES[c]: rep[q * nc + c] for c in range(nc)} for q, k in enumerate(kinds)}

    def _stream(self):
        return N.stream_handle(self.device)

    def _dev(self, t, shape=None):
        t = torch.as_tensor(t)
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
        return t.to(device=f"cuda:{self.device}", dtype=torch.float32).contiguous()

    # -- model / gradient -------------------------------------------------
    def set_model(self, lam, mu, rho):
        """lambda, mu and rho (g/cc) on the padded grid [nz][nx][ny]."""
        g = self.grid
        l, m, r = (self._dev(a, g.shape) for a in (lam, mu, rho))
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_set_model(self._h, N.vptr(l), N.vptr(m), N.vptr(r), 0, self._stream()),
                    "efwi_set_model")
            # the copies are queued on the stream; keep the sources alive until then
            torch.cuda.current_stream(self.device).synchronize()

    def zero_gradient(self):
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_zero_gradient(self._h, self._stream()), "efwi_zero_gradient")

    def gradient(self):
        """(g_lambda, g_mu, g_rho) accumulated since the last zero_gradient()."""
        g = self.grid
        out = [torch.empty(g.shape, dtype=torch.float32, device=f"cuda:{self.device}") for _ in range(3)]
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_get_gradient(self._h, *(N.vptr(t) for t in out), 0, self._stream()),
                    "efwi_get_gradient")
        return tuple(out)

    def set_illumination(self, on: bool = True):
        """Accumulate the source illumination sum_t |v|^2 of every shot's forward
        field from now on (cleared by zero_gradient)."""
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_set_illumination(self._h, int(bool(on)), self._stream()), "efwi_set_illumination")

    def illumination(self):
        """The accumulated source illumination (padded grid, float32 CUDA tensor)."""
        g = self.grid
        out = torch.empty(g.shape, dtype=torch.float32, device=f"cuda:{self.device}")
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_get_illumination(self._h, N.vptr(out), 0, self._stream()), "efwi_get_illumination")
        return out

    # -- shots ------------------------------------------------------------
    def run_shot(self, shot: ShotSpec, calc_id: int, obs=None, want_syn=False, want_res=False):
        """Forward (+ misfit, + backward for calc_id 1) of one shot.

        Returns (misfit, syn, res): misfit is the shot's FP64 sum of squared
        windowed residuals (None for calc_id 2); syn/res are [3][nrec][nt]
        CUDA tensors when requested.
        """
        g = self.grid
        nrec = shot.nrec
        dev = f"cuda:{self.device}"
        shp = (3, nrec, g.nt)
        obs_t = None
        if calc_id != N.CALC_SYNTHETIC:
            if obs is None:
                raise ValueError("observed data required for calc_id 0/1")
            obs_t = self._dev(obs, shp)
        syn = torch.empty(shp, dtype=torch.float32, device=dev) if (want_syn or calc_id == N.CALC_SYNTHETIC) else None
        res = torch.empty(shp, dtype=torch.float32, device=dev) if want_res else None
        j = C.c_double(0.0)
        cs = shot.c_struct()
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_run_shot(self._h, C.byref(cs), int(calc_id), N.vptr(obs_t), N.vptr(syn),
                                          N.vptr(res), C.byref(j), self._stream()), "efwi_run_shot")
        return (None if calc_id == N.CALC_SYNTHETIC else j.value), syn, res

    def run_batch(self, shots, calc_id: int, obs=None, want_syn=False, use_graph=True):
        """All `shots` together: one persistent launch per stencil step over
        every (shot, tile) item, each time step replayed from a captured
        hipGraph (use_graph). Returns (misfits, syns): a list of FP64 misfits
        (None for calc_id 2) and, when requested, [3][nrec][nt] CUDA tensors.
        calc_id 1 keeps one gradient per shot (batch_gradient)."""
        g = self.grid
        n = len(shots)
        dev = f"cuda:{self.device}"
        keep = []
        obs_p = (C.c_void_p * n)()
        syn_p = (C.c_void_p * n)()
        syns = []
        for k, sh in enumerate(shots):
            shp = (3, sh.nrec, g.nt)
            if calc_id != N.CALC_SYNTHETIC:
                if obs is None:
                    raise ValueError("observed data required for calc_id 0/1")
                o = self._dev(obs[k], shp)
                keep.append(o)
                obs_p[k] = N.vptr(o)
            if want_syn or calc_id == N.CALC_SYNTHETIC:
                y = torch.empty(shp, dtype=torch.float32, device=dev)
                syns.append(y)
                syn_p[k] = N.vptr(y)
            else:
                syns.append(None)
        cs = (N.Shot * n)()
        for k, sh in enumerate(shots):
            cs[k] = sh.c_struct()
            keep.append(sh._keep)
        mis = (C.c_double * n)()
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_run_batch(self._h, cs, n, int(calc_id), obs_p, syn_p, mis, int(bool(use_graph)),
                                           self._stream()), "efwi_run_batch")
        misfits = None if calc_id == N.CALC_SYNTHETIC else [mis[k] for k in range(n)]
        return misfits, (syns if (want_syn or calc_id == N.CALC_SYNTHETIC) else None)

    def batch_group(self, shots_per_launch: int = 0):
        """How run_batch runs: groups of this many shots per persistent launch,
        -1 for one shot after another on the single-shot kernels, 0 for the
        automatic choice (single-shot kernels where one shot fills the
        device). Results do not depend on it."""
        N.check(N.lib().efwi_batch_group(self._h, int(shots_per_launch)), "efwi_batch_group")

    def batch_config(self, take: int = 2, static_pct: int = 0):
        """Work distribution of the batched launches: each persistent workgroup
        runs static_pct percent of an even share of its XCD's queue, then
        claims the rest `take` tiles per atomic."""
        N.check(N.lib().efwi_batch_config(self._h, int(take), int(static_pct)), "efwi_batch_config")

    def batch_gradient(self, shot: int):
        """(g_lambda, g_mu, g_rho) of one shot of the last run_batch."""
        g = self.grid
        out = [torch.empty(g.shape, dtype=torch.float32, device=f"cuda:{self.device}") for _ in range(3)]
        with torch.cuda.device(self.device):
            N.check(N.lib().efwi_batch_gradient(self._h, int(shot), *(N.vptr(t) for t in out), 0, self._stream()),
                    "efwi_batch_gradient")
        return tuple(out)
