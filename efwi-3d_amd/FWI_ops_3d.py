"""Drop-in replacement of the reference's Ops/FWI/FWI_ops_3d.py on gfx950.

A user of the reference switches by putting this directory on sys.path in
place of Ops/FWI; the module names, classes, functions and argument orders
are the reference's:

  fwi_ops.forward(Lambda, Mu, Rho, Stf, gpu_id, Shot_ids, para_fname) -> [misfit]
  fwi_ops.backward(Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname)
        -> [misfit, grad_Lambda, grad_Mu, grad_Rho, grad_Stf]
  fwi_ops.obscalc(Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname) -> None
  FWIFunction.apply(Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname) -> misfit
  FWI_3d(Vp, Vs, Rho, Stf, opt, Mask=None).forward(Shot_ids, ngpu=1)
  FWI_obscalc(Vp, Vs, Rho, Stf, opt).forward(Shot_ids, ngpu=1)

(Src/Torch_Fwi.cpp:12-131, Ops/FWI/FWI_ops_3d.py:34-108.) The native work
runs in libefwi.so; there is no CPU path. Differences from the reference,
all deliberate:
  * errors raise EfwiError / ContractError instead of exiting the process;
  * FWIFunction.backward scales the gradients by grad_output (the reference
    ignores it, which only agrees when the loss is the misfit itself);
  * CUDA tensors are accepted and stay on the device (zero copy); CPU tensors
    take the reference's host-pointer path through efwi_cufd;
  * the misfit is reduced in FP64 (the reference sums in FP32) and returned
    as float32;
  * FWI_3d's default Mask is 3-D (the reference's 2-D default cannot
    broadcast against the 3-D model).
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import threading

import numpy as np
import torch
import torch.nn as nn

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

import fwi_utils as ft  # noqa: E402
from efwi import _native as N  # noqa: E402
from efwi import io as eio  # noqa: E402
from efwi.propagator import Propagator  # noqa: E402


def _shot_bars(group_size: int, ngpu: int):
    """Contiguous shot chunks per GPU (Src/Torch_Fwi.cpp:53-66): the bars of
    torch.linspace(0, group_size, ngpu+1), truncated to int."""
    bars = torch.linspace(0, group_size, ngpu + 1, dtype=torch.float32)
    return [(int(bars[i].item()), int(bars[i + 1].item())) for i in range(ngpu)]


def _check_ngpu(ngpu: int, group_size: int):
    if ngpu < 1:
        raise N.EfwiError("ngpu must be at least 1")
    if ngpu > group_size:
        raise N.EfwiError("The number of GPUs should be smaller than the number of shots!")
    if ngpu > torch.cuda.device_count():
        raise N.EfwiError(f"ngpu={ngpu} but only {torch.cuda.device_count()} HIP device(s) are visible")


def _f32c(t):
    return t.detach().to(torch.float32).contiguous()


class _DeviceCache:
    """One Propagator per (grid, device), reused across calls (the reference
    re-allocates everything on every call)."""

    def __init__(self):
        self._lock = threading.Lock()
        self._props = {}

    def get(self, grid, device):
        with self._lock:
            key = (grid, device)
            p = self._props.get(key)
            if p is None:
                for k in [k for k in self._props if k[1] == device]:
                    self._props.pop(k).close()
                p = Propagator(grid, device)
                self._props[key] = p
            return p


_cache = _DeviceCache()


def _window_factors(grid, shot):
    """The misfit's factors w(t)^2 (x weights) per trace, [nrec][nt] or [nt]
    (Src/utilities.cu:524-614), as the library applies them."""
    def fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))
    if grid.if_win:
        out = np.zeros((shot.nrec, grid.nt), np.float32)
        ws, we = (np.ascontiguousarray(a, np.float32) for a in (shot.win_start, shot.win_end))
        wt = None if shot.weights is None else np.ascontiguousarray(shot.weights, np.float32)
        N.check(N.lib().efwi_window_factors(grid.nt, grid.dt, shot.nrec, fp(ws), fp(we), None if wt is None else fp(wt),
                                            float(shot.src_weight), fp(out)), "efwi_window_factors")
        return out
    out = np.zeros(grid.nt, np.float32)
    N.check(N.lib().efwi_taper_w2(grid.nt, grid.dt, 0.005, fp(out)), "efwi_taper_w2")
    return out


def _run_device(Lambda, Mu, Rho, Stf, calc_id, device, shot_ids, para_fname):
    """The cufd pipeline for device tensors through the context API, with the
    host path's files: synthetic data for calc_id 2, and for gradients with a
    scratch_dir_name the windowed residual, synthetic and observed data
    (Src/libCUFD.cu:660-725)."""
    para = eio.read_para(para_fname)
    grid = eio.grid_from_para(para)
    survey = eio.read_survey(para["survey_fname"])
    stf = Stf.detach().to("cpu", torch.float32).numpy()
    prop = _cache.get(grid, device)
    scratch = para.get("scratch_dir_name") if calc_id == N.CALC_GRADIENT else None
    with torch.cuda.device(device):
        prop.set_model(Lambda, Mu, Rho)
        prop.zero_gradient()
        total = 0.0
        for sid in shot_ids:
            shot = eio.shot_from_survey(survey, sid, grid.npml, stf[sid], grid.if_win)
            obs = None
            if calc_id != N.CALC_SYNTHETIC:
                obs = torch.from_numpy(eio.read_shot_data(para["data_dir_name"], sid, shot.nrec, grid.nt))
            j, syn, res = prop.run_shot(shot, calc_id, obs, want_syn=scratch is not None, want_res=scratch is not None)
            if calc_id == N.CALC_SYNTHETIC:
                eio.write_shot_data(para["data_dir_name"], sid, syn.cpu().numpy())
            else:
                total += j
            if scratch is not None:
                fac = _window_factors(grid, shot)
                eio.write_shot_data(scratch, sid, res.cpu().numpy(), prefix="Residual_Shot_")
                eio.write_shot_data(scratch, sid, syn.cpu().numpy() * fac, prefix="Syn_Shot_")
                eio.write_shot_data(scratch, sid, obs.numpy() * fac, prefix="CondObs_Shot_")
        grads = prop.gradient() if calc_id == N.CALC_GRADIENT else None
        torch.cuda.synchronize(device)
    return total, grads


def _run_host(Lambda, Mu, Rho, Stf, calc_id, gpu_id, shot_ids, para_fname, want_grad):
    """The reference's host-pointer path: one efwi_cufd call (Src/libCUFD.cu:37)."""
    lam, mu, rho, stf = (_f32c(t).cpu() for t in (Lambda, Mu, Rho, Stf))
    ids = np.ascontiguousarray(np.asarray(shot_ids, dtype=np.int32))
    misfit = C.c_float(0.0)
    g = [torch.zeros_like(lam) for _ in range(3)] if want_grad else [None, None, None]
    gp = [N.fptr(t) if t is not None else None for t in g]
    rc = N.lib().efwi_cufd(C.byref(misfit), *gp, N.fptr(lam), N.fptr(mu), N.fptr(rho), N.fptr(stf),
                           int(calc_id), int(gpu_id), int(ids.size),
                           ids.ctypes.data_as(C.POINTER(C.c_int)), para_fname.encode())
    N.check(rc, "efwi_cufd")
    return float(misfit.value), g


class _FwiOps:
    """The pybind module `fwi` of the reference (Src/Torch_Fwi.cpp:127-131)."""

    def forward(self, Lambda, Mu, Rho, Stf, gpu_id, Shot_ids, para_fname):
        ids = [int(i) for i in torch.as_tensor(Shot_ids).tolist()]
        if Lambda.is_cuda:
            j, _ = _run_device(Lambda, Mu, Rho, Stf, N.CALC_MISFIT, Lambda.device.index, ids, para_fname)
            return [torch.tensor([j], dtype=torch.float32, device=Lambda.device)]
        j, _ = _run_host(Lambda, Mu, Rho, Stf, N.CALC_MISFIT, gpu_id, ids, para_fname, False)
        return [torch.tensor([j], dtype=torch.float32)]

    def backward(self, Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname):
        ids = [int(i) for i in torch.as_tensor(Shot_ids).tolist()]
        _check_ngpu(int(ngpu), len(ids))
        if Lambda.is_cuda:
            j, g = _run_device(Lambda, Mu, Rho, Stf, N.CALC_GRADIENT, Lambda.device.index, ids, para_fname)
            return [torch.tensor([j], dtype=torch.float32, device=Lambda.device), g[0], g[1], g[2],
                    torch.zeros_like(Stf)]
        results = [None] * int(ngpu)
        errors = []

        def work(i, a, b):
            try:
                results[i] = _run_host(Lambda, Mu, Rho, Stf, N.CALC_GRADIENT, i, ids[a:b], para_fname, True)
            except Exception as e:  # re-raised on the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(i, a, b)) for i, (a, b) in
                   enumerate(_shot_bars(len(ids), int(ngpu)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        misfit = np.float32(0.0)
        gl, gm, gr = (torch.zeros_like(_f32c(t).cpu()) for t in (Lambda, Mu, Rho))
        for j, g in results:  # host sum over GPUs, Src/Torch_Fwi.cpp:89-94
            gl += g[0]; gm += g[1]; gr += g[2]
            misfit = np.float32(misfit + np.float32(j))
        return [torch.tensor([float(misfit)], dtype=torch.float32), gl, gm, gr, torch.zeros_like(Stf)]

    def obscalc(self, Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname):
        ids = [int(i) for i in torch.as_tensor(Shot_ids).tolist()]
        _check_ngpu(int(ngpu), len(ids))
        if Lambda.is_cuda:
            _run_device(Lambda, Mu, Rho, Stf, N.CALC_SYNTHETIC, Lambda.device.index, ids, para_fname)
            return None
        errors = []

        def work(i, a, b):
            try:
                _run_host(Lambda, Mu, Rho, Stf, N.CALC_SYNTHETIC, i, ids[a:b], para_fname, False)
            except Exception as e:
                errors.append(e)

        threads = [threading.Thread(target=work, args=(i, a, b)) for i, (a, b) in
                   enumerate(_shot_bars(len(ids), int(ngpu)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return None


fwi_ops = _FwiOps()


class FWIFunction(torch.autograd.Function):
    """Misfit as a differentiable function of lambda, mu, rho
    (Ops/FWI/FWI_ops_3d.py:34-44): the forward runs the full gradient
    pipeline and caches the gradients; the backward returns them."""

    @staticmethod
    def forward(ctx, Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname):
        outputs = fwi_ops.backward(Lambda, Mu, Rho, Stf, ngpu, Shot_ids, para_fname)
        ctx.outputs = outputs[1:]
        return outputs[0]

    @staticmethod
    def backward(ctx, grad_misfit):
        gl, gm, gr, gs = ctx.outputs
        s = grad_misfit.reshape(()).to(gl.device)
        return gl * s, gm * s, gr * s, gs * s.to(gs.device), None, None, None


class FWI_3d(torch.nn.Module):
    """Vp/Vs/Rho (kg/m^3) on the physical grid -> misfit (Ops/FWI/FWI_ops_3d.py:47-94)."""

    def __init__(self, Vp, Vs, Rho, Stf, opt, Mask=None):
        super().__init__()
        self.nz, self.nx, self.ny, self.nPml = opt['nz'], opt['nx'], opt['ny'], opt['nPml']
        self.Bounds = {}
        Vp_pad, Vs_pad, Rho_pad = ft.padding_3d(Vp, Vs, Rho, self.nz, self.nx, self.ny, self.nPml)
        self.register_buffer('Vp_ref', Vp_pad.clone().detach())
        self.register_buffer('Vs_ref', Vs_pad.clone().detach())
        self.register_buffer('Rho_ref', Rho_pad.clone().detach())
        self.Vp = nn.Parameter(Vp) if Vp.requires_grad else Vp
        self.Vs = nn.Parameter(Vs) if Vs.requires_grad else Vs
        self.Rho = nn.Parameter(Rho) if Rho.requires_grad else Rho
        if Mask is None:
            p = self.nPml
            self.Mask = torch.ones((self.nz + 2 * p, self.nx + 2 * p, self.ny + 2 * p),
                                   dtype=torch.float32, device=Vp.device)
        else:
            self.Mask = Mask
        self.Stf = Stf
        self.para_fname = opt['para_fname']

    def forward(self, Shot_ids, ngpu=1):
        Vp_pad, Vs_pad, Rho_pad = ft.padding_3d(self.Vp, self.Vs, self.Rho, self.nz, self.nx, self.ny, self.nPml)
        Vp_m = self.Mask * Vp_pad + (1.0 - self.Mask) * self.Vp_ref
        Vs_m = self.Mask * Vs_pad + (1.0 - self.Mask) * self.Vs_ref
        Rho_m = self.Mask * Rho_pad + (1.0 - self.Mask) * self.Rho_ref
        Rho = Rho_m / 1e3
        Lambda = (Vp_m ** 2 - 2.0 * Vs_m ** 2) * Rho
        Mu = Vs_m ** 2 * Rho
        return FWIFunction.apply(Lambda, Mu, Rho, self.Stf, ngpu, Shot_ids, self.para_fname)


class FWI_obscalc(torch.nn.Module):
    """Synthetic data generation (Ops/FWI/FWI_ops_3d.py:96-108); Rho in g/cc."""

    def __init__(self, Vp, Vs, Rho, Stf, opt):
        super().__init__()
        Vp_pad, Vs_pad, Rho_pad = ft.padding_3d(Vp, Vs, Rho, opt['nz'], opt['nx'], opt['ny'], opt['nPml'])
        self.Lambda = (Vp_pad ** 2 - 2.0 * Vs_pad ** 2) * Rho_pad
        self.Mu = Vs_pad ** 2 * Rho_pad
        self.Rho = Rho_pad
        self.Stf = Stf
        self.para_fname = opt['para_fname']

    def forward(self, Shot_ids, ngpu=1):
        fwi_ops.obscalc(self.Lambda, self.Mu, self.Rho, self.Stf, ngpu, Shot_ids, self.para_fname)
