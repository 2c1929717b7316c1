"""The reference's Python op API end to end on the GPU, through its file
contract (para/survey JSON written by fwi_utils, Shot_*.bin data):
FWI_obscalc writes the observed data, FWI_3d + FWIFunction returns the misfit
and autograd gradients through padding, mask and the lambda/mu/rho formulas.
Checked against the oracle on the same inputs."""
import json
import os

import numpy as np
import pytest
import torch

from util import rel_l2

pytestmark = pytest.mark.gpu


def _experiment(tmp_path, n=20, npml=5, nt=120, if_win=False):
    import fwi_utils as ft
    from efwi import synth
    nz = nx = ny = n
    rng = np.random.default_rng(7)
    vp = synth.layered(nz, nx, ny, [2200.0, 2700.0, 3100.0], [6, 13]).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    rho = (310.0 * vp ** 0.25).astype(np.float32)  # kg/m^3
    vp0 = (vp * (1 + 0.03 * rng.standard_normal(vp.shape))).astype(np.float32)
    para = str(tmp_path / "para.json")
    survey = str(tmp_path / "survey.json")
    data = str(tmp_path / "Data")
    ft.paraGen_3d(nz + 2 * npml, nx + 2 * npml, ny + 2 * npml, 10.0, 10.0, 10.0, nt, 0.001, 15.0, npml, para,
                  survey, data, if_win=if_win, scratch_dir_name=str(tmp_path / "scratch"))
    xs = np.array([[5, 14]]); zs = np.array([[2, 2]]); ys = np.array([[10, 10]])
    gx, gy = np.meshgrid(np.arange(1, nx - 1, 3), np.arange(1, ny - 1, 3), indexing="ij")
    rx, ry = gx.ravel(), gy.ravel()
    rz = np.full(rx.shape, 1)
    ft.surveyGen_3d(zs, xs, ys, rz, rx, ry, survey)
    if if_win:
        s = json.load(open(survey))
        for k in range(s["nShots"]):
            m = s[f"shot{k}"]["nrec"]
            s[f"shot{k}"]["win_start"] = list(np.linspace(0.0, 0.02, m))
            s[f"shot{k}"]["win_end"] = list(np.linspace(0.09, 0.11, m))
            s[f"shot{k}"]["weights"] = list(np.linspace(0.5, 1.5, m))
            s[f"shot{k}"]["src_weight"] = 0.8 + 0.1 * k
        json.dump(s, open(survey, "w"))
    stf = np.stack([synth.ricker_integrated(15.0, nt, 1e-3)] * 2).astype(np.float32)
    opt = {"nz": nz, "nx": nx, "ny": ny, "nPml": npml, "para_fname": para}
    return dict(vp=vp, vs=vs, rho=rho, vp0=vp0, stf=stf, opt=opt, para=para, survey=survey, data=data, nt=nt,
                npml=npml)


def _oracle_reference(orc, e, shot_ids):
    """Misfit and gradients of the same experiment with the oracle."""
    from efwi import io as eio
    from efwi import synth
    p = e["npml"]
    grid = eio.grid_from_para(eio.read_para(e["para"]))
    g = orc.Grid(grid.nz, grid.nx, grid.ny, p, grid.nt, grid.dz, grid.dx, grid.dy, grid.dt, grid.f0)
    survey = eio.read_survey(e["survey"])
    shots = [eio.shot_from_survey(survey, s, p, e["stf"][s], grid.if_win) for s in shot_ids]
    osh = orc.Shots(src=np.array([s.src for s in shots]), stf=np.stack([s.stf for s in shots]),
                    rec=[s.rec for s in shots],
                    win_start=[s.win_start for s in shots] if grid.if_win else [],
                    win_end=[s.win_end for s in shots] if grid.if_win else [],
                    weights=[s.weights for s in shots] if grid.if_win else [],
                    src_weight=np.array([s.src_weight for s in shots], np.float32))
    true = synth.lame(synth.pad_replicate(e["vp"], p), synth.pad_replicate(e["vs"], p),
                      synth.pad_replicate(e["rho"], p))
    obs = orc.run(g, osh, *true, calc_id=2).syn
    obs = [o.astype(np.float32) for o in obs]
    cur = synth.lame(synth.pad_replicate(e["vp0"], p), synth.pad_replicate(e["vs"], p),
                     synth.pad_replicate(e["rho"], p))
    return obs, orc.run(g, osh, *cur, obs=obs, calc_id=1)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
@pytest.mark.parametrize("if_win", [False, True])
def test_fwi3d_through_files_matches_oracle(orc, tmp_path, device, if_win):
    import FWI_ops_3d as F
    from efwi import io as eio
    print(f"[gpu] drop-in device={device} if_win={if_win}", flush=True)
    e = _experiment(tmp_path, if_win=if_win)
    ids = torch.tensor([0, 1], dtype=torch.int32)
    dev = torch.device(device)
    obscalc = F.FWI_obscalc(torch.tensor(e["vp"], device=dev), torch.tensor(e["vs"], device=dev),
                            torch.tensor((e["rho"] / np.float32(1e3)).astype(np.float32), device=dev), torch.tensor(e["stf"]), e["opt"])
    obscalc(ids, ngpu=1)
    obs_ref, ref = _oracle_reference(orc, e, [0, 1])
    survey = eio.read_survey(e["survey"])
    for k in (0, 1):
        got = eio.read_shot_data(e["data"], k, survey[f"shot{k}"]["nrec"], e["nt"])
        assert rel_l2(got, obs_ref[k]) <= 1e-5
    vp = torch.tensor(e["vp0"], device=dev, requires_grad=True)
    fwi = F.FWI_3d(vp, torch.tensor(e["vs"], device=dev), torch.tensor(e["rho"], device=dev), torch.tensor(e["stf"]),
                   e["opt"])
    loss = fwi(ids, ngpu=1)
    loss.backward()
    assert abs(loss.item() - ref.misfit64) <= 1e-6 * ref.misfit64
    # chain rule of the op's parameterisation: dJ/dvp = dJ/dlam * 2 vp rho + dJ/dmu * 0 (vs fixed)
    p = e["npml"]
    from efwi import synth
    rho_pad = synth.pad_replicate(e["rho"], p).astype(np.float32) / np.float32(1e3)
    vp_pad = synth.pad_replicate(e["vp0"], p).astype(np.float32)
    gl_pad = ref.glam * 2 * vp_pad * rho_pad
    # replicate padding sends the padded cells' gradient back to the edge cells
    t = torch.tensor(gl_pad, dtype=torch.float64)
    n = e["vp"].shape[0]
    back = torch.zeros((n, n, n), dtype=torch.float64, requires_grad=True)
    padded = torch.nn.functional.pad(back.view(1, 1, n, n, n), (p,) * 6, mode="replicate").view(t.shape)
    (padded * t).sum().backward()
    assert rel_l2(fwi.Vp.grad.cpu().numpy(), back.grad.numpy()) <= 1e-5


def test_cufd_scratch_files_and_gradient_scaling(orc, tmp_path):
    """efwi_cufd (host path) writes Residual/Syn/CondObs scratch files; the
    autograd backward scales the cached gradients by grad_output."""
    import FWI_ops_3d as F
    from efwi import io as eio
    e = _experiment(tmp_path)
    ids = torch.tensor([1], dtype=torch.int32)
    F.FWI_obscalc(torch.tensor(e["vp"]), torch.tensor(e["vs"]), torch.tensor((e["rho"] / np.float32(1e3)).astype(np.float32)),
                  torch.tensor(e["stf"]), e["opt"])(ids)
    vp = torch.tensor(e["vp0"], requires_grad=True)
    fwi = F.FWI_3d(vp, torch.tensor(e["vs"]), torch.tensor(e["rho"]), torch.tensor(e["stf"]), e["opt"])
    loss = fwi(ids)
    (3.0 * loss).backward()
    g3 = fwi.Vp.grad.clone()
    fwi.Vp.grad = None
    fwi(ids).backward()
    assert torch.allclose(g3, 3.0 * fwi.Vp.grad, rtol=1e-6, atol=0)
    _, ref = _oracle_reference(orc, e, [1])
    m = eio.read_survey(e["survey"])["shot1"]["nrec"]
    for c in "xyz":
        f = tmp_path / "scratch" / f"Residual_Shot_{c}1.bin"
        assert f.exists() and os.path.getsize(f) == m * e["nt"] * 4
    rx = np.fromfile(tmp_path / "scratch" / "Residual_Shot_x1.bin", np.float32).reshape(m, e["nt"])
    assert rel_l2(rx, ref.res[0][0]) <= 1e-5


@pytest.mark.parametrize("if_win", [False, True])
def test_device_path_writes_the_host_paths_scratch_files(tmp_path, if_win):
    """With CUDA tensors the drop-in op runs the device-resident path; with a
    scratch_dir_name in the para file it writes the same Residual/Syn/CondObs
    files as the host path (efwi_cufd, Src/libCUFD.cu:660-701), bit for bit."""
    import FWI_ops_3d as F
    print(f"[gpu] scratch files host vs device, if_win={if_win}", flush=True)
    files = {}
    for dev in ("cpu", "cuda"):
        d = tmp_path / dev
        d.mkdir()
        e = _experiment(d, if_win=if_win)
        ids = torch.tensor([0, 1], dtype=torch.int32)
        F.FWI_obscalc(torch.tensor(e["vp"]), torch.tensor(e["vs"]),
                      torch.tensor((e["rho"] / np.float32(1e3)).astype(np.float32)), torch.tensor(e["stf"]),
                      e["opt"])(ids)
        dv = torch.device(dev)
        fwi = F.FWI_3d(torch.tensor(e["vp0"], device=dv), torch.tensor(e["vs"], device=dv),
                       torch.tensor(e["rho"], device=dv), torch.tensor(e["stf"]), e["opt"])
        fwi(ids)
        out = {}
        for p in sorted((d / "scratch").iterdir()):
            out[p.name] = np.fromfile(p, np.float32)
        files[dev] = out
    assert len(files["cpu"]) == 2 * 3 * 3 and set(files["cpu"]) == set(files["cuda"])
    for name, a in files["cpu"].items():
        assert np.array_equal(a, files["cuda"][name]), name
        assert np.abs(a).max() > 0, name
