"""Host-side contract (no GPU needed): the C ABI library loads and exports its
header, the host utilities equal the oracle's bit for bit, and the Python
drop-in (fwi_utils, the para/survey/.bin contract, FWI_ops_3d's host logic)
behaves like the reference's."""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def _header_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(efwi_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_header_symbol():
    from efwi import _native as N
    lib = N.lib()
    names = _header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(N.SIGNATURES), "ctypes signatures out of sync with include/*.h"
    assert lib.efwi_version() >= 100


@pytest.mark.parametrize("n,npml,h,f0,dt", [(84, 10, 10.0, 10.0, 1e-3), (240, 20, 12.5, 8.0, 1e-3),
                                            (160, 20, 12.5, 8.0, 1e-3), (360, 20, 10.0, 10.0, 8e-4)])
def test_library_cpml_profile_equals_oracle(orc, n, npml, h, f0, dt):
    from efwi import _native as N
    out = np.zeros(6 * n, np.float32)
    N.check(N.lib().efwi_cpml_profile(n, npml, h, f0, dt, out.ctypes.data_as(C.POINTER(C.c_float))))
    assert np.array_equal(out.reshape(6, n), orc.cpml_profile(n, npml, h, f0, dt))


@pytest.mark.parametrize("nt,dt,ratio", [(500, 1e-3, 0.005), (3000, 1e-3, 0.005), (3000, 1e-3, 0.001),
                                         (4000, 8e-4, 0.005)])
def test_library_taper_equals_oracle(orc, nt, dt, ratio):
    from efwi import _native as N
    out = np.zeros(nt, np.float32)
    N.check(N.lib().efwi_taper_w2(nt, dt, ratio, out.ctypes.data_as(C.POINTER(C.c_float))))
    assert np.array_equal(out, orc.taper_w2(nt, dt, ratio))


def _courant(vmax, dt, dz, dx, dy):
    from efwi import _native as N
    out = C.c_double(0.0)
    N.check(N.lib().efwi_courant_number(vmax, dt, dz, dx, dy, C.byref(out)))
    return out.value


def test_courant_number_closed_form_and_configs():
    """efwi_set_model's stability guard (Src/utilities.cu:186-202 with the 3-D
    factor sqrt(3) in place of sqrt(2)): the closed form, its scaling in vmax,
    dt and the smallest step, and the numbers of the five workloads (stable:
    0.78, 0.70 and 0.68 for 64^3/10 m, 120x200x200/12.5 m and 160x320x320/10 m)."""
    k = np.sqrt(3.0) * (9.0 / 8.0 + 1.0 / 24.0)
    assert _courant(3000.0, 1e-3, 10.0, 10.0, 10.0) == pytest.approx(3000 * float(np.float32(1e-3)) * k / 10, rel=1e-12)
    # the smallest of the three steps sets it, whichever axis it is on
    for steps in ((5.0, 10.0, 10.0), (10.0, 5.0, 10.0), (10.0, 10.0, 5.0)):
        assert _courant(3000.0, 1e-3, *steps) == pytest.approx(2 * _courant(3000.0, 1e-3, 10, 10, 10), rel=1e-12)
    # never below the reference's 2-D form sqrt(2)(9/8 + 1/24) / min(dz, dx)
    two_d = 3000.0 * float(np.float32(1e-3)) * np.sqrt(2.0) * (1 / 24 + 9 / 8) / 10.0
    assert _courant(3000.0, 1e-3, 10.0, 10.0, 10.0) > two_d
    # the workloads: vp max (anomaly included) against h and dt
    assert _courant(3500 * 1.10, 1e-3, 10, 10, 10) == pytest.approx(0.778, abs=2e-3)
    assert _courant(4000 * 1.08, 1e-3, 12.5, 12.5, 12.5) == pytest.approx(0.698, abs=2e-3)
    assert _courant(4200, 8e-4, 10, 10, 10) == pytest.approx(0.679, abs=2e-3)
    # the stability edge: vp dt / h = 1 / (sqrt(3) (9/8 + 1/24))
    edge = 10.0 / (float(np.float32(1e-3)) * k)
    assert _courant(edge * (1 - 1e-6), 1e-3, 10, 10, 10) < 1.0 < _courant(edge * (1 + 1e-6), 1e-3, 10, 10, 10)
    from efwi import _native as N
    assert N.lib().efwi_courant_number(3000.0, 0.0, 10.0, 10.0, 10.0, C.byref(C.c_double())) == -1


def test_cufd_reports_errors_instead_of_exiting(tmp_path):
    from efwi import _native as N
    lib = N.lib()
    n = 16
    a = np.ones(n ** 3, np.float32)
    fp = a.ctypes.data_as(C.POINTER(C.c_float))
    ids = np.zeros(1, np.int32)
    ip = ids.ctypes.data_as(C.POINTER(C.c_int))
    m = C.c_float(0)
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 7, 0, 1, ip, b"/nonexistent.json")
    assert rc == -1 and b"calc_id" in lib.efwi_last_error()
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(tmp_path / "nope.json").encode())
    assert rc == -3 and b"nope.json" in lib.efwi_last_error()
    para = tmp_path / "para.json"
    para.write_text(json.dumps({"nz": n, "nx": n, "ny": n, "dz": 10.0, "dx": 10.0, "dy": 10.0, "nSteps": 10,
                                "dt": 0.001, "f0": 10.0, "nPoints_pml": 4,
                                "survey_fname": str(tmp_path / "missing_survey.json"),
                                "data_dir_name": str(tmp_path)}))
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(para).encode())
    assert rc == -3 and b"missing_survey.json" in lib.efwi_last_error()
    (tmp_path / "bad.json").write_text("{\"nz\": 16, ")
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(tmp_path / "bad.json").encode())
    assert rc == -3 and b"malformed" in lib.efwi_last_error()
    # runaway nesting is refused, not recursed into
    (tmp_path / "deep.json").write_text("{\"a\": " + "[" * 100000 + "]" * 100000 + "}")
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(tmp_path / "deep.json").encode())
    assert rc == -3 and b"nesting" in lib.efwi_last_error()
    # integer keys must hold integers in int range
    base = {"nz": n, "nx": n, "ny": n, "dz": 10.0, "dx": 10.0, "dy": 10.0, "nSteps": 10, "dt": 0.001, "f0": 10.0,
            "nPoints_pml": 4, "survey_fname": str(tmp_path / "survey.json"), "data_dir_name": str(tmp_path)}
    for bad in ({"nz": 16.5}, {"nx": 1e30}, {"nSteps": -3e12}):
        para.write_text(json.dumps({**base, **bad}))
        rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(para).encode())
        assert rc == -3 and b"missing one of" in lib.efwi_last_error(), bad
    # a survey shot with a negative receiver count or a fractional coordinate
    para.write_text(json.dumps(base))
    shot = {"z_src": 2, "x_src": 8, "y_src": 8, "nrec": -1, "z_rec": [1], "x_rec": [1], "y_rec": [1]}
    (tmp_path / "survey.json").write_text(json.dumps({"shot0": shot}))
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(para).encode())
    assert rc == -3 and b"negative nrec" in lib.efwi_last_error()
    (tmp_path / "survey.json").write_text(json.dumps({"shot0": {**shot, "nrec": 1, "x_rec": [1.5]}}))
    rc = lib.efwi_cufd(C.byref(m), fp, fp, fp, fp, fp, fp, fp, 1, 0, 1, ip, str(para).encode())
    assert rc == -3 and b"z_rec/x_rec/y_rec" in lib.efwi_last_error()


def test_nn_entry_points_refuse_invalid_arguments():
    """The learned preconditioner's entry points check their arguments before
    any device work: each invalid call returns EFWI_NN_ERR_ARG (1) on the host.
    (The pointers stand for device memory that is never touched.)"""
    from efwi import _native as N
    lib = N.lib()
    P = C.c_void_p(256)   # aligned, non-NULL, never dereferenced
    Z = None
    # convolutions: NULL input, channel count not a multiple of 16, an odd
    # volume read through upsampling, bad second source or second output
    assert lib.efwi_nn_conv3(Z, 16, 16, P, Z, P, 16, 16, Z, 0, 4, 4, 4, 0, 1, Z) == 1
    assert lib.efwi_nn_conv3(P, 16, 8, P, Z, P, 16, 16, Z, 0, 4, 4, 4, 0, 1, Z) == 1
    assert lib.efwi_nn_conv3(P, 16, 16, P, Z, P, 16, 16, P, 8, 4, 4, 4, 0, 1, Z) == 1   # mask stride < channels
    assert lib.efwi_nn_conv3_cat(P, 32, 32, 1, P, 16, 16, P, Z, P, 16, 16, Z, 0, 0, Z, 0, 5, 4, 4, 0, 1, Z) == 1
    assert lib.efwi_nn_conv3_cat(P, 32, 32, 0, P, 16, 8, P, Z, P, 16, 16, Z, 0, 0, Z, 0, 4, 4, 4, 0, 1, Z) == 1
    assert lib.efwi_nn_conv3_cat(P, 32, 32, 0, Z, 0, 0, P, Z, P, 16, 16, P, 16, 8, Z, 0, 4, 4, 4, 0, 1, Z) == 1
    # weight gradients
    assert lib.efwi_nn_wgrad(P, 16, 16, P, 16, 16, Z, 0, P, Z, 4, 4, 4, 1, Z) == 1               # no scratch
    assert lib.efwi_nn_wgrad(P, 16, 16, P, 16, 16, P, 1, P, Z, 4, 4, 4, 1, Z) == 1               # scratch too small
    assert lib.efwi_nn_wgrad_cat(P, 16, 16, P, 16, 16, 0, Z, 16, 16, P, 1 << 30, P, Z, 4, 4, 4, 1, Z) == 1
    assert lib.efwi_nn_wgrad_cat(P, 16, 16, P, 16, 16, 1, P, 16, 16, P, 1 << 30, P, Z, 3, 4, 4, 1, Z) == 1
    assert lib.efwi_nn_wgrad_noalloc(P, 16, 16, P, 16, 8, P, Z, 4, 4, 4, 1, Z) == 1
    # batched flip, pooling gradient, fused loss tail
    ptrs = (C.c_void_p * 1)(256)
    ints = (C.c_int * 1)(16)
    assert lib.efwi_nn_flip_weights_n(0, ptrs, ints, ints, ptrs, Z) == 1
    assert lib.efwi_nn_flip_weights_n(17, ptrs, ints, ints, ptrs, Z) == 1
    assert lib.efwi_nn_pool2_bwd(P, 16, P, 16, 16, 3, 4, 4, 1, Z, 0, Z) == 1
    assert lib.efwi_nn_pool2_bwd(P, 16, P, 16, 16, 4, 4, 4, 1, P, 8, Z) == 1
    assert lib.efwi_nn_loss_head_bwd(P, 16, P, P, Z, P, P, 64, 0.0, P, 16, P, P, P, Z) == 1
    assert lib.efwi_nn_loss_head_bwd(P, 12, P, P, Z, P, P, 64, 1.0, P, 16, P, P, P, Z) == 1


def test_create_without_device_fails_cleanly():
    """On a machine without a HIP device, efwi_create returns an error code."""
    code = ("import sys, ctypes as C; sys.path.insert(0, %r); from efwi import _native as N; import torch;"
            "cfg = N.Config(32, 32, 32, 4, 10, 10., 10., 10., 1e-3, 10., 0, 0, 0); h = C.c_void_p();"
            "rc = N.lib().efwi_create(C.byref(cfg), C.byref(h));"
            "print('RC', rc, torch.cuda.is_available())" % os.path.join(ROOT, "efwi-3d_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rc, avail = r.stdout.split("RC")[-1].split()
    if avail == "False":
        assert int(rc) != 0


# ---------------------------------------------------------------------------
# fwi_utils (Ops/FWI/fwi_utils.py)
# ---------------------------------------------------------------------------
def test_padding_3d_is_replicate_padding():
    import fwi_utils as ft
    rng = np.random.default_rng(0)
    nz, nx, ny, p = 5, 6, 7, 3
    a = [torch.tensor(rng.standard_normal((nz, nx, ny)), dtype=torch.float32, requires_grad=True) for _ in range(3)]
    out = ft.padding_3d(*a, nz, nx, ny, p)
    for src, o in zip(a, out):
        ref = np.pad(src.detach().numpy(), p, mode="edge")
        assert o.shape == (nz + 2 * p, nx + 2 * p, ny + 2 * p)
        assert np.array_equal(o.detach().numpy(), ref)
    out[0].sum().backward()
    # each physical cell receives the count of padded cells replicating it
    g = a[0].grad.numpy()
    assert g[0, 0, 0] == (p + 1) ** 3 and g[2, 2, 2] == 1


def test_paragen_and_surveygen_formats(tmp_path):
    import fwi_utils as ft
    from efwi import io as eio
    para = tmp_path / "para.json"
    survey = tmp_path / "survey.json"
    data = tmp_path / "Data"
    ft.paraGen_3d(84, 84, 84, 10.0, 10.0, 10.0, 500, 0.001, 10.0, 10, str(para), str(survey), str(data))
    p = json.load(open(para))
    assert set(p) == {"nz", "nx", "ny", "dz", "dx", "dy", "nSteps", "dt", "f0", "nPoints_pml", "survey_fname",
                      "data_dir_name"}
    assert os.path.isdir(data)
    g = eio.grid_from_para(eio.read_para(str(para)))
    assert (g.nz, g.nx, g.ny, g.npml, g.nt) == (84, 84, 84, 10, 500) and g.dt == float(np.float32(0.001))
    ft.paraGen_3d(84, 84, 84, 10.0, 10.0, 10.0, 500, 0.001, 10.0, 10, str(para), str(survey), str(data),
                  if_win=True, scratch_dir_name=str(tmp_path / "scr"))
    p = json.load(open(para))
    assert p["if_win"] is True and p["scratch_dir_name"].endswith("scr")
    # sources on a 2 x 3 grid: shot i takes [i % 2][i // 2] (Ops/FWI/fwi_utils.py:76-80)
    xs = np.array([[1, 2, 3], [4, 5, 6]])
    zs = np.full((2, 3), 2)
    ys = np.array([[7, 8, 9], [10, 11, 12]])
    rx = np.array([3, 4, 5]); rz = np.array([1, 1, 1]); ry = np.array([6, 7, 8])
    ft.surveyGen_3d(zs, xs, ys, rz, rx, ry, str(survey))
    s = json.load(open(survey))
    assert s["nShots"] == 6
    for i in range(6):
        assert s[f"shot{i}"]["x_src"] == xs[i % 2][i // 2]
        assert s[f"shot{i}"]["y_src"] == ys[i % 2][i // 2]
        assert s[f"shot{i}"]["nrec"] == 3
    shot = eio.shot_from_survey(s, 3, 10, np.zeros(5))
    assert shot.src == (2 + 10, xs[1][1] + 10, ys[1][1] + 10)
    assert np.array_equal(shot.rec, np.stack([rz, rx, ry], 1) + 10)


def test_sourcegene_is_integrated_ricker():
    import fwi_utils as ft
    from efwi import synth
    f, nt, dt = 10.0, 500, 1e-3
    s = ft.sourceGene(f, nt, dt)
    assert np.array_equal(s, synth.ricker_integrated(f, nt, dt))
    # int_{-inf}^{t} (1 - 2e u^2) exp(-e u^2) du = t exp(-e t^2), e = pi^2 f^2
    e = np.pi ** 2 * f ** 2
    t = np.arange(nt) * dt - 1.2 / f
    exact = t * np.exp(-e * t ** 2)
    # left Riemann sum of a smooth pulse: error O(dt) relative to the peak
    assert np.max(np.abs(s - exact)) < 0.05 * np.max(np.abs(exact))
    assert np.argmax(s) == np.argmax(exact) or abs(np.argmax(s) - np.argmax(exact)) <= 1


def test_bin_data_roundtrip_and_size_check(tmp_path):
    from efwi import io as eio
    rng = np.random.default_rng(1)
    d = rng.standard_normal((3, 7, 11)).astype(np.float32)
    eio.write_shot_data(str(tmp_path), 4, d)
    for c in "xyz":  # [nrec][nt] float32 per component file
        assert os.path.getsize(tmp_path / f"Shot_{c}4.bin") == 7 * 11 * 4
    assert np.array_equal(eio.read_shot_data(str(tmp_path), 4, 7, 11), d)
    with pytest.raises(eio.ContractError):
        eio.read_shot_data(str(tmp_path), 4, 7, 12)
    with pytest.raises(eio.ContractError):
        eio.read_shot_data(str(tmp_path), 5, 7, 11)


def test_shot_bars_follow_linspace_truncation():
    import FWI_ops_3d as F
    assert F._shot_bars(10, 3) == [(0, 3), (3, 6), (6, 10)]
    assert F._shot_bars(4, 4) == [(0, 1), (1, 2), (2, 3), (3, 4)]
    with pytest.raises(F.N.EfwiError):
        F._check_ngpu(5, 4)
