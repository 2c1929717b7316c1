"""The reference's on-disk contract: para / survey JSON and per-shot .bin data.

  para JSON    written by fwi_utils.paraGen_3d, read by Parameter
               (Src/Parameter.cpp:41-184); nz/nx/ny are padded sizes.
  survey JSON  written by fwi_utils.surveyGen_3d, read by Src_Rec
               (Src/Src_Rec.cu:70-206); coordinates are physical-grid indices,
               the padded index adds nPoints_pml.
  data         data_dir/Shot_{x,y,z}<id>.bin: float32 [nrec][nSteps]
               (read at Src/libCUFD.cu:277-308, written at :704-725).
"""
from __future__ import annotations

import json
import os

import numpy as np

from .propagator import GridSpec, ShotSpec


class ContractError(RuntimeError):
    """A para/survey/data file that does not meet the reference's contract."""


_PARA_KEYS_INT = ("nz", "nx", "ny", "nSteps", "nPoints_pml")
_PARA_KEYS_NUM = ("dz", "dx", "dy", "dt", "f0")


def read_para(fname: str) -> dict:
    try:
        with open(fname) as fp:
            p = json.load(fp)
    except OSError as e:
        raise ContractError(f"cannot open parameter file {fname}: {e}") from e
    for k in _PARA_KEYS_INT + _PARA_KEYS_NUM + ("survey_fname", "data_dir_name"):
        if k not in p:
            raise ContractError(f"{fname}: missing key {k!r}")
    return p


def grid_from_para(p: dict, formulation: int = 0) -> GridSpec:
    f32 = lambda v: float(np.float32(v))
    return GridSpec(int(p["nz"]), int(p["nx"]), int(p["ny"]), int(p["nPoints_pml"]), int(p["nSteps"]),
                    f32(p["dz"]), f32(p["dx"]), f32(p["dy"]), f32(p["dt"]), f32(p["f0"]),
                    formulation=formulation, if_win=bool(p.get("if_win", False)))


def read_survey(fname: str) -> dict:
    try:
        with open(fname) as fp:
            return json.load(fp)
    except OSError as e:
        raise ContractError(f"cannot open survey file {fname}: {e}") from e


def shot_from_survey(survey: dict, shot_id: int, npml: int, stf_row, if_win: bool = False) -> ShotSpec:
    key = f"shot{int(shot_id)}"
    if key not in survey:
        raise ContractError(f"survey has no object {key}")
    s = survey[key]
    for k in ("z_src", "x_src", "y_src", "nrec", "z_rec", "x_rec", "y_rec"):
        if k not in s:
            raise ContractError(f"{key}: missing {k!r}")
    n = int(s["nrec"])
    rec = np.stack([np.asarray(s["z_rec"][:n]), np.asarray(s["x_rec"][:n]),
                    np.asarray(s["y_rec"][:n])], axis=1).astype(np.int32) + npml
    kw = {}
    if if_win:
        if "win_start" not in s or "win_end" not in s:
            raise ContractError(f"{key}: if_win needs win_start and win_end")
        kw["win_start"] = np.asarray(s["win_start"][:n], np.float32)
        kw["win_end"] = np.asarray(s["win_end"][:n], np.float32)
    if "weights" in s:
        kw["weights"] = np.asarray(s["weights"][:n], np.float32)
    kw["src_weight"] = float(np.float32(s.get("src_weight", 1.0)))
    return ShotSpec((int(s["z_src"]) + npml, int(s["x_src"]) + npml, int(s["y_src"]) + npml),
                    np.asarray(stf_row, np.float32), rec, **kw)


def shot_files(data_dir: str, shot_id: int, prefix: str = "Shot_"):
    return [os.path.join(data_dir, f"{prefix}{c}{int(shot_id)}.bin") for c in "xyz"]


def read_shot_data(data_dir: str, shot_id: int, nrec: int, nt: int) -> np.ndarray:
    """[3][nrec][nt] float32 from Shot_{x,y,z}<id>.bin."""
    out = np.empty((3, nrec, nt), np.float32)
    for c, f in enumerate(shot_files(data_dir, shot_id)):
        try:
            a = np.fromfile(f, dtype=np.float32)
        except OSError as e:
            raise ContractError(f"cannot read {f}: {e}") from e
        if a.size != nrec * nt:
            raise ContractError(f"{f}: expected {nrec * nt} floats, found {a.size}")
        out[c] = a.reshape(nrec, nt)
    return out


def write_shot_data(data_dir: str, shot_id: int, data, prefix: str = "Shot_") -> None:
    data = np.asarray(data, np.float32)
    for c, f in enumerate(shot_files(data_dir, shot_id, prefix)):
        np.ascontiguousarray(data[c]).tofile(f)
