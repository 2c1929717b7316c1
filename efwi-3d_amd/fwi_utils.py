"""Drop-in replacement of the reference's Ops/FWI/fwi_utils.py.

Same functions, argument order and file formats: replicate padding of the
model, the para and survey JSON writers, and the integrated Ricker wavelet.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch
import torch.nn.functional as F


def padding_3d(cp, cs, den, nz, nx, ny, nPml):
    """Replicate-pad vp, vs, rho by nPml cells on every face
    (Ops/FWI/fwi_utils.py:9-20). Returns [nz+2p][nx+2p][ny+2p] views that keep
    the autograd graph."""
    pad = (nPml, nPml, nPml, nPml, nPml, nPml)
    out = []
    for a in (cp, cs, den):
        t = F.pad(a.view(1, 1, nz, nx, ny), pad=pad, mode="replicate")
        out.append(t.view(nz + 2 * nPml, nx + 2 * nPml, ny + 2 * nPml))
    return tuple(out)


def paraGen_3d(nz, nx, ny, dz, dx, dy, nSteps, dt, f0, nPml, para_fname, survey_fname,
               data_dir_name, if_win=False, filter_para=None, if_src_update=False,
               scratch_dir_name='', if_cross_misfit=False):
    """Write the para JSON (Ops/FWI/fwi_utils.py:22-60). nz/nx/ny are the
    padded sizes. Creates data_dir_name (and scratch_dir_name if given)."""
    para = {'nz': int(nz), 'nx': int(nx), 'ny': int(ny), 'dz': float(dz), 'dx': float(dx),
            'dy': float(dy), 'nSteps': int(nSteps), 'dt': float(dt), 'f0': float(f0),
            'nPoints_pml': int(nPml)}
    if if_win is not False:
        para['if_win'] = True
    if filter_para is not None:
        para['filter'] = filter_para
    if if_src_update is not False:
        para['if_src_update'] = True
    para['survey_fname'] = survey_fname
    para['data_dir_name'] = data_dir_name
    os.makedirs(data_dir_name, exist_ok=True)
    if if_cross_misfit is not False:
        para['if_cross_misfit'] = True
    if scratch_dir_name != '':
        para['scratch_dir_name'] = scratch_dir_name
        os.makedirs(scratch_dir_name, exist_ok=True)
    with open(para_fname, 'w') as fp:
        json.dump(para, fp)


def surveyGen_3d(z_src, x_src, y_src, z_rec, x_rec, y_rec, survey_fname, Windows=None,
                 Weights=None, Src_Weights=None, Src_rxz=None, Rec_rxz=None):
    """Write the survey JSON (Ops/FWI/fwi_utils.py:63-88): the sources form a
    2-D grid of arrays [nx_src][ny_src]; shot i takes element
    [i % nx_src][i // nx_src] and every shot shares the same receivers.
    Coordinates are physical-grid indices (the native side adds nPml)."""
    x_src = np.asarray(x_src).tolist()
    z_src = np.asarray(z_src).tolist()
    y_src = np.asarray(y_src).tolist()
    x_rec = np.asarray(x_rec).tolist()
    z_rec = np.asarray(z_rec).tolist()
    y_rec = np.asarray(y_rec).tolist()
    nsrc = len(x_src) * len(x_src[0])
    nrec = len(x_rec)
    survey = {'nShots': nsrc}
    for i in range(nsrc):
        a, b = i % len(x_src), i // len(x_src)
        survey['shot' + str(i)] = {'z_src': z_src[a][b], 'x_src': x_src[a][b], 'y_src': y_src[a][b],
                                   'nrec': nrec, 'z_rec': z_rec, 'x_rec': x_rec, 'y_rec': y_rec}
    with open(survey_fname, 'w') as fp:
        json.dump(survey, fp)


def sourceGene(f, nStep, delta_t):
    """Ricker wavelet with delay 1.2/f, integrated by a cumulative sum and
    scaled by dt (Ops/FWI/fwi_utils.py:90-105). float64 numpy array."""
    e = np.pi * np.pi * f * f
    t_delay = 1.2 / f
    source = np.zeros(nStep)
    for it in range(nStep):
        t = delta_t * it - t_delay
        source[it] = (1 - 2 * e * t ** 2) * np.exp(-e * t ** 2)
    for it in range(1, nStep):
        source[it] = source[it] + source[it - 1]
    return source * delta_t
