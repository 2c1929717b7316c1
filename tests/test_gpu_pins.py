"""Source and receiver terms of the GPU path through the reference's file
contract, against the closed-form explosion (tests/analytic.py): the survey
JSON holds physical-grid coordinates to which the reader adds nPml for the
source and every receiver (Src/Src_Rec.cu:78-118), the stress source is
1500^2 s dt (Src/utilities.cu:442-479), the buoyancy 1/rho with rho in g/cc
(FWI_obscalc, Ops/FWI/FWI_ops_3d.py:96-108, Src/Model.cu:65), and slot it + 1
holds v after step it (Src/utilities.cu:481-495). A slip in any of them moves
the fitted arrival by a sample or more, or the amplitude by far more than
the 8 % the test allows."""
import numpy as np
import pytest
import torch

from analytic import Explosion, components

pytestmark = pytest.mark.gpu

N_PHYS, NPML, NT, H, DT, F0, VP, RHO = 48, 12, 260, 5.0, 1e-3, 25.0, 2000.0, 2.0
SRC = (9, 14, 11)  # physical (z, x, y)
OFFS = [(0, 10, 0), (0, 0, 14), (8, 8, 0), (-6, 6, 6), (0, 20, 5), (12, 0, -9), (10, 12, 0)]


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_obscalc_matches_exact_explosion(tmp_path, device):
    import FWI_ops_3d as F
    import fwi_utils as ft
    from efwi import io as eio
    from efwi import synth
    print(f"[gpu] explosion through the file contract, tensors on {device}", flush=True)
    n, p = N_PHYS, NPML
    para, survey, data = str(tmp_path / "para.json"), str(tmp_path / "survey.json"), str(tmp_path / "Data")
    ft.paraGen_3d(n + 2 * p, n + 2 * p, n + 2 * p, H, H, H, NT, DT, F0, p, para, survey, data)
    rec = np.array([[SRC[0] + a, SRC[1] + b, SRC[2] + c] for a, b, c in OFFS])
    assert rec.min() >= 0 and rec.max() < n
    ft.surveyGen_3d(np.array([[SRC[0]]]), np.array([[SRC[1]]]), np.array([[SRC[2]]]),
                    rec[:, 0], rec[:, 1], rec[:, 2], survey)
    stf = synth.ricker_integrated(F0, NT, DT).astype(np.float32)
    stf[0] = 0.0
    dev = torch.device(device)
    vs = VP / 1.73
    shp = (n, n, n)
    F.FWI_obscalc(torch.full(shp, VP, device=dev), torch.full(shp, vs, device=dev), torch.full(shp, RHO, device=dev),
                  torch.tensor(stf[None]), {"nz": n, "nx": n, "ny": n, "nPml": p, "para_fname": para})(
        torch.tensor([0], dtype=torch.int32))
    syn = eio.read_shot_data(data, 0, len(OFFS), NT)  # [3][nrec][nt], x y z
    A = Explosion(stf, DT, RHO, VP, H)
    fits = [(k, ci, A.fit(syn[ci, k], NT, rel, axis)) for k, ci, rel, axis in components(OFFS)]
    for k, ci, f in fits:
        print(f"receiver {k} component {'xyz'[ci]}: shift {f['shift']:+.2f} amp {f['amp']:.3f} err {f['err']:.3f}")
    assert len(fits) >= 10
    for k, ci, f in fits:
        assert abs(f["shift"]) <= 0.4 and 0.92 <= f["amp"] <= 1.08 and f["err"] <= 0.08 and f["err0"] <= 0.10, (k, ci, f)
