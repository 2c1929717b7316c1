"""The stability guard of efwi_set_model on the GPU (Src/libCUFD.cu:91,
Src/utilities.cu:186-202): a model whose Courant number exceeds 1 is refused
with EFWI_ERR_ARG instead of running into inf/NaN, the context keeps no model,
and a stable model set afterwards runs normally."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_unstable_model_is_refused_and_stable_one_runs():
    from efwi import _native as N
    from efwi import synth
    from efwi.propagator import Propagator
    w = synth.small(n=24, npml=5, nt=40)
    g = w.grid
    lam, mu, rho = (np.array(a) for a in w.model)
    vp = np.sqrt((lam.astype(np.float64) + 2 * mu) / rho).max()
    k = np.sqrt(3.0) * (9 / 8 + 1 / 24) * g.dt / min(g.dz, g.dx, g.dy)
    print(f"[gpu] guard: grid {g.shape}, vp max {vp:.1f}, Courant {vp * k:.3f}", flush=True)
    assert vp * k < 1.0
    # one cell fast enough to push the Courant number to 1.05
    c = (g.nz // 2, g.nx // 2, g.ny // 2)
    vp_c = np.sqrt((float(lam[c]) + 2 * float(mu[c])) / float(rho[c]))
    scale = (1.05 / (vp_c * k)) ** 2
    bad_lam = lam.copy()
    bad_mu = mu.copy()
    bad_lam[c] *= scale
    bad_mu[c] *= scale
    with Propagator(g) as P:
        with pytest.raises(N.EfwiError, match="Courant number 1.05"):
            P.set_model(torch.from_numpy(bad_lam), torch.from_numpy(bad_mu), torch.from_numpy(rho))
        with pytest.raises(N.EfwiError, match="efwi_set_model was not called"):
            P.run_shot(w.shots[0], 2, None, want_syn=True)
        # the same cell just below the edge is accepted and runs
        scale2 = (0.99 / (vp_c * k)) ** 2
        bad_lam[c] = lam[c] * scale2
        bad_mu[c] = mu[c] * scale2
        P.set_model(torch.from_numpy(bad_lam), torch.from_numpy(bad_mu), torch.from_numpy(rho))
        _, syn, _ = P.run_shot(w.shots[0], 2, None, want_syn=True)
        assert torch.isfinite(syn).all() and syn.abs().max() > 0
