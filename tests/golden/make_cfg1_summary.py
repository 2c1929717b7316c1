"""Writes tests/golden/cfg1_summary.npz (run from the repository root; ~2-4
minutes of CPU time):

    python tests/golden/make_cfg1_summary.py

The full-size plain-method workload (64^3 physical grid + 10 PML cells, 500
steps, 4096 receivers, efwi.synth.cfg1) through the fp32 oracle: observed
data from the anomaly-free model, then misfit and gradient of the model with
the anomaly. The arrays are too large for a fixture, so it keeps per-trace
energies, a set of full traces, per-plane sums and sums of squares of the
three gradients, and the misfits."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "efwi-3d_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

TRACES = np.array([0, 1, 63, 64, 2016, 2080, 4095])


def summarize(syn, res, misfit64, glam, gmu, grho):
    out = {"trace_energy": np.sum(np.asarray(syn, np.float64) ** 2, axis=2),
           "traces": np.asarray(syn, np.float32)[:, TRACES, :],
           "res_energy": np.sum(np.asarray(res, np.float64) ** 2, axis=2),
           "misfit64": np.float64(misfit64)}
    for k, g in (("glam", glam), ("gmu", gmu), ("grho", grho)):
        g = np.asarray(g, np.float64)
        out[k + "_plane_sum"] = g.sum(axis=(1, 2))
        out[k + "_plane_sq"] = (g ** 2).sum(axis=(1, 2))
    return out


def main():
    import oracle
    from efwi import synth
    from util import oracle_run
    w = synth.cfg1()
    obs = [o.astype(np.float32) for o in oracle_run(oracle, w, model=w.true, calc_id=2).syn]
    r = oracle_run(oracle, w, obs=obs, calc_id=1)
    out = summarize(r.syn[0], r.res[0], r.misfit64, r.glam, r.gmu, r.grho)
    out["obs_trace_energy"] = np.sum(np.asarray(obs[0], np.float64) ** 2, axis=2)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg1_summary.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes; misfit", r.misfit64, "loop seconds", r.seconds)


if __name__ == "__main__":
    main()
