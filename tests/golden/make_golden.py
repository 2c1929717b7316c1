"""Writes the golden fixtures of the oracle (run from the repository root):

    python tests/golden/make_golden.py

Each fixture holds, for one seeded small workload (efwi.synth.small), a
SHA-256 of the float32 inputs (lambda, mu, rho of the true and current
models, source time function, receiver table) and the oracle's fp32 outputs:
the observed data (forward of the true model), the synthetic data, the
adjoint source, the misfit (FP64 sum and the reference's FP32 block sum) and
the lambda/mu/rho gradients on the interior. The reference itself ships no
vectors for this path (it has no tests), so these fixtures freeze the
oracle's restatement: tests check the oracle and the GPU library against them.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "efwi-3d_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

CASES = {
    "small_reference": dict(n=24, npml=4, nt=110, nrec_side=8, seed=3, formulation=0),
    "small_consistent": dict(n=24, npml=4, nt=110, nrec_side=8, seed=4, formulation=1),
}


def input_digest(w):
    h = hashlib.sha256()
    for a in (*w.true, *w.model):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    for s in w.shots:
        h.update(np.ascontiguousarray(s.stf, np.float32).tobytes())
        h.update(np.ascontiguousarray(s.rec, np.int32).tobytes())
        h.update(np.asarray(s.src, np.int32).tobytes())
    return h.hexdigest()


def compute(name):
    import oracle
    from efwi import synth
    from util import oracle_run
    kw = CASES[name]
    w = synth.small(**kw)
    obs = oracle_run(oracle, w, model=w.true, calc_id=2).syn
    obs32 = [o.astype(np.float32) for o in obs]
    r = oracle_run(oracle, w, obs=obs32, calc_id=1)
    p = w.grid.npml
    inner = (slice(p, -p),) * 3
    return w, dict(digest=np.array(input_digest(w)), obs=obs32[0], syn=r.syn[0].astype(np.float32),
                   res=r.res[0].astype(np.float32), misfit64=np.float64(r.misfit64),
                   misfit32=np.float32(r.misfit32), glam=r.glam[inner].astype(np.float32),
                   gmu=r.gmu[inner].astype(np.float32), grho=r.grho[inner].astype(np.float32))


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for name in CASES:
        _, out = compute(name)
        path = os.path.join(here, name + ".npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
