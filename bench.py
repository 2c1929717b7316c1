"""Benchmark: fwd+adj+grad of the single-shot production configuration on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2] [--nt NT]
                    [--no-cpu-baseline] [--no-profile] [--no-autotune]

A "step" is one gradient evaluation of one shot: the forward loop, the misfit
and adjoint source, and the backward loop (reverse reconstruction + adjoint +
lambda/mu/rho imaging) over all nt time steps of the 200x200x120 (padded
240x240x160), 3000-step workload. Inputs (model, observed data) are resident
in HBM before the timed region. Prints one JSON line.

Other workloads: --workload cfg3 (16 shots, batched launches against the same
shots one by one, in shots/s) and --workload cfg4 (320x320x160, 4000 steps, 4
shots one after another with sqrt(nt) checkpoints; a step is the summed
gradient of the 4 shots).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "efwi-3d_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic FP32 traffic with an ideal cache (every field read and written
# once) of the reference's kernels, in floats per cell: (per padded cell, per
# interior cell). The backward kernels here fuse the reverse and the adjoint
# kernel of the reference: velocity_bwd = velocity(rev) 18 + velocity_adj 15,
# stress_bwd = stress(rev) 27 + stress_adj 18.
ALG_FLOATS = {"stress_fwd": (17, 0), "velocity_fwd": (13, 0), "velocity_bwd": (15, 18),
              "stress_bwd": (18, 27), "velocity_adj": (15, 0), "stress_adj": (18, 0)}


def alg_bytes_per_launch(kernel, grid):
    per_pad, per_int = ALG_FLOATS[kernel]
    return 4.0 * (per_pad * grid.ncells + per_int * grid.interior_cells())


def alg_bytes_pipeline(grid):
    npad, nint, nt = grid.ncells, grid.interior_cells(), grid.nt
    return 120.0 * npad * (nt - 1) + 180.0 * nint * (nt - 1) + 132.0 * npad * nt


def ring_bytes(grid):
    """Bytes of the boundary ring of all nt - 1 steps (5-cell band around the
    C-PML-free interior, 9 fields), as the library sizes it."""
    p = grid.npml
    nzb, nxb, nyb = (n - 2 * p + 4 for n in grid.shape)
    band = 10 * nxb * nyb + 10 * (nzb - 10) * nyb + 10 * (nzb - 10) * (nxb - 10)
    return band * 9 * (grid.nt - 1) * 4


def hbm_traffic(kernel, grid, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/hbm_traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, each scaled by the factor the same counter reads for a plain
    copy of known size), or None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("workload") != workload or tuple(t.get("grid", ())) != tuple(grid.shape):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return 1024.0 * (k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"])


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, nsteps):
    """The CPU oracle (fp32, OpenMP over all host threads) on the first
    `nsteps` time steps of the same workload (fwd + adjoint + gradient)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    g = w.grid
    og = oracle.Grid(g.nz, g.nx, g.ny, g.npml, nsteps, g.dz, g.dx, g.dy, g.dt, g.f0)
    sh = w.shots[0]
    osh = oracle.Shots(src=np.array([sh.src]), stf=sh.stf[None, :nsteps], rec=[sh.rec])
    obs = [np.zeros((3, sh.nrec, nsteps), np.float32)]
    r = oracle.run(og, osh, *w.model, obs=obs, calc_id=1, precision="fp32")
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    gcell = g.ncells * nsteps / r.seconds / 1e9
    return {"value": gcell, "unit": "Gcell-updates/s", "cores": cores, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"{w.name} grid {g.nz}x{g.nx}x{g.ny}, first {nsteps} of {g.nt} time steps, "
                      f"fwd+adj+grad, 1 shot, fp32 oracle, {r.seconds:.1f} s"}


def bench_batch(args, w):
    """Batched shots (cfg3): all shots through one run_batch call, timed in
    each of its modes (the automatic choice; persistent (shot, tile) launches
    over groups of shots with one captured hipGraph per time step; one shot
    after another on the single-shot kernels), the best one reported; the same
    shots through run_shot one after another for comparison."""
    import torch
    from efwi.propagator import Propagator
    g = w.grid
    n = len(w.shots)
    print(f"[bench] {w.name}: padded grid {g.shape}, nt={g.nt}, {n} shots, nrec {min(s.nrec for s in w.shots)}-"
          f"{max(s.nrec for s in w.shots)}", file=sys.stderr, flush=True)
    P = Propagator(g, 0)
    P.set_model(*(torch.from_numpy(a) for a in w.true))
    P.batch_group(-1)
    _, obs = P.run_batch(w.shots, 2)
    P.set_model(*(torch.from_numpy(a) for a in w.model))
    if not args.no_autotune:
        tune = P.autotune(reps=3)
        print("[bench] autotune:", {k: min(v, key=v.get) for k, v in tune.items()}, file=sys.stderr, flush=True)
    tuned = P.tiles()
    kinds = ("stress_fwd", "velocity_fwd", "stress_bwd", "velocity_bwd")

    def timed(cfg):
        group, tile, take = cfg
        for k in kinds:
            P.set_tile(k, *(tuned[k] if tile is None else tile))
        P.batch_group(group)
        P.batch_config(take, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.run_batch(w.shots, 1, obs)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def label(c):
        mode = {0: "auto", -1: "single-shot kernels"}.get(c[0], f"persistent groups of {c[0]}")
        return mode + ("" if c[1] is None else f", tile {c[1][0]}x{c[1][1]}, take {c[2]}")

    trials = {}
    cands = [(0, None, 2), (-1, None, 2)] + [(grp, tile, take) for grp in (n, max(1, n // 4))
                                              for tile in ((2, 1), (2, 2)) for take in (1, 4)]
    cands += [(1, (2, 2), 1)]
    for cfg in cands:
        timed(cfg)  # warm-up (graph capture)
        trials[cfg] = timed(cfg)
        print(f"[bench] {label(cfg)}: {n / trials[cfg]:.3f} shots/s", file=sys.stderr, flush=True)
    best = min(trials, key=trials.get)
    for _ in range(args.warmup):
        timed(best)
    t0 = time.perf_counter()
    for k in range(args.steps):
        dt = timed(best)
        print(f"[bench] step {k}: {dt:.3f} s", file=sys.stderr, flush=True)
    dt_step = (time.perf_counter() - t0) / args.steps
    for k, tile in tuned.items():
        P.set_tile(k, *tile)
    # the same shots one by one through run_shot (one timed pass, profiled
    # afterwards for the per-kernel roofline of the single-shot kernels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P.zero_gradient()
    for s, o in zip(w.shots, obs):
        P.run_shot(s, 1, o)
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t0
    roof, kernels = None, {}
    if not args.no_profile:
        P.profile(True)
        P.profile_read(reset=True)
        P.zero_gradient()
        for s, o in zip(w.shots, obs):
            P.run_shot(s, 1, o)
        prof = P.profile_read(reset=True)
        P.profile(False)
        for k, (ms, cnt) in prof.items():
            if k in ALG_FLOATS and cnt:
                kernels[k] = {"avg_us": round(ms / cnt * 1e3, 2), "launches": cnt,
                              "alg_GBps": round(alg_bytes_per_launch(k, g) / (ms / cnt * 1e-3) / 1e9, 1)}
        if kernels:
            dom = max(kernels, key=lambda k: prof[k][0])
            ach = kernels[dom]["alg_GBps"]
            roof = {"bound": "hbm", "kernel": dom + " (single-shot kernels, per shot)", "achieved": ach,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "alg_bytes": round(alg_bytes_per_launch(dom, g))}
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        c = cpu_baseline(w, min(args.cpu_steps, g.nt))
        # the oracle's cell-update rate as shots of this workload per second
        cpu = dict(c, value=round(c["value"] * 1e9 / (g.ncells * g.nt), 6), unit="shots/s",
                   gcell_updates_per_s=c["value"])
    pipe_gbs = alg_bytes_pipeline(g) * n / dt_step / 1e9
    out = {
        "metric": "batched fwd+adj+grad shots/s",
        "value": round(n / dt_step, 4),
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {g.nz - 2 * g.npml}x{g.nx - 2 * g.npml}x{g.ny - 2 * g.npml} physical + "
                               f"{g.npml} PML cells, padded {g.nz}x{g.nx}x{g.ny}, nt={g.nt}, {n} shots in one "
                               f"run_batch call, {sum(s.nrec for s in w.shots)} receivers in all, {label(best)}",
                   "batch_trials_shots_per_s": {label(c): round(n / t, 4) for c, t in trials.items()},
                   "gcell_updates_per_s": round(g.ncells * g.nt * n / dt_step / 1e9, 3),
                   "pipeline_alg_GBps": round(pipe_gbs, 1),
                   "pipeline_hbm_frac": round(pipe_gbs / HBM_PEAK_GBS, 4),
                   "sequential_shots_per_s": round(n / dt_seq, 4),
                   "kernels": kernels},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    P.close()


def bench_learned(args):
    """Learned preconditioner: training pairs from the plain method on random
    models, training on the library's BF16 MFMA kernels, then inversions of
    held-out models with plain L-BFGS and with the learned preconditioner."""
    import numpy as np
    import torch
    from efwi import learned as L
    from efwi.inversion import lbfgs_invert
    from efwi.propagator import Propagator
    log = lambda m: print(m, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    tdims = (96, 128, 128)
    data, fdims, t_gen = L.generate(args.train_models, dims=tdims, npml=20, nt=600, seed=11, log=log)
    log(f"[bench] {len(data)} training pairs at {tdims} in {t_gen:.1f} s")
    net, losses, t_train = L.train(data, fdims, epochs=args.epochs, log=log)
    # kernel throughput of a training step, timed with events on torch's stream
    V = fdims[0] * fdims[1] * fdims[2]
    mask = torch.ones(V, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    nstep = 10
    for k in range(nstep):
        feat, _, logs = data[k % len(data)]
        net.train_step(feat, logs, mask, V, fdims, lr=0.0, log_scaling=True)
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / nstep
    from efwi.unet import training_flops
    tflops = training_flops(fdims) / (step_ms * 1e-3) / 1e12
    del data
    torch.cuda.empty_cache()
    # held-out inversions at the production size
    edims, npml = (120, 200, 200), 20
    rng = np.random.default_rng(2024)
    results = []
    for m in range(args.eval_models):
        survey = L.surface_survey(*edims, npml, args.eval_nt, 12.5, 1e-3, 8.0, 8, rng)
        vp, vs, rho, vp0 = L.random_model(rng, *edims, sigma=8.0)
        if m < args.eval_skip:  # drawn (the sequence stays the same) but not inverted
            continue
        row = {}
        for method in ("plain", "learned"):
            prop = Propagator(survey.grid, 0)
            prob = L.make_problem(prop, survey, vp, vs, rho, dev)
            v0 = L._padded(vp0, npml, dev)
            if method == "plain":
                _, hist = lbfgs_invert(prob, v0, args.iters, log=log)
                counts = None
            else:
                _, hist, counts = L.learned_invert(prob, v0, args.iters, net, log=log)
            prop.close()
            row[method] = {"f": hist.f, "t": hist.t, "nfev": hist.nfev[-1], "counts": counts}
            log(f"[bench] model {m} {method}: misfit {hist.f[0]:.4e} -> {hist.f[-1]:.4e} in {hist.t[-1]:.1f} s, "
                f"{hist.nfev[-1]} evaluations {counts or ''}")
        target = row["plain"]["f"][-1]
        reach = next((t for f, t in zip(row["learned"]["f"], row["learned"]["t"]) if f <= target), None)
        row["time_ratio"] = None if reach is None else reach / row["plain"]["t"][-1]
        results.append(row)
    ratios = [r["time_ratio"] for r in results]
    reached = [x for x in ratios if x is not None]
    out = {
        "metric": "wall time of the learned-preconditioned inversion to the plain L-BFGS final misfit, "
                  "relative to plain L-BFGS",
        "value": round(float(np.mean(reached)), 4) if ratios and len(reached) == len(ratios) else None,
        "unit": "ratio",
        "n_gpus": 1, "steps": args.iters, "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network), fp32 (stencil)",
        "data": "synthetic",
        "config": {"workload": f"cfg5: U-Net (3 levels, 16 base channels, 3x3x3, BF16 MFMA, {net.nparams()} "
                               f"parameters) trained on {args.train_models} random layered+anomaly models at "
                               f"{tdims[0]}x{tdims[1]}x{tdims[2]} ({args.epochs} epochs); {args.eval_models - args.eval_skip} "
                               f"held-out models (from model {args.eval_skip}) at {edims[0]}x{edims[1]}x{edims[2]} + {npml} PML, 8 shots, "
                               f"nt={args.eval_nt}, {args.iters} iterations each method",
                   "time_ratios": ratios,
                   "final_misfit_ratio_learned_over_plain": [r["learned"]["f"][-1] / r["plain"]["f"][-1] for r in results],
                   "plain": [{"f_rel": [round(x / r["plain"]["f"][0], 5) for x in r["plain"]["f"]],
                              "seconds": round(r["plain"]["t"][-1], 2), "nfev": r["plain"]["nfev"]} for r in results],
                   "learned": [{"f_rel": [round(x / r["learned"]["f"][0], 5) for x in r["learned"]["f"]],
                                "seconds": round(r["learned"]["t"][-1], 2), "nfev": r["learned"]["nfev"],
                                "steps": r["learned"]["counts"]} for r in results],
                   "training": {"pairs_s": round(t_gen, 1), "train_s": round(t_train, 1),
                                "loss_first": losses[0], "loss_last_epoch": float(np.mean(losses[-args.train_models:]))}},
        "roofline": {"bound": "mfma", "kernel": "U-Net training step (conv3 forward/input/weight gradients)",
                     "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 5), "traffic": None, "step_ms": round(step_ms, 2)},
        "cpu_baseline": None if args.no_cpu_baseline else unet_cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


def unet_cpu_baseline(dims=(96, 128, 128), cin=16, cout=16):
    """The CPU oracle (oracle/nn_oracle.py, numpy in float64) on a
    bounded sample of the U-Net's work: the forward, input-gradient and
    weight-gradient convolutions of one 16 -> 16 layer at the full 96x128x128,
    reported in the roofline's unit (TFLOP/s of matrix-core-equivalent work)."""
    import time
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nn_oracle as O
    rng = np.random.default_rng(0)
    x = rng.standard_normal((*dims, cin))
    w = rng.standard_normal((27, cout, cin)) * 0.1
    g = rng.standard_normal((*dims, cout))
    t0 = time.perf_counter()
    O.conv3(x, w)
    O.conv3_dgrad(g, w)
    O.conv3_wgrad(g, x)
    sec = time.perf_counter() - t0
    flops = 3 * 2.0 * 27 * cin * cout * dims[0] * dims[1] * dims[2]
    try:  # the matrix products run on numpy's BLAS threads
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except ImportError:
        cores = 1
    return {"value": flops / sec / 1e12, "unit": "TFLOP/s", "cores": cores, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"nn_oracle forward + input gradient + weight gradient of one {cin}->{cout} 3x3x3 convolution "
                      f"at {dims[0]}x{dims[1]}x{dims[2]} voxels (float64, numpy matrix products), {sec:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--nt", type=int, default=None, help="override the number of time steps")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--train-models", type=int, default=256, help="cfg5: training models")
    ap.add_argument("--epochs", type=int, default=6, help="cfg5: training epochs")
    ap.add_argument("--eval-models", type=int, default=2, help="cfg5: held-out models")
    ap.add_argument("--eval-nt", type=int, default=800, help="cfg5: time steps of the held-out inversions")
    ap.add_argument("--eval-skip", type=int, default=0,
                    help="cfg5: draw but do not invert the first K held-out models (to split a long evaluation)")
    ap.add_argument("--iters", type=int, default=20, help="cfg5: inversion iterations")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tile", action="append", default=[], metavar="KERNEL=ROWSxZ",
                    help="fix a launch shape, e.g. stress_fwd=2x4 (repeatable)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (scale of the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 2 ** 20 // 4
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4,
                          "copies": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this deployment has one GPU and the path "
              "shards by shot across GPUs (see the multi-GPU design); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from efwi import synth
    from efwi.propagator import Propagator

    if args.workload == "cfg5":
        return bench_learned(args)
    w = synth.WORKLOADS[args.workload](**({} if args.nt is None else {"nt": args.nt}))
    g = w.grid
    if len(w.shots) > 1 and g.memory_mode != 1:
        return bench_batch(args, w)
    # one step: the summed gradient of every shot of the workload, shot after shot
    shots = w.shots
    nrec = sum(s.nrec for s in shots)
    print(f"[bench] {w.name}: padded grid {g.shape}, nt={g.nt}, {len(shots)} shot(s), nrec={nrec}",
          file=sys.stderr, flush=True)
    P = Propagator(g, 0)
    # observed data from the true model (untimed set-up)
    P.set_model(*(torch.from_numpy(a) for a in w.true))
    obs = [P.run_shot(s, 2)[1] for s in shots]
    P.set_model(*(torch.from_numpy(a) for a in w.model))
    torch.cuda.synchronize()

    tune = None
    if not args.no_autotune:
        tune = P.autotune(reps=3)
        print("[bench] autotune:", {k: min(v, key=v.get) for k, v in tune.items()}, file=sys.stderr, flush=True)
    for spec in args.tile:
        k, shape = spec.split("=")
        r, z = shape.split("x")
        P.set_tile(k, int(r), int(z))
    tiles = P.tiles()

    def step():
        P.zero_gradient()
        return sum(P.run_shot(s, 1, o)[0] for s, o in zip(shots, obs))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print(f"[bench] warmup done, device bytes {P.device_bytes / 2**30:.1f} GiB", file=sys.stderr, flush=True)
    if not args.no_profile:
        P.profile(True)
        P.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        j = step()
        print(f"[bench] step {k} misfit {j:.6e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = P.profile_read(reset=True) if not args.no_profile else {}
    dt_step = (t1 - t0) / args.steps
    gcell = g.ncells * g.nt * len(shots) / dt_step / 1e9
    pipe_gbs = alg_bytes_pipeline(g) * len(shots) / dt_step / 1e9
    device_gb = P.device_bytes / 1e9

    roof = None
    kernels = {}
    for k, (ms, n) in prof.items():
        if k in ALG_FLOATS and n:
            avg_ms = ms / n
            ach = alg_bytes_per_launch(k, g) / (avg_ms * 1e-3) / 1e9
            kernels[k] = {"avg_us": round(avg_ms * 1e3, 2), "launches": n, "share": ms / (dt_step * args.steps * 1e3),
                          "alg_GBps": round(ach, 1)}
    if kernels:
        dom = max(kernels, key=lambda k: prof[k][0])
        ach = kernels[dom]["alg_GBps"]
        traffic = hbm_traffic(dom, g, w.name)
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else round(traffic),
                "alg_bytes": round(alg_bytes_per_launch(dom, g))}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, args.cpu_steps)

    if g.memory_mode == 1:
        seg = g.checkpoint_every or int(np.ceil(np.sqrt(g.nt - 1)))
        res, nseg = P.resident_segments()
        recon = (f"sqrt(nt) checkpoints (segments of {seg} steps; the last {res} of {nseg} segments keep the "
                 f"forward pass's boundary ring, the others rebuild it by replay); "
                 f"the ring of all steps would need {ring_bytes(g) / 1e9:.0f} GB")
    else:
        recon = "ring reconstruction"
    out = {
        "metric": "fwd+adj+grad Gcell-updates/s & %HBM peak",
        "value": round(gcell, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {g.nz - 2 * g.npml}x{g.nx - 2 * g.npml}x{g.ny - 2 * g.npml} physical "
                               f"(z,x,y) + {g.npml} PML cells, padded {g.nz}x{g.nx}x{g.ny}, nt={g.nt}, "
                               f"{len(shots)} shot(s), {nrec} receivers, fwd+adj+grad, {recon}",
                   "device_GB": round(device_gb, 2),
                   "tiles": {k: f"{r}x{z}" for k, (r, z) in tiles.items()},
                   "pipeline_alg_GBps": round(pipe_gbs, 1),
                   "pipeline_hbm_frac": round(pipe_gbs / HBM_PEAK_GBS, 4),
                   "kernels": kernels},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    P.close()


if __name__ == "__main__":
    main()
