#!/usr/bin/env python3
"""Benchmark of the moving-mesh hot path on one MI355X.

Workload (the metric's configuration): 3D, 1,000,000 generators uniform in
the unit cube, reflective box, Sod along x (rho,P = 1,1 | 0.125,0.1, v=0,
gamma=1.4), CFL 0.3, generators moving with the fluid.  A step is one
moving-mesh step: primitives + CFL dt + HLL update + drift + one full Voronoi
rebuild (grid, KNN, clip, restarts, face tables), all resident in HBM.

Two runs of the same steps from the same initial state: fixed k=32 with
certificate restarts (the plain method), and learned k (a classifier trained
on the GPU, in this run, on the first mesh, choosing each seed's first k;
the certificate restarts what it underestimates).  value = hydro cells /
step time of the learned run; the fixed-k run, the speedup and a check that
both runs end in bit-identical states are reported beside it.

Prints one JSON line.  Extra fields: per-stage times, the "1 rebuild + 10 HLL
steps" schedule, the roofline of the dominant kernel (the first clip pass)
and a CPU baseline (the FP64 oracle on a bounded sample of the same
workload, timed on this host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# Algorithmic HBM bytes per cell of the clip kernel (3D, Poisson-Voronoi
# F~15.54 faces, F/2 owned): read the k-row (4k) and the seed (12), write the
# cell record (volume 8, centroid 12, range 8) and the owned face records
# (ids + area 12, normal 12, centroid 12): 4k + 12 + 28 + 7.77 * 36.
CLIP_FIXED_BYTES_PER_CELL = 12 + 28 + 7.77 * 36
# Whole step (grid 42.6 + KNN 140 + clip 448 + HLL 500 + drift 60) at k = 32.
PATH_BYTES_PER_CELL = 1190.0
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30,
                    help="size of the streaming-copy probe (measured HBM bandwidth)")
    ap.add_argument("--train-steps", type=int, default=400, help="Adam steps of the k classifier")
    ap.add_argument("--no-boxes", action="store_true", help="skip the 256-box Kelvin-Helmholtz batch timing")
    ap.add_argument("--only", choices=["plain", "learned", "batched"], default=None,
                    help="run one variant only (for profiling): the fixed-k host step, the learned-k host "
                         "step, or the batched fixed- and learned-k steps")
    return ap.parse_args()


def clip_traffic(launches):
    """HBM bytes per launch of the first clip pass at this workload: the
    FETCH_SIZE + WRITE_SIZE of its launches in one step (rocprofv3 --pmc,
    scaled by a copy kernel of known size), from the committed summary in
    profiles/, divided by the launches per step; None without a summary."""
    path = os.path.join(ROOT, "profiles", "k_voronoi_1m_traffic.json")
    if not os.path.exists(path) or launches <= 0:
        return None
    with open(path) as f:
        d = json.load(f)
    per_step = d.get("learned_first_pass_bytes_per_step")
    return None if per_step is None else round(per_step / launches, 1)


def cpu_baseline(n_sample, k):
    """The oracle (FP64 C++ restatement, OpenMP) on a bounded sample: one mesh
    build with restarts + one CFL dt + one HLL step, same workload type."""
    from oracle import oracle as O
    from proteusgpu_amd.generators import sod_state, uniform_points

    pts = uniform_points(n_sample, 3, 2024)
    rho, vel, E = sod_state(pts, 1.4)
    t0 = time.perf_counter()
    m = O.Mesh(pts, "reflective", k=k, restart=True)
    dt = m.dt_cfl(1.4, 0.3, rho, vel, E)
    m.hydro_step(1.4, dt, rho, vel, E, flux_mode=1)
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n_sample / el, "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"oracle: 3D {n_sample} uniform seeds, reflective, k={k}+restarts, 1 rebuild + "
                      f"1 HLL step ({el:.2f} s)"}


def phase(name):
    """Progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"[bench] {name}", file=sys.stderr, flush=True)


def run_steps(ctx, sim, warmup, steps):
    for _ in range(warmup):
        sim.step()
    ctx.sync()
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        stages.append(sim.times())
    ctx.sync()
    el = time.perf_counter() - t0
    mean = {k: float(np.mean([s[k] for s in stages])) for k in stages[0] if k.endswith("_ms")}
    mean["n_restarted"] = float(np.mean([s["n_restarted"] for s in stages]))
    mean["row_entries"] = float(np.mean([s["row_entries"] for s in stages]))
    return el * 1e3 / steps, mean


def schedule_ms(ctx, sim):
    """One full rebuild plus 10 static-mesh HLL steps."""
    ctx.sync()
    t1 = time.perf_counter()
    sim.step()
    sim.hydro_steps(10)
    ctx.sync()
    return (time.perf_counter() - t1) * 1e3


def run_batch(ctx, B, warmup, steps):
    for _ in range(warmup):
        B.step()
    ctx.sync()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu_ms.append(B.step())
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / steps, float(np.mean(gpu_ms))


def profile_variant(a, ctx, api, pts, rho, vel, E, kw):
    """One variant alone, so that a profiler sees only its kernels."""
    out = {"variant": a.only, "n": a.n, "steps": a.steps}
    km = None
    if a.only in ("learned", "batched"):
        km = api.KModel(ctx, seed=1)
        t = api.Simulation(ctx, pts, rho, vel, E, **kw)
        out["train_loss"] = km.train_mesh(t, steps=a.train_steps, batch=4096, lr=3e-3)
        t.free()
    if a.only in ("plain", "learned"):
        sim = api.Simulation(ctx, pts, rho, vel, E, **kw)
        if km is not None:
            sim.set_kmodel(km)
        ms, st = run_steps(ctx, sim, a.warmup, a.steps)
        out.update(ms_per_step=round(ms, 3), stages_ms={k: round(v, 3) for k, v in st.items() if k.endswith("_ms")},
                   mesh=sim.mesh_info())
        sim.free()
    else:
        for name, model in (("batched_fixed_k", None), ("batched_learned_k", km)):
            B = api.Batch(ctx, np.array([a.n]), pts, rho, vel, E, **kw)
            if model is not None:
                B.set_kmodel(model)
            ms, gpu = run_batch(ctx, B, a.warmup, a.steps)
            out[name] = {"ms_per_step": round(ms, 3), "gpu_ms_per_step": round(gpu, 3)}
            B.free()
    out["measured_copy_GBs"] = round(ctx.stream_copy_probe(a.probe_bytes, 10), 1)
    print(json.dumps(out))
    if km is not None:
        km.free()
    ctx.close()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    from proteusgpu_amd import api
    from proteusgpu_amd.generators import sod_state, uniform_points

    gamma, cfl = 1.4, 0.3
    pts = uniform_points(a.n, 3, 42)
    rho, vel, E = sod_state(pts, gamma)
    ctx = api.Context(0)
    kw = dict(boundary="reflective", k=a.k, gamma=gamma, cfl=cfl, flux="physical", move_mesh=True)

    if a.only:
        return profile_variant(a, ctx, api, pts, rho, vel, E, kw)

    # 1. the plain method: fixed k=32 + certificate restarts, host-driven step
    phase("fixed k, host-driven step")
    plain = api.Simulation(ctx, pts, rho, vel, E, **kw)
    ms_plain, st_plain = run_steps(ctx, plain, a.warmup, a.steps)
    sched_plain = schedule_ms(ctx, plain)
    # the same with restarts that clip again from the first neighbour
    phase("fixed k, host-driven step, restarts from scratch")
    api.set_clip_resume(False)
    scratch = api.Simulation(ctx, pts, rho, vel, E, **kw)
    ms_scratch, st_scratch = run_steps(ctx, scratch, a.warmup, a.steps)
    api.set_clip_resume(True)
    state_scratch = scratch.get()
    scratch.free()

    # 2. learned k: the classifier is trained on the GPU, in this run, on the
    #    first mesh of the workload (certified k of every seed as labels)
    phase("training the k classifier")
    km = api.KModel(ctx, seed=1)
    trainer = api.Simulation(ctx, pts, rho, vel, E, **kw)
    ctx.sync()
    t0 = time.perf_counter()
    loss = km.train_mesh(trainer, steps=a.train_steps, batch=4096, lr=3e-3)
    ctx.sync()
    train_ms = (time.perf_counter() - t0) * 1e3
    # learned k in the host-driven step (stage times)
    phase("learned k, host-driven step")
    trainer.set_kmodel(km)
    ms_lsim, st_lsim = run_steps(ctx, trainer, a.warmup, a.steps)
    info_l = trainer.mesh_info()
    trainer.free()

    # 3. the batched step: one box, the whole step in one hipGraph launch,
    #    fixed k and learned k
    sizes = np.array([a.n])
    phase("batched step, fixed k")
    bfix = api.Batch(ctx, sizes, pts, rho, vel, E, **kw)
    tune_fix, _ = bfix.autotune(reps=1)
    ms_bfix, gpu_bfix = run_batch(ctx, bfix, a.warmup, a.steps)
    phase("batched step, learned k")
    blearn = api.Batch(ctx, sizes, pts, rho, vel, E, **kw)
    blearn.set_kmodel(km)
    tune_learn, _ = blearn.autotune(reps=1)
    ms_blearn, gpu_blearn = run_batch(ctx, blearn, a.warmup, a.steps)

    # 256 uneven 2D Kelvin-Helmholtz boxes in one captured step (batched
    # boxes, reported beside the headline; not part of `value`)
    boxes = None
    if not a.no_boxes:
        phase("256 Kelvin-Helmholtz boxes, batched")
        from proteusgpu_amd.generators import batch_points, box_sizes, kelvin_helmholtz_state
        bs = box_sizes(256, 1000, 65536, 11)
        bp = batch_points(bs, 2, 500)
        br, bv, bE = kelvin_helmholtz_state(bp)
        kh = api.Batch(ctx, bs, bp, br, bv, bE, boundary="periodic", buff=0.1, k=16, gamma=5.0 / 3.0, cfl=0.3,
                       flux="physical", move_mesh="regularized")
        tune_kh, _ = kh.autotune(reps=1)
        ms_kh, gpu_kh = run_batch(ctx, kh, a.warmup, a.steps)
        boxes = {"workload": "256 periodic 2D boxes, sizes log-uniform in [1000, 65536], Kelvin-Helmholtz, "
                             "regularised moving mesh, k = 16", "cells": int(bs.sum()),
                 "ms_per_step": round(ms_kh, 3), "gpu_ms_per_step": round(gpu_kh, 3),
                 "cells_per_s": round(float(bs.sum()) / (ms_kh * 1e-3), 1), "tuning": tune_kh}
        kh.free()

    # every run took the same steps from the same state: identical results
    # (the batched runs did warmup + steps; a fresh host-driven run is
    # brought to the same step count)
    same = True
    phase("identity check")
    chk = api.Simulation(ctx, pts, rho, vel, E, **kw)
    for _ in range(a.warmup + a.steps):
        chk.step()
    ref = chk.get()
    for got in (bfix.get(), blearn.get(), state_scratch):
        same = same and all(np.array_equal(x, y) for x, y in zip(got, ref))
    chk.free()

    # roofline of the dominant kernel: the first clip pass (k_voronoi<3,16,4>,
    # one launch per k bucket) of the learned step
    n_seeds = info_l["n_seeds"]
    rows_first = float(sum(int(c) * b for c, b in zip(info_l["n_class"], (16, 24, 32, 48, 64, 96))))
    launches = int(sum(1 for c in info_l["n_class"] if c > 0))
    clip_bytes = 4.0 * rows_first + n_seeds * CLIP_FIXED_BYTES_PER_CELL
    clip_ms = st_lsim["clip_ms"]
    achieved = clip_bytes / (clip_ms * 1e-3) / 1e9
    copy_gbs = ctx.stream_copy_probe(a.probe_bytes, 10)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": clip_traffic(launches),
            "kernel": "k_voronoi<3,16,4> (first clip pass, one launch per k bucket)",
            "launches_per_step": launches, "bytes_per_launch": round(clip_bytes / max(launches, 1), 1),
            "avg_launch_ms": round(clip_ms / max(launches, 1), 4),
            "bytes_per_cell": round(clip_bytes / n_seeds, 1),
            "measured_copy_GBs": round(copy_gbs, 1), "frac_of_measured_copy": round(achieved / copy_gbs, 5)}
    best = ms_blearn
    path_gbs = PATH_BYTES_PER_CELL * a.n / (best * 1e-3) / 1e9
    cpu = None
    if not a.no_cpu:
        phase("cpu baseline")
        cpu = cpu_baseline(a.cpu_sample, a.k)
    tot = blearn.totals()[0]
    out = {
        "metric": "cells/s per Voronoi rebuild+HLL step, 1M 3D seeds; learned-k vs fixed-k speedup",
        "value": round(a.n / (best * 1e-3), 1), "unit": "cells/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(best, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D {a.n} uniform generators, reflective box, Sod along x (gamma 1.4), "
                               f"CFL 0.3, moving mesh, one full Voronoi rebuild per step; batched step (one "
                               f"hipGraph launch) with k per seed from the learned classifier (trained in this "
                               f"run on the first mesh) and certificate restarts; the plain fixed k={a.k} step "
                               f"timed beside",
                   "n_cells": a.n, "k_fixed": a.k},
        "roofline": roof, "cpu_baseline": cpu,
        "learned_k_speedup": round(ms_plain / best, 4),
        "learned_k_speedup_vs_scratch_restarts": round(ms_scratch / best, 4),
        "results_identical": same,
        "plain_fixed_k": {"ms_per_step": round(ms_plain, 3), "cells_per_s": round(a.n / (ms_plain * 1e-3), 1),
                          "stages_ms": {k: round(v, 3) for k, v in st_plain.items() if k.endswith("_ms")},
                          "restarted_seeds_per_step": st_plain["n_restarted"],
                          "row_entries_per_step": st_plain["row_entries"],
                          "schedule_1rebuild_10hll_ms": round(sched_plain, 3)},
        "plain_fixed_k_scratch_restarts": {
            "ms_per_step": round(ms_scratch, 3), "cells_per_s": round(a.n / (ms_scratch * 1e-3), 1),
            "stages_ms": {k: round(v, 3) for k, v in st_scratch.items() if k.endswith("_ms")},
            "restarted_seeds_per_step": st_scratch["n_restarted"]},
        "learned_k_host_step": {"ms_per_step": round(ms_lsim, 3),
                                "stages_ms": {k: round(v, 3) for k, v in st_lsim.items() if k.endswith("_ms")},
                                "restarted_seeds_per_step": st_lsim["n_restarted"],
                                "row_entries_per_step": st_lsim["row_entries"],
                                "seeds_per_bucket": [int(c) for c in info_l["n_class"]]},
        "batched_fixed_k": {"ms_per_step": round(ms_bfix, 3), "gpu_ms_per_step": round(gpu_bfix, 3),
                            "tuning": tune_fix},
        "batched_learned_k": {"ms_per_step": round(ms_blearn, 3), "gpu_ms_per_step": round(gpu_blearn, 3),
                              "tuning": tune_learn},
        "batched_boxes_kh256": boxes,
        "training": {"ms": round(train_ms, 2), "adam_steps": a.train_steps, "batch": 4096,
                     "final_loss": round(loss, 4)},
        "path_roofline": {"bytes_per_cell": PATH_BYTES_PER_CELL, "achieved_GBs": round(path_gbs, 2),
                          "frac": round(path_gbs / HBM_PEAK_GBS, 5)},
        "conserved_totals": [float(x) for x in tot],
    }
    print(json.dumps(out))
    for o in (plain, bfix, blearn, km):
        o.free()
    ctx.close()


if __name__ == "__main__":
    main()
