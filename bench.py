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
# Single-thread time of the reference compiled from its own sources over the
# oracle's time for the same work (2D 10k periodic mesh, k = 16, 20 HLL
# steps; oracle/calibrate.py on an Intel Xeon @ 2.70GHz).
REF_OVER_ORACLE = 1.426


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--cpu-sample", type=int, default=100_000, help="seeds of the one-thread CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30,
                    help="size of the streaming-copy probe (measured HBM bandwidth)")
    ap.add_argument("--train-steps", type=int, default=400, help="Adam steps of the k classifier")
    ap.add_argument("--no-boxes", action="store_true", help="skip the 256-box Kelvin-Helmholtz batch timing")
    ap.add_argument("--table", action="store_true",
                    help="measure the five workload configurations instead (one JSON line each)")
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_schedule(O, pts, rho, vel, E, k, gamma, cfl, hll_steps):
    """The oracle (FP64 C++ restatement, OpenMP) on the 1 rebuild + n HLL
    steps schedule: (stage seconds, final state, mesh)."""
    t0 = time.perf_counter()
    m = O.Mesh(pts, "reflective", k=k, restart=True)
    t_build = time.perf_counter() - t0
    st = {f"{key}": v for key, v in m.times().items()}
    st["build_s"] = t_build
    r, v, e = rho.copy(), vel.copy(), E.copy()
    t_dt = t_hll = 0.0
    dts = []
    for _ in range(hll_steps):
        t1 = time.perf_counter()
        dt = m.dt_cfl(gamma, cfl, r, v, e)
        t2 = time.perf_counter()
        m.hydro_step(gamma, dt, r, v, e, flux_mode=1)
        t3 = time.perf_counter()
        t_dt += t2 - t1
        t_hll += t3 - t2
        dts.append(dt)
    st["dt_s"] = t_dt / max(hll_steps, 1)
    st["hll_s"] = t_hll / max(hll_steps, 1)
    return st, (r, v, e, dts), m


def cpu_baseline(n, k, gamma, cfl, pts, rho, vel, E, sample):
    """The oracle on the host cores: the full workload's rebuild + one HLL
    step on all cores (one warm-up, median of 3, stage times), and a bounded
    sample on one core.  cells/s = cells / (rebuild + dt + one HLL step)."""
    from oracle import oracle as O
    cores = O.get_threads()
    oracle_schedule(O, pts[:sample], rho[:sample], vel[:sample], E[:sample], k, gamma, cfl, 1)  # warm-up
    runs = []
    for _ in range(3):
        st, _, _ = oracle_schedule(O, pts, rho, vel, E, k, gamma, cfl, 1)
        runs.append(st)
    med = {key: float(np.median([r[key] for r in runs])) for key in runs[0]}
    step_s = med["build_s"] + med["dt_s"] + med["hll_s"]
    O.set_threads(1)
    st1, _, _ = oracle_schedule(O, pts[:sample], rho[:sample], vel[:sample], E[:sample], k, gamma, cfl, 1)
    O.set_threads(cores)
    step1 = st1["build_s"] + st1["dt_s"] + st1["hll_s"]
    return {"value": round(n / step_s, 1), "unit": "cells/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"oracle (FP64 restatement of the reference path): the full workload, 3D {n} uniform seeds, "
                      f"reflective, k={k} + certificate restarts, one rebuild + dt + one HLL step, all {cores} "
                      f"threads, median of 3 after a warm-up ({step_s:.2f} s); one thread on the first {sample} seeds",
            "stages_s": {key: round(v, 4) for key, v in med.items()},
            "one_core": {"cells_per_s": round(sample / step1, 1), "cells": sample,
                         "stages_s": {key: round(v, 4) for key, v in st1.items()}},
            "reference_over_oracle_1core": REF_OVER_ORACLE}


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
            tune, _ = B.autotune(reps=1)  # the configuration the full run times
            ms, gpu = run_batch(ctx, B, a.warmup, a.steps)
            out[name] = {"ms_per_step": round(ms, 3), "gpu_ms_per_step": round(gpu, 3), "tuning": tune}
            B.free()
    out["measured_copy_GBs"] = round(ctx.stream_copy_probe(a.probe_bytes, 10), 1)
    print(json.dumps(out))
    if km is not None:
        km.free()
    ctx.close()


def clip_valu(launches, avg_launch_ms):
    """VALU issue share of the first clip pass: its measured VALU
    instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU, committed in
    profiles/) x 4 cycles per wave instruction over the SIMD-cycles of one
    launch (1024 SIMDs x 2.4 GHz x the launch time measured here); None
    without a summary."""
    path = os.path.join(ROOT, "profiles", "k_voronoi_1m_valu.json")
    if not os.path.exists(path) or launches <= 0:
        return None
    with open(path) as f:
        d = json.load(f)
    per_launch = d["valu_insts_per_step"] / launches
    simd_cycles = 1024 * 2.4e9 * avg_launch_ms * 1e-3
    return {"insts_per_launch": round(per_launch, 1), "issue_share": round(4.0 * per_launch / simd_cycles, 4),
            "insts_per_seed": round(d["valu_insts_per_step"] / d["seeds_per_step"], 1),
            "source": "profiles/k_voronoi_1m_valu.json"}


def stage_dict(st):
    return {k: round(v, 3) for k, v in st.items() if k.endswith("_ms")}


def config_c1(ctx, api, gamma, cfl):
    """Config C1: 2D, 10,000 uniform generators, reflective box, Sod, k = 16,
    20 moving steps at CFL 0.3: GPU step time, the oracle's on the host (all
    threads and one), cell by cell difference after the 20 steps."""
    from oracle import oracle as O
    from proteusgpu_amd.generators import sod_state, uniform_points
    phase("config C1 (2D 10k Sod, 20 moving steps)")
    p1 = uniform_points(10000, 2, 1001)
    r1, v1, e1 = sod_state(p1, gamma)
    kw1 = dict(boundary="reflective", k=16, gamma=gamma, cfl=cfl, flux="physical", move_mesh="fluid")
    g1 = api.Simulation(ctx, p1, r1, v1, e1, **kw1)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(20):
        g1.step()
    ctx.sync()
    g1_ms = (time.perf_counter() - t0) * 1e3 / 20
    oms = {}
    cores = O.get_threads()
    for threads in (cores, 1):
        O.set_threads(threads)
        o1 = O.Sim(p1, r1, v1, e1, boundary="reflective", k=16, gamma=gamma, cfl=cfl, flux="physical",
                   move="fluid")
        t0 = time.perf_counter()
        for _ in range(20):
            o1.step()
        oms[threads] = (time.perf_counter() - t0) * 1e3 / 20
        O.set_threads(cores)
    xg, rg, vg, eg = g1.get()
    xo, ro, vo, eo = o1.get()
    out = {"gpu_ms_per_step": round(g1_ms, 3), "gpu_cells_per_s": round(1e4 / (g1_ms * 1e-3), 1),
           "oracle_ms_per_step": {str(t): round(v, 2) for t, v in oms.items()},
           "oracle_cells_per_s": {str(t): round(1e4 / (v * 1e-3), 1) for t, v in oms.items()},
           "state_rel_err_vs_oracle": {"rho": float(np.max(np.abs(rg - ro) / ro)),
                                       "E": float(np.max(np.abs(eg - eo) / eo)),
                                       "x": float(np.max(np.abs(xg - xo)))},
           "topology_bit_exact": bool(np.array_equal(g1.mesh_copy_out()["neighbor_cell"],
                                                     o1.mesh().get("neighbor_cell")))}
    g1.free()
    return out


def oracle_steps(O, mk, steps, threads):
    """Mean seconds per oracle step of the simulation mk() makes (its
    creation, the first rebuild, is timed apart)."""
    cores = O.get_threads()
    O.set_threads(threads)
    t0 = time.perf_counter()
    sim = mk()
    t_make = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    t_step = (time.perf_counter() - t0) / steps
    O.set_threads(cores)
    return t_make, t_step


def table(a):
    """The workload configurations, one JSON line each: GPU throughput of the
    moving-mesh step (rebuild included), the oracle on this host's cores on
    a bounded sample (named in each line), and the parity figures that fit
    the oracle's time."""
    from oracle import oracle as O
    from proteusgpu_amd import api
    from proteusgpu_amd.generators import (batch_points, box_sizes, gaussian_mixture, jittered_lattice,
                                           kelvin_helmholtz_state, sod_state, uniform_points)
    ctx = api.Context(0)
    cores = O.get_threads()
    gamma, cfl = 1.4, 0.3
    emit = lambda d: print(json.dumps(d), flush=True)  # noqa: E731

    c1 = config_c1(ctx, api, gamma, cfl)
    emit({"config": "C1", "workload": "2D 10000 uniform seeds, reflective, Sod, k=16, 20 moving steps", **c1,
          "cpu_cores": cores, "cpu": cpu_model()})

    # C2: 1M 3D Sod, fixed k = 32: one rebuild + 10 static HLL steps
    phase("C2: 1M 3D Sod, 1 rebuild + 10 HLL")
    pts = uniform_points(1_000_000, 3, 42)
    rho, vel, E = sod_state(pts, gamma)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=32, gamma=gamma, cfl=cfl,
                         flux="physical", move_mesh=True)
    sched = [schedule_ms(ctx, sim) for _ in range(3)]
    sim.free()
    st, _, _ = oracle_schedule(O, pts, rho, vel, E, 32, gamma, cfl, 10)
    emit({"config": "C2", "workload": "3D 1000000 uniform seeds, reflective, Sod, k=32 + restarts, "
                                      "one rebuild + 10 static HLL steps",
          "gpu_ms": round(float(np.median(sched)), 3),
          "oracle_s": round(st["build_s"] + 10 * (st["dt_s"] + st["hll_s"]), 3), "cpu_cores": cores,
          "oracle_stages_s": {k: round(v, 4) for k, v in st.items()}})

    # C3: 4M clustered, periodic, learned k vs fixed k: mesh builds (a
    # simulation refuses meshes with uncertified cells, and cells beside the
    # densest clusters stay uncertified at k_max = 1024)
    phase("C3: 4M clustered, learned vs fixed k")
    n3 = 4_000_000
    p3 = gaussian_mixture(n3, 64, 0.002, 0.05, 0.2, 3)
    km = api.KModel(ctx, seed=9)
    sub = api.Mesh(ctx, p3[::10], boundary="periodic", k=32)
    km.train_mesh(sub, steps=300, batch=4096, lr=3e-3)
    sub.free()
    res, fails = {}, {}
    for name in ("fixed", "learned"):
        ts = []
        for _ in range(3):
            ctx.sync()
            t0 = time.perf_counter()
            m3 = api.Mesh(ctx, p3, boundary="periodic", k=32, k_max=1024, kmodel=km if name == "learned" else None)
            ctx.sync()
            ts.append((time.perf_counter() - t0) * 1e3)
            info = {"n_failed": m3.n_failed, "n_seeds": m3.n_seeds}
            m3.free()
        res[name] = float(np.median(ts))
        fails[name] = info["n_failed"]
    km.free()
    sub_n = 400_000
    ps = gaussian_mixture(sub_n, 64, 0.002, 0.05, 0.2, 3)
    t0 = time.perf_counter()
    om = O.Mesh(ps, "periodic", buff=0.04, k=32, restart=True, k_max=1024)
    o_s = time.perf_counter() - t0
    emit({"config": "C3", "workload": f"3D {n3} points (64 Gaussians, sigma in [0.002, 0.05], 20% uniform), "
                                      "periodic, k=32 + restarts to 1024, one mesh build",
          "gpu_ms_per_build": {k: round(v, 3) for k, v in res.items()},
          "gpu_cells_per_s": {k: round(n3 / (v * 1e-3), 1) for k, v in res.items()},
          "gpu_note": "host-driven mesh build from host points (copy to the device included), median of 3",
          "learned_speedup": round(res["fixed"] / res["learned"], 4),
          "uncertified": {k: int(v) for k, v in fails.items()}, "seeds": int(info["n_seeds"]),
          "oracle_build_s": round(o_s, 3), "oracle_cells_per_s": round(sub_n / o_s, 1),
          "oracle_sample": f"one mesh build of {sub_n} points of the same mixture, {cores} threads",
          "oracle_uncertified": int(om.n_failed)})

    # C4: 256 Kelvin-Helmholtz boxes, batched
    phase("C4: 256 KH boxes, batched")
    bs = box_sizes(256, 1000, 65536, 11)
    bp = batch_points(bs, 2, 500)
    br, bv, bE = kelvin_helmholtz_state(bp)
    kw4 = dict(boundary="periodic", buff=0.1, k=16, gamma=5.0 / 3.0, cfl=0.3, flux="physical",
               move_mesh="regularized")
    kh = api.Batch(ctx, bs, bp, br, bv, bE, **kw4)
    kh.autotune(reps=1)
    ms4, gpu4 = run_batch(ctx, kh, 3, 20)
    kh.free()
    off = np.concatenate([[0], np.cumsum(bs)])
    pick = np.random.default_rng(5).choice(len(bs), 16, replace=False)
    o_cells = int(bs[pick].sum())
    o_t = 0.0
    for b in pick:
        sl = slice(off[b], off[b + 1])
        _, t_step = oracle_steps(O, lambda: O.Sim(bp[sl], br[sl], bv[sl], bE[sl], boundary="periodic", buff=0.1,
                                                  k=16, gamma=5.0 / 3.0, cfl=0.3, move="regularized"), 3, cores)
        o_t += t_step
    emit({"config": "C4", "workload": f"256 periodic 2D boxes ({int(bs.sum())} cells, sizes log-uniform in "
                                      "[1000, 65536]), Kelvin-Helmholtz, regularised moving mesh, k=16",
          "gpu_ms_per_step": round(ms4, 3), "gpu_cells_per_s": round(float(bs.sum()) / (ms4 * 1e-3), 1),
          "oracle_cells_per_s": round(o_cells / o_t, 1),
          "oracle_sample": f"16 of the boxes ({o_cells} cells) one after another, 3 steps each, {cores} threads"})

    # C5: 8M Sedov, learned k, batched step
    phase("C5: 8M Sedov")
    g5, m5 = 5.0 / 3.0, 200
    p5 = jittered_lattice(m5, 3, 23)
    n5 = len(p5)
    d2 = np.sum((p5 - 0.5) ** 2, axis=1)
    hot = np.argsort(d2, kind="stable")[:32]
    r5 = np.ones(n5)
    v5 = np.zeros((n5, 3))
    e5 = np.full(n5, 1e-5 / (g5 - 1.0))
    e5[hot] += (1.0 / 32.0) * m5 ** 3
    kw5 = dict(boundary="reflective", k=32, gamma=g5, cfl=0.3, flux="physical", move_mesh="regularized")
    B5 = api.Batch(ctx, np.array([n5]), p5, r5, v5, e5, **kw5)
    ms5, _ = run_batch(ctx, B5, 2, 10)
    B5.free()
    m_s = 100
    ps5 = jittered_lattice(m_s, 3, 23)
    ns5 = len(ps5)
    rs5, vs5 = np.ones(ns5), np.zeros((ns5, 3))
    es5 = np.full(ns5, 1e-5 / (g5 - 1.0))
    t_make, t_step = oracle_steps(O, lambda: O.Sim(ps5, rs5, vs5, es5, boundary="reflective", k=32, gamma=g5,
                                                   cfl=0.3, move="regularized"), 1, cores)
    emit({"config": "C5", "workload": f"3D {n5} jittered-lattice seeds, reflective, Sedov (32 hot cells), "
                                      "regularised moving mesh, batched step, fixed k=32 + restarts",
          "gpu_ms_per_step": round(ms5, 3), "gpu_cells_per_s": round(n5 / (ms5 * 1e-3), 1),
          "oracle_cells_per_s": round(ns5 / t_step, 1),
          "oracle_sample": f"100^3 jittered lattice, one moving step, {cores} threads"})
    ctx.close()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if a.table:
        return table(a)
    from proteusgpu_amd import api
    from proteusgpu_amd.generators import sod_state, uniform_points

    gamma, cfl = 1.4, 0.3
    pts = uniform_points(a.n, 3, 42)
    rho, vel, E = sod_state(pts, gamma)
    ctx = api.Context(0)
    kw = dict(boundary="reflective", k=a.k, gamma=gamma, cfl=cfl, flux="physical", move_mesh=True)

    if a.only:
        return profile_variant(a, ctx, api, pts, rho, vel, E, kw)

    # 1. the plain method: fixed k + certificate restarts (resuming the saved
    #    cell), host-driven step, at the configured k and at 48 and 64
    fixed = {}
    for kk in sorted({a.k, 48, 64}):
        phase(f"fixed k={kk}, host-driven step")
        sim = api.Simulation(ctx, pts, rho, vel, E, **dict(kw, k=kk))
        ms, st = run_steps(ctx, sim, a.warmup, a.steps)
        fixed[kk] = {"ms_per_step": round(ms, 3), "cells_per_s": round(a.n / (ms * 1e-3), 1),
                     "stages_ms": stage_dict(st), "restarted_seeds_per_step": st["n_restarted"],
                     "row_entries_per_step": st["row_entries"]}
        if kk == a.k:
            plain = sim
            fixed[kk]["schedule_1rebuild_10hll_ms"] = round(schedule_ms(ctx, plain), 3)
        else:
            sim.free()
    ms_plain = fixed[a.k]["ms_per_step"]
    # the same with restarts that clip again from the first neighbour (the
    # restart of the reference's semantics)
    phase("fixed k, host-driven step, restarts from scratch")
    api.set_clip_resume(False)
    scratch = api.Simulation(ctx, pts, rho, vel, E, **kw)
    ms_scratch, st_scratch = run_steps(ctx, scratch, a.warmup, a.steps)
    api.set_clip_resume(True)
    state_scratch = scratch.get()
    scratch.free()

    # 2. learned k: the classifier is trained on the GPU, in this run, on the
    #    first mesh of a different point set (seed 43) of the same workload
    #    (certified k of every seed as labels), then used on this one
    phase("training the k classifier")
    km = api.KModel(ctx, seed=1)
    tp = uniform_points(a.n, 3, 43)
    tr, tv, tE = sod_state(tp, gamma)
    trainer = api.Simulation(ctx, tp, tr, tv, tE, **kw)
    ctx.sync()
    t0 = time.perf_counter()
    loss = km.train_mesh(trainer, steps=a.train_steps, batch=4096, lr=3e-3)
    ctx.sync()
    train_ms = (time.perf_counter() - t0) * 1e3
    trainer.free()
    phase("learned k, host-driven step")
    lsim = api.Simulation(ctx, pts, rho, vel, E, **kw)
    lsim.set_kmodel(km)
    ms_lsim, st_lsim = run_steps(ctx, lsim, a.warmup, a.steps)
    info_l = lsim.mesh_info()
    lsim.free()

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

    # roofline of the dominant kernel: the first clip pass (k_voronoi<3,16,S>,
    # one launch per k bucket) of the learned step
    n_seeds = info_l["n_seeds"]
    rows_first = float(sum(int(c) * b for c, b in zip(info_l["n_class"], (16, 24, 32, 48, 64, 96))))
    launches = int(sum(1 for c in info_l["n_class"] if c > 0))
    clip_bytes = 4.0 * rows_first + n_seeds * CLIP_FIXED_BYTES_PER_CELL
    clip_ms = st_lsim["clip_ms"]
    achieved = clip_bytes / (clip_ms * 1e-3) / 1e9
    copy_gbs = ctx.stream_copy_probe(a.probe_bytes, 10)
    avg_launch = clip_ms / max(launches, 1)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": clip_traffic(launches),
            "kernel": "k_voronoi<3,16,S> (first clip pass of the learned step, one launch per k bucket)",
            "launches_per_step": launches, "bytes_per_launch": round(clip_bytes / max(launches, 1), 1),
            "avg_launch_ms": round(avg_launch, 4), "bytes_per_cell": round(clip_bytes / n_seeds, 1),
            "measured_copy_GBs": round(copy_gbs, 1), "frac_of_measured_copy": round(achieved / copy_gbs, 5),
            "valu": clip_valu(launches, avg_launch)}
    best = ms_blearn
    path_gbs = PATH_BYTES_PER_CELL * a.n / (best * 1e-3) / 1e9

    # 4. the FP64 oracle on the host: CPU baseline, and the state error of the
    #    "1 rebuild + 10 HLL steps" schedule (static mesh after one build from
    #    the initial state) against the GPU's
    cpu, err = None, None
    if not a.no_cpu:
        phase("cpu baseline and state error vs the FP64 oracle")
        from oracle import oracle as O
        cpu = cpu_baseline(a.n, a.k, gamma, cfl, pts, rho, vel, E, min(a.cpu_sample, a.n))
        m = api.Mesh(ctx, pts, boundary="reflective", k=a.k, restart=True)
        hs = api.HydroState(ctx, m, rho, vel, E)
        mo = O.Mesh(pts, "reflective", k=a.k, restart=True)
        same_topo = bool(np.array_equal(m.copy_out()["neighbor_cell"], mo.get("neighbor_cell")))
        dts_g = []
        r_o, v_o, e_o = rho.copy(), vel.copy(), E.copy()
        for _ in range(10):
            dg = hs.dt_cfl(gamma, cfl)
            do = mo.dt_cfl(gamma, cfl, r_o, v_o, e_o)
            dts_g.append(abs(dg - do) / do)
            hs.step(gamma, do, "physical")
            mo.hydro_step(gamma, do, r_o, v_o, e_o, flux_mode=1)
        r_g, v_g, e_g = hs.copy_out()
        err = {"schedule": "one rebuild (fixed k + restarts) then 10 static HLL steps from the initial state, "
                           "both at the oracle's dt", "topology_bit_exact": same_topo,
               "rho": float(np.max(np.abs(r_g - r_o) / np.abs(r_o))),
               "E": float(np.max(np.abs(e_g - e_o) / np.abs(e_o))),
               "v": float(np.max(np.abs(v_g - v_o)) / np.max(np.abs(v_o))),
               "dt": float(max(dts_g))}
        hs.free()
        m.free()
    # 5. config C1 beside the headline (its GPU step, the oracle's, cell by
    #    cell difference)
    c1 = None if a.no_cpu else config_c1(ctx, api, gamma, cfl)
    tot = blearn.totals()[0]
    best_fixed_k = min(fixed, key=lambda kk: fixed[kk]["ms_per_step"])
    ms_best_fixed = fixed[best_fixed_k]["ms_per_step"]
    out = {
        "metric": "cells/s per Voronoi rebuild+HLL step, 1M 3D seeds; learned-k vs fixed-k speedup",
        "value": round(a.n / (best * 1e-3), 1), "unit": "cells/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(best, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D {a.n} uniform generators, reflective box, Sod along x (gamma 1.4), "
                               f"CFL 0.3, moving mesh, one full Voronoi rebuild per step; batched step (one "
                               f"hipGraph launch) with k per seed from the learned classifier (trained in this "
                               f"run on the first mesh of another point set) and certificate restarts; "
                               f"fixed-k steps at k = {', '.join(str(x) for x in sorted(fixed))} timed beside",
                   "n_cells": a.n, "k_fixed": a.k},
        "roofline": roof, "cpu_baseline": cpu, "state_rel_err_vs_oracle": err, "config_c1": c1,
        "learned_k": {
            "host_ms_per_step": round(ms_lsim, 3), "batched_ms_per_step": round(ms_blearn, 3),
            "speedup_host_vs_fixed_k": round(ms_plain / ms_lsim, 4),
            "speedup_batched_vs_fixed_k_batched": round(ms_bfix / ms_blearn, 4),
            "best_fixed_k": best_fixed_k,
            "learned_vs_best_fixed": round(ms_best_fixed / ms_lsim, 4),
            "speedup_vs_scratch_restarts_host": round(ms_scratch / ms_lsim, 4),
            "target_1_5x_vs_best_fixed_met": bool(ms_best_fixed / ms_lsim >= 1.5),
            "stages_ms": stage_dict(st_lsim), "restarted_seeds_per_step": st_lsim["n_restarted"],
            "row_entries_per_step": st_lsim["row_entries"],
            "seeds_per_bucket": [int(c) for c in info_l["n_class"]],
            "training": {"ms": round(train_ms, 2), "adam_steps": a.train_steps, "batch": 4096,
                         "final_loss": round(loss, 4), "points": "seed 43 (evaluated on seed 42)"}},
        "results_identical": same,
        "fixed_k_host": {str(kk): v for kk, v in fixed.items()},
        "fixed_k_scratch_restarts_host": {
            "ms_per_step": round(ms_scratch, 3), "cells_per_s": round(a.n / (ms_scratch * 1e-3), 1),
            "stages_ms": stage_dict(st_scratch), "restarted_seeds_per_step": st_scratch["n_restarted"]},
        "batched_fixed_k": {"ms_per_step": round(ms_bfix, 3), "gpu_ms_per_step": round(gpu_bfix, 3),
                            "tuning": tune_fix},
        "batched_learned_k": {"ms_per_step": round(ms_blearn, 3), "gpu_ms_per_step": round(gpu_blearn, 3),
                              "tuning": tune_learn},
        "batched_boxes_kh256": boxes,
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
