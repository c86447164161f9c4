#!/usr/bin/env python3
"""Calibration of the CPU baseline (test infrastructure): the reference
compiled from its own sources (oracle/_ref, `make -C oracle ref`, only where
the reference tree is mounted) timed beside the oracle's restatement of the
same work, so that the oracle's timings on the GPU host (where the reference
never goes) translate into the reference's.

Work at config-1 size: 2D, 10,000 uniform generators, the reference's
periodic mesh (buff 0.04, fixed k = 16, no restarts), then 20 CFL-limited
HLL steps on the static mesh (gamma 1.4, CFL 0.3, the reference's face
orientation) -- compute_periodic_mesh + dt_CFL + hydro_step, the only path
the reference has.  Both run single-threaded (the reference's Voronoi stage
is serial by construction; its build here has no OpenMP).  Median of 3
after one warm-up.

    python oracle/calibrate.py
"""
import json
import os
import platform
import statistics
import struct
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[0] = ROOT  # the package `oracle`, not this directory's oracle.py

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402
from proteusgpu_amd.generators import sod_state, uniform_points  # noqa: E402


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_reference(pts, rho, vel, E, steps):
    exe = os.path.join(HERE, "_ref", "ref_2d_k16")
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "in.bin")
        with open(inp, "wb") as f:
            f.write(struct.pack("<iq", pts.shape[1], pts.shape[0]))
            for a in (pts, rho, vel, E):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        t0 = time.perf_counter()
        subprocess.run([exe, "periodic", inp, d, "1.4", "0.3", str(steps), "0.04"], check=True,
                       capture_output=True)
        return time.perf_counter() - t0


def time_oracle(pts, rho, vel, E, steps):
    r, v, e = rho.copy(), vel.copy(), E.copy()
    t0 = time.perf_counter()
    m = O.Mesh(pts, "periodic", buff=0.04, k=16, restart=False)
    for _ in range(steps):
        dt = m.dt_cfl(1.4, 0.3, r, v, e)
        m.hydro_step(1.4, dt, r, v, e, flux_mode=0)
    return time.perf_counter() - t0


def main():
    if not os.path.exists(os.path.join(HERE, "_ref", "ref_2d_k16")):
        sys.exit("oracle/_ref is not built (make -C oracle ref, where the reference is mounted)")
    pts = uniform_points(10000, 2, 1001)
    rho, vel, E = sod_state(pts, 1.4)
    O.set_threads(1)
    res = {}
    for name, fn in (("reference", time_reference), ("oracle", time_oracle)):
        fn(pts, rho, vel, E, 20)
        res[name] = statistics.median(fn(pts, rho, vel, E, 20) for _ in range(3))
    out = {"work": "2D 10000 uniform seeds, periodic mesh k=16 (no restarts) + 20 HLL steps, 1 thread",
           "cpu": cpu_model(), "reference_s": round(res["reference"], 4), "oracle_s": round(res["oracle"], 4),
           "reference_over_oracle": round(res["reference"] / res["oracle"], 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
