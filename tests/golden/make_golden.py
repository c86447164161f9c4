#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference itself.

Runs the harness binaries of oracle/_ref (built from the reference's own
sources by `make -C oracle ref`, in a container where the reference tree is
mounted) on inputs from the project's seeded generator and stores their
outputs.  Only data is stored: inputs are regenerated from (generator, seed)
and pinned by a SHA-256 of their bytes.  Large arrays are stored in full where
they pin topology and as a leading slice plus a SHA-256 of the whole array
otherwise.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from proteusgpu_amd.generators import sod_state, uniform_points  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(binary, mode, pts, extra=(), hydro=None):
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "in.bin")
        with open(inp, "wb") as f:
            f.write(struct.pack("<iq", pts.shape[1], pts.shape[0]))
            f.write(pts.astype("<f8").tobytes())
            if hydro is not None:
                for a in hydro:
                    f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        r = subprocess.run([os.path.join(REF, binary), mode, inp, d, *map(str, extra)],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{binary} {mode} failed: {r.returncode}\n{r.stderr[-2000:]}")
        out = {}
        for line in open(os.path.join(d, "index.txt")):
            name, dt, _ = line.split()
            out[name] = np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)
        return out


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"{name}.npz  {os.path.getsize(path) / 1024:.0f} KiB")


def knn_case(name, dim, n, k, seed):
    pts = uniform_points(n, dim, seed)
    o = run(f"ref_{dim}d_k{k}", "knn", pts)
    kn = o["knearest"].reshape(n, k)
    save(name, dim=dim, n=n, k=k, seed=seed, pts_sha=sha(pts), N_grid=o["N_grid"][0],
         permutation=o["permutation"], sorted_sha=sha(o["sorted_points"]),
         knearest_head=kn[:3000], knearest_sha=sha(kn))


def mesh_case(name, dim, n, k, seed, head_cells):
    pts = uniform_points(n, dim, seed)
    o = run(f"ref_{dim}d_k{k}", "mesh", pts)
    fc, fp = o["face_counts"], o["face_ptr"]
    nf_head = int(fp[head_cells - 1] + fc[head_cells - 1]) if head_cells < n else len(o["neighbor_cell"])
    save(name, dim=dim, n=n, k=k, seed=seed, pts_sha=sha(pts), face_counts=fc.astype(np.int16),
         neighbor_cell=o["neighbor_cell"], volumes=o["volumes"],
         head_cells=head_cells, face_area_head=o["face_area"][:nf_head],
         face_normal_head=o["face_normal"].reshape(-1, 3)[:nf_head],
         face_area_sha=sha(o["face_area"]), face_normal_sha=sha(o["face_normal"]),
         seeds_sha=sha(o["seeds"]))


def periodic_case(name, dim, n, k, seed, buff, gamma, cfl, steps):
    pts = uniform_points(n, dim, seed)
    rho, vel, E = sod_state(pts, gamma)
    o = run(f"ref_{dim}d_k{k}", "periodic", pts, extra=(gamma, cfl, steps, buff), hydro=(rho, vel, E))
    nh = int(o["mesh_header"][2])
    fc, fp = o["face_counts"], o["face_ptr"]
    nf_h = int(fp[nh - 1] + fc[nh - 1])
    save(name, dim=dim, n=n, k=k, seed=seed, buff=buff, gamma=gamma, cfl=cfl, steps=steps,
         pts_sha=sha(pts), n_seeds=o["mesh_header"][0], ghost_ids=o["ghost_ids"],
         face_counts=fc.astype(np.int16), neighbor_cell_hydro=o["neighbor_cell"][:nf_h],
         neighbor_cell_sha=sha(o["neighbor_cell"]), volumes=o["volumes"],
         face_area_sha=sha(o["face_area"]), dts=o["dts"], rho=o["rho"], vel=o["vel"], E=o["E"])


def main():
    if not os.path.isdir(REF):
        sys.exit("oracle/_ref is not built (make -C oracle ref, where the reference is mounted)")
    knn_case("knn_2d_n2000_k16", 2, 2000, 16, 101)
    knn_case("knn_3d_n13000_k32", 3, 13000, 32, 102)
    mesh_case("mesh_2d_n2000_k32", 2, 2000, 32, 103, head_cells=2000)
    mesh_case("mesh_3d_n13000_k64", 3, 13000, 64, 104, head_cells=1500)
    periodic_case("periodic_2d_n10000_k48", 2, 10000, 48, 105, 0.04, 1.4, 0.3, 4)
    periodic_case("periodic_3d_n20000_k128", 3, 20000, 128, 106, 0.12, 1.4, 0.3, 3)


if __name__ == "__main__":
    main()
