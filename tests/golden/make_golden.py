#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference itself.

Runs the harness binaries of oracle/_ref (built from the reference's own
sources by `make -C oracle ref`, in a container where the reference tree is
mounted) on inputs from the project's seeded generator and stores their
outputs.  Only data is stored: inputs are regenerated from (generator, seed)
and pinned by a SHA-256 of their bytes.  Large arrays are stored in full where
they pin topology and as a leading slice plus a SHA-256 of the whole array
otherwise.

    python tests/golden/make_golden.py            (every fixture)
    python tests/golden/make_golden.py files      (only the file-layout fixture)
"""
import hashlib
import json
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


H5DUMP = "/opt/conda/bin/h5dump"
MAKE_IC = os.path.join(ROOT, "proteusgpu_amd", "host", "make_ic")


def h5_header(path):
    """h5dump -H of an HDF5 file (groups, datasets, dtypes, shapes and
    attributes), without the line naming the file."""
    r = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True)
    return "\n".join(r.stdout.splitlines()[1:])


def h5_attrs(path):
    """h5dump -A of the /header group: its attributes with their values."""
    r = subprocess.run([H5DUMP, "-A", "-g", "/header", path], capture_output=True, text=True, check=True)
    return "\n".join(r.stdout.splitlines()[1:])


def h5_sha(path, name):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "x.bin")
        subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
        return hashlib.sha256(open(out, "rb").read()).hexdigest()


def files_case(name, periodic_golden):
    """The reference's own main loop (IC reader, periodic mesh with its KNN
    dump, CFL loop, snapshot writer) on an IC written by the project's
    make_ic, run to the time the golden periodic case reaches in its steps.
    Stores the structure of both files (h5dump -H) and SHA-256 of the
    datasets a port must reproduce bit for bit."""
    g = np.load(os.path.join(HERE, periodic_golden + ".npz"))
    n, seed, gamma, cfl, buff = int(g["n"]), int(g["seed"]), float(g["gamma"]), float(g["cfl"]), float(g["buff"])
    t_end = 0.0
    for dt in g["dts"]:
        t_end += float(dt)
    with tempfile.TemporaryDirectory() as d:
        ic = os.path.join(d, "IC.hdf5")
        subprocess.run([MAKE_IC, ic, "3", str(n), str(seed), "sod", repr(gamma)], check=True, capture_output=True)
        outdir = os.path.join(d, "out") + "/"
        par = os.path.join(d, "param.txt")
        with open(par, "w") as f:
            f.write(f"ic_file = {ic}\noutput_directory = {outdir}\noutput_mesh_file = mesh.hdf5\n"
                    f"output_knn_file = knn.hdf5\n")
        r = subprocess.run([os.path.join(REF, "ref_files_3d_k128"), "files", par, repr(gamma), repr(cfl), repr(t_end),
                            repr(buff)], capture_output=True, text=True, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"ref_files failed: {r.returncode}\n{r.stderr[-2000:]}")
        mesh, knn = os.path.join(outdir, "mesh.hdf5"), os.path.join(outdir, "knn.hdf5")
        # the reference's snapshot state is the golden periodic state
        for key, ds in (("rho", "/hydro/rho"), ("E", "/hydro/Energy"), ("vel", "/hydro/vel")):
            assert h5_sha(mesh, ds) == sha(g[key]), key
        out = {"n": n, "seed": seed, "gamma": gamma, "cfl": cfl, "buff": buff, "t_end": t_end, "k": 128,
               "mesh_header": h5_header(mesh), "knn_header": h5_header(knn),
               "mesh_attrs": h5_attrs(mesh), "knn_attrs": h5_attrs(knn),
               "mesh_sha": {ds: h5_sha(mesh, ds) for ds in ("/cells/cell_ids", "/cells/face_counts",
                                                              "/cells/faces/neighbor_cell")},
               "knn_sha": {ds: h5_sha(knn, ds) for ds in ("/knn/points", "/knn/nearest", "/knn/permutation")}}
    path = os.path.join(HERE, name + ".json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{name}.json  {os.path.getsize(path) / 1024:.0f} KiB")


def main():
    if not os.path.isdir(REF):
        sys.exit("oracle/_ref is not built (make -C oracle ref, where the reference is mounted)")
    if sys.argv[1:] == ["files"]:
        files_case("files_3d_n20000_k128", "periodic_3d_n20000_k128")
        return
    knn_case("knn_2d_n2000_k16", 2, 2000, 16, 101)
    knn_case("knn_3d_n13000_k32", 3, 13000, 32, 102)
    mesh_case("mesh_2d_n2000_k32", 2, 2000, 32, 103, head_cells=2000)
    mesh_case("mesh_3d_n13000_k64", 3, 13000, 64, 104, head_cells=1500)
    periodic_case("periodic_2d_n10000_k48", 2, 10000, 48, 105, 0.04, 1.4, 0.3, 4)
    periodic_case("periodic_3d_n20000_k128", 3, 20000, 128, 106, 0.12, 1.4, 0.3, 3)
    files_case("files_3d_n20000_k128", "periodic_3d_n20000_k128")


if __name__ == "__main__":
    main()
