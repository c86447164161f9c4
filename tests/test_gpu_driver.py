"""The reference-shaped driver end to end: the seeded IC writer (make_ic)
writes an IC file in the reference layout, ProteusGPU reads it with the
reference's param.txt keys, builds the periodic mesh, runs the CFL-limited
HLL loop of main.cpp:38-84 and writes the snapshot (and, built with
WRITE_KNN_OUTPUT, the KNN dump).  The files must hold exactly what the
library gives through its own API for the same points: cell volumes, face
areas, neighbour rows and the hydro state after the same steps.
"""
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from proteusgpu_amd import api

pytestmark = pytest.mark.gpu

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "proteusgpu_amd", "host")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def h5(path, name, dtype, tmp):
    out = os.path.join(tmp, name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype)


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_driver_snapshot_and_knn_dump(ctx, tmp_path):
    for exe in ("ProteusGPU_knn", "make_ic"):
        assert os.path.exists(os.path.join(HOST, exe)), f"{exe} not built (make -C proteusgpu_amd/host)"
    tmp = str(tmp_path)
    ic = os.path.join(tmp, "IC.hdf5")
    subprocess.run([os.path.join(HOST, "make_ic"), ic, "3", "20000", "7", "sod", "1.6666666666666667"], check=True)
    outdir = os.path.join(tmp, "out") + "/"
    with open(os.path.join(tmp, "param.txt"), "w") as f:
        f.write(f"ic_file = {ic}\noutput_directory = {outdir}\noutput_mesh_file = mesh.hdf5\n"
                f"output_knn_file = knn.hdf5\nt_end = 0.002\ncfl = 0.4\nboundary = periodic\nbuff = 0.12\n")
    print("driver: running ProteusGPU_knn", flush=True)
    r = subprocess.run([os.path.join(HOST, "ProteusGPU_knn"), os.path.join(tmp, "param.txt")], cwd=tmp,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0
    mesh_f, knn_f = os.path.join(outdir, "mesh.hdf5"), os.path.join(outdir, "knn.hdf5")
    pts = h5(ic, "/seedpos", np.float64, tmp).reshape(-1, 3)
    rho0 = h5(ic, "/rho", np.float64, tmp)
    vel0 = h5(ic, "/vel", np.float64, tmp).reshape(-1, 3)
    E0 = h5(ic, "/Energy", np.float64, tmp)
    n = len(pts)
    assert n == 20000 and np.all((pts >= 0) & (pts < 1))
    # the same mesh through the API (k = _K_ = 32, restarts on; 20k seeds
    # need a ghost layer wider than the default 0.04)
    m = api.Mesh(ctx, pts, boundary="periodic", buff=0.12, k=32, restart=True)
    g = m.copy_out()
    vol = h5(mesh_f, "/cells/volumes", np.float64, tmp)
    assert np.array_equal(vol, g["volumes"][:n])
    area = h5(mesh_f, "/cells/faces/area", np.float64, tmp)
    assert np.array_equal(area, g["face_area"])
    nb = h5(mesh_f, "/cells/faces/neighbor_cell", np.int32, tmp)
    assert np.array_equal(nb, g["neighbor_cell"])
    # the KNN dump: the first n rows of the build's sorted problem
    P, rows, perm = m.knn()
    assert np.array_equal(h5(knn_f, "/knn/nearest", np.uint32, tmp).reshape(n, 32), rows[:n])
    assert np.array_equal(h5(knn_f, "/knn/permutation", np.uint32, tmp), perm[:n])
    assert np.array_equal(h5(knn_f, "/knn/points", np.float64, tmp).reshape(n, 3), P[:n])
    # the hydro loop of main.cpp:38-84 on the same mesh (reference flux)
    st = api.HydroState(ctx, m, rho0, vel0, E0)
    t, t_end, gam = 0.0, 0.002, 5.0 / 3.0
    while t < t_end:
        dt = st.dt_cfl(gam, 0.4)
        if t + dt > t_end:
            dt = t_end - t
        st.step(gam, dt, "reference")
        t += dt
    rho, vel, E = st.copy_out()
    assert np.array_equal(h5(mesh_f, "/hydro/rho", np.float64, tmp), rho)
    assert np.array_equal(h5(mesh_f, "/hydro/vel", np.float64, tmp).reshape(n, 3), vel)
    assert np.array_equal(h5(mesh_f, "/hydro/Energy", np.float64, tmp), E)


def test_reference_shaped_entry_points():
    """voronoi::allocate_vmesh / compute_cells / unpermute_vmesh and
    hydro::riemann_hll through the C++ drop-in layer (src/check/shim_check.cpp)."""
    exe = os.path.join(HOST, "shim_check")
    assert os.path.exists(exe), "shim_check not built (make -C proteusgpu_amd/host)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "SHIM OK" in r.stdout


def _h5_text(path, *args):
    r = subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, check=True)
    return "\n".join(r.stdout.splitlines()[1:])


def _h5_sha(path, name, tmp):
    out = os.path.join(tmp, "sha.bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return hashlib.sha256(open(out, "rb").read()).hexdigest()


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_driver_files_match_reference_writers(tmp_path):
    """The driver built with the reference's mesh semantics at _K_ = 128 (no
    certificate restarts, the reference's face orientation, KNN dump on), on
    the IC of the golden case periodic_3d_n20000_k128 run to the time the
    reference reaches in its three CFL steps, against what the reference's
    own main loop and writers produced from the same make_ic file
    (tests/golden/files_3d_n20000_k128.json, written by
    oracle/ref_harness.cpp's files mode; the reference's readICFile read
    make_ic's file there):
      * snapshot and KNN dump have the reference's groups, dataset names,
        dtypes, shapes and header attributes (h5dump -H, -A);
      * cell ids, face counts and neighbour ids of the snapshot and the
        points, rows and permutation of the KNN dump are the reference's
        bytes (SHA-256);
      * the hydro state is the reference's post-step state (the golden
        periodic case) within the hydro tolerances: rtol 1e-11, velocity
        1e-9 relative / 1e-12 absolute."""
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    fx = json.load(open(os.path.join(here, "files_3d_n20000_k128.json")))
    g = np.load(os.path.join(here, "periodic_3d_n20000_k128.npz"))
    exe = os.path.join(HOST, "ProteusGPU_k128")
    assert os.path.exists(exe), "ProteusGPU_k128 not built (make -C proteusgpu_amd/host)"
    tmp = str(tmp_path)
    ic = os.path.join(tmp, "IC.hdf5")
    n = fx["n"]
    subprocess.run([os.path.join(HOST, "make_ic"), ic, "3", str(n), str(fx["seed"]), "sod", repr(fx["gamma"])],
                   check=True)
    outdir = os.path.join(tmp, "out") + "/"
    with open(os.path.join(tmp, "param.txt"), "w") as f:
        f.write(f"ic_file = {ic}\noutput_directory = {outdir}\noutput_mesh_file = mesh.hdf5\n"
                f"output_knn_file = knn.hdf5\nt_end = {fx['t_end']!r}\ncfl = {fx['cfl']!r}\n"
                f"gamma = {fx['gamma']!r}\nboundary = periodic\nbuff = {fx['buff']!r}\n")
    print("driver: running ProteusGPU_k128 to t_end", fx["t_end"], flush=True)
    r = subprocess.run([exe, os.path.join(tmp, "param.txt")], cwd=tmp, capture_output=True, text=True, timeout=300)
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 0
    mesh_f, knn_f = os.path.join(outdir, "mesh.hdf5"), os.path.join(outdir, "knn.hdf5")
    assert _h5_text(mesh_f, "-H") == fx["mesh_header"]
    assert _h5_text(knn_f, "-H") == fx["knn_header"]
    assert _h5_text(mesh_f, "-A", "-g", "/header") == fx["mesh_attrs"]
    assert _h5_text(knn_f, "-A", "-g", "/header") == fx["knn_attrs"]
    for ds, want in fx["mesh_sha"].items():
        assert _h5_sha(mesh_f, ds, tmp) == want, ds
    for ds, want in fx["knn_sha"].items():
        assert _h5_sha(knn_f, ds, tmp) == want, ds
    np.testing.assert_allclose(h5(mesh_f, "/cells/volumes", np.float64, tmp), g["volumes"][:n], rtol=1e-12)
    np.testing.assert_allclose(h5(mesh_f, "/hydro/rho", np.float64, tmp), g["rho"], rtol=1e-11)
    np.testing.assert_allclose(h5(mesh_f, "/hydro/Energy", np.float64, tmp), g["E"], rtol=1e-11)
    np.testing.assert_allclose(h5(mesh_f, "/hydro/vel", np.float64, tmp), g["vel"], rtol=1e-9, atol=1e-12)
