"""The oracle reproduces the reference's own outputs bit for bit.

Fixtures in tests/golden were written by tests/golden/make_golden.py from the
reference's sources (knn::solve, voronoi::compute_mesh,
voronoi::compute_periodic_mesh, hydro::dt_CFL + hydro::hydro_step) on the
project's seeded inputs.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd.generators import sod_state, uniform_points

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"))


@pytest.mark.parametrize("name", ["knn_2d_n2000_k16", "knn_3d_n13000_k32"])
def test_knn_golden(name):
    g = load(name)
    dim, n, k, seed = int(g["dim"]), int(g["n"]), int(g["k"]), int(g["seed"])
    pts = uniform_points(n, dim, seed)
    assert sha(pts) == str(g["pts_sha"]), "generator drifted"
    N, s, perm, kn = O.knn(pts, k)
    assert N == int(g["N_grid"])
    assert np.array_equal(perm, g["permutation"])
    assert sha(s) == str(g["sorted_sha"])
    assert np.array_equal(kn[:3000], g["knearest_head"])
    assert sha(kn) == str(g["knearest_sha"])


@pytest.mark.parametrize("name", ["mesh_2d_n2000_k32", "mesh_3d_n13000_k64"])
def test_mesh_golden(name):
    g = load(name)
    dim, n, k, seed = int(g["dim"]), int(g["n"]), int(g["k"]), int(g["seed"])
    pts = uniform_points(n, dim, seed)
    assert sha(pts) == str(g["pts_sha"])
    m = O.Mesh(pts, "none", k=k, restart=False)
    assert np.array_equal(m.get("face_counts"), g["face_counts"].astype(np.int64))
    assert np.array_equal(m.get("neighbor_cell"), g["neighbor_cell"])
    assert np.array_equal(m.get("volumes"), g["volumes"])
    area, nrm = m.get("face_area"), m.get("face_normal")
    h = len(g["face_area_head"])
    assert np.array_equal(area[:h], g["face_area_head"])
    assert np.array_equal(nrm[:h], g["face_normal_head"])
    assert sha(area) == str(g["face_area_sha"])
    assert sha(nrm) == str(g["face_normal_sha"])
    assert sha(m.get("seeds")) == str(g["seeds_sha"])
    # failed cells of the reference have no faces; the oracle reports them
    assert np.array_equal(m.get("status") != 4, g["face_counts"] == 0)


@pytest.mark.parametrize("name", ["periodic_2d_n10000_k48", "periodic_3d_n20000_k128"])
def test_periodic_hydro_golden(name):
    g = load(name)
    dim, n, k, seed = int(g["dim"]), int(g["n"]), int(g["k"]), int(g["seed"])
    buff, gamma, cfl, steps = float(g["buff"]), float(g["gamma"]), float(g["cfl"]), int(g["steps"])
    pts = uniform_points(n, dim, seed)
    assert sha(pts) == str(g["pts_sha"])
    rho, vel, E = sod_state(pts, gamma)
    m = O.Mesh(pts, "periodic", buff=buff, k=k, restart=False)
    assert m.n_seeds == int(g["n_seeds"])
    assert np.array_equal(m.get("ghost_ids"), g["ghost_ids"])
    assert np.array_equal(m.get("face_counts"), g["face_counts"].astype(np.int64))
    nb = m.get("neighbor_cell")
    assert np.array_equal(nb[: len(g["neighbor_cell_hydro"])], g["neighbor_cell_hydro"])
    assert sha(nb) == str(g["neighbor_cell_sha"])
    assert np.array_equal(m.get("volumes"), g["volumes"])
    assert sha(m.get("face_area")) == str(g["face_area_sha"])
    dts = []
    for _ in range(steps):
        dt = m.dt_cfl(gamma, cfl, rho, vel, E)
        dts.append(dt)
        m.hydro_step(gamma, dt, rho, vel, E, flux_mode=0)
    assert np.array_equal(np.array(dts), g["dts"])
    assert np.array_equal(rho, g["rho"])
    assert np.array_equal(vel.ravel(), g["vel"])
    assert np.array_equal(E, g["E"])
