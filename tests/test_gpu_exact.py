"""The exact-predicate fallback on the GPU (pg_mesh_params.exact).

Lattices make the reference's conflict predicate degenerate (cospherical
neighbours): the reference leaves those cells failed with
needs_exact_predicates (cell.cpp:216-218, voronoi.cpp:104-110).  With exact=1
the library clips them again with the exact predicate and symbolic
tie-breaking (exact.h); the oracle applies the same rule through its own
exact arithmetic (sign-magnitude integers, Leibniz determinants, global
coordinates; pinned by rational arithmetic in test_oracle_exact.py).  Here:
topology, statuses and kcert bit-exact against the oracle, geometry against
the oracle and against the lattice's closed forms, inert on generic input,
and a moving-mesh run from an exact lattice against the oracle's.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import sod_state, uniform_points

pytestmark = pytest.mark.gpu


def lattice(dim, m):
    g = (np.arange(m) + 0.5) / m
    return np.stack([a.ravel() for a in np.meshgrid(*([g] * dim), indexing="ij")], 1)


def _cmp(g, ref, rtol):
    for key in ("status", "kcert", "face_counts", "face_ptr", "neighbor_cell", "face_normal"):
        assert np.array_equal(g[key], ref.get(key)), key
    np.testing.assert_array_equal(g["seeds"], ref.get("seeds"))
    np.testing.assert_allclose(g["volumes"], ref.get("volumes"), rtol=rtol, atol=1e-15)
    # zero-area faces: their areas are round-off of either side's fan sum
    scale = np.max(ref.get("face_area"))
    np.testing.assert_allclose(g["face_area"], ref.get("face_area"), rtol=rtol, atol=1e-13 * scale)


CASES = [(2, 40, 16, "reflective"), (2, 40, 16, "periodic"), (3, 16, 32, "reflective"), (3, 20, 32, "periodic")]


@pytest.mark.parametrize("dim,m,k,boundary", CASES)
def test_exact_lattice_equals_oracle(ctx, dim, m, k, boundary):
    print(f"exact lattice dim={dim} m={m} {boundary}", flush=True)
    pts = lattice(dim, m)
    n = len(pts)
    buff = 0.12 if dim == 3 else 0.06
    g = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k, restart=True, exact=True)
    ref = O.Mesh(pts, boundary, buff=buff, k=k, restart=True, exact=True)
    assert g.n_failed == ref.n_failed == 0
    assert g.n_exact > 0.5 * n and g.displacement == 0.0
    out = g.copy_out()
    # geometry rtol 1e-11: each lattice vertex is a cluster of coincident
    # vertices summed in each side's own order
    _cmp(out, ref, 1e-11)
    h = 1.0 / m
    np.testing.assert_allclose(out["volumes"][:n], h ** dim, rtol=2e-12)
    assert abs(np.sum(out["volumes"][:n]) - 1.0) < 1e-11


def test_exact_is_inert_on_generic_input(ctx):
    pts = uniform_points(20000, 3, 31)
    a = api.Mesh(ctx, pts, boundary="periodic", buff=0.12, k=32, restart=True)
    b = api.Mesh(ctx, pts, boundary="periodic", buff=0.12, k=32, restart=True, exact=True)
    assert b.n_exact == 0
    ga, gb = a.copy_out(), b.copy_out()
    for key in ga:
        assert np.array_equal(ga[key], gb[key]), key


def test_moving_mesh_from_an_exact_lattice(ctx):
    """Moving-mesh Sod from an exact 2D lattice with the exact fallback:
    every rebuild certifies, state and generators follow the oracle's run
    cell by cell, and mass and energy are conserved."""
    pts = lattice(2, 32)
    rho, vel, E = sod_state(pts, 1.4)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=16, gamma=1.4, cfl=0.3, exact=True)
    ref = O.Sim(pts, rho, vel, E, boundary="reflective", k=16, gamma=1.4, cfl=0.3, exact=True)
    t0 = sim.totals()
    for step in range(4):
        dt = sim.step()
        dr = ref.step()
        print(f"step {step}: dt {dt:.6e} oracle {dr:.6e}", flush=True)
        np.testing.assert_allclose(dt, dr, rtol=1e-10)
    x, r, v, e = sim.get()
    xo, ro, vo, eo = ref.get()
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-13)
    np.testing.assert_allclose(r, ro, rtol=1e-9)
    np.testing.assert_allclose(e, eo, rtol=1e-9)
    t1 = sim.totals()
    np.testing.assert_allclose(t1[0], t0[0], rtol=1e-13)
    np.testing.assert_allclose(t1[-1], t0[-1], rtol=1e-13)
