"""Degenerate inputs on the GPU: exact lattices and the degeneracy fallback.

An exact cubic lattice has cospherical neighbours, so the reference's
conflict predicate is zero (below its eps) and the cells fail with
needs_exact_predicates (cell.cpp:187-217); the reference leaves them failed
(voronoi.cpp:104-110).  Without the fallback the library reproduces those
statuses bit for bit.  With pg_mesh_params.perturb the build is repeated on
generators displaced by the oracle's displace(): the resulting mesh must equal
the oracle's mesh of the same displaced points (topology, statuses, seeds
bit-exact; geometry rtol 1e-11, see below), which also pins the device
displacement.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import sod_state

pytestmark = pytest.mark.gpu


def lattice(dim, m):
    g = (np.arange(m) + 0.5) / m
    return np.stack([a.ravel() for a in np.meshgrid(*([g] * dim), indexing="ij")], 1)


def _cmp_mesh(g, ref, vol_atol, rtol=1e-12):
    for key in ("status", "kcert", "face_counts", "face_ptr", "neighbor_cell", "face_normal"):
        assert np.array_equal(g[key], ref.get(key)), key
    np.testing.assert_array_equal(g["seeds"], ref.get("seeds"))
    np.testing.assert_allclose(g["volumes"], ref.get("volumes"), rtol=rtol, atol=vol_atol)
    np.testing.assert_allclose(g["face_area"], ref.get("face_area"), rtol=rtol, atol=1e-300)


CASES = [(2, 40, 16, "reflective"), (2, 40, 16, "periodic"), (3, 24, 32, "reflective"), (3, 24, 32, "periodic")]


@pytest.mark.parametrize("dim,m,k,boundary", CASES)
def test_exact_lattice_fails_as_the_reference(ctx, dim, m, k, boundary):
    print(f"exact lattice dim={dim} m={m} {boundary}", flush=True)
    pts = lattice(dim, m)
    buff = 0.12 if dim == 3 else 0.06
    g = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k, restart=True)
    ref = O.Mesh(pts, boundary, buff=buff, k=k, restart=True)
    assert g.n_failed == ref.n_failed > 0.9 * len(pts)
    assert g.displacement == 0.0 and g.n_degenerate == 0  # the fallback is off by default
    _cmp_mesh(g.copy_out(), ref, 1e-15 if dim == 2 else 0.0)


@pytest.mark.parametrize("dim,m,k,boundary", CASES)
def test_fallback_equals_oracle_on_displaced_points(ctx, dim, m, k, boundary):
    print(f"fallback lattice dim={dim} m={m} {boundary}", flush=True)
    pts = lattice(dim, m)
    n = len(pts)
    buff = 0.12 if dim == 3 else 0.06
    g = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k, restart=True, perturb=1e-6)
    assert g.n_failed == 0
    assert g.n_degenerate > 0.9 * n
    # first retry: amplitude perturb x n^(-1/dim)
    np.testing.assert_allclose(g.displacement, 1e-6 * n ** (-1.0 / dim), rtol=1e-15)
    d = O.displace(pts, g.displacement, 0, boundary == "periodic")
    ref = O.Mesh(d, boundary, buff=buff, k=k, restart=True)
    assert ref.n_failed == 0
    out = g.copy_out()
    # geometry rtol 1e-11: every lattice vertex became a cluster of vertices
    # within ~1e-6 h of each other, and the fan sums (summed in each side's
    # own vertex order) cancel about a hundred times more than on generic
    # inputs (largest difference seen: 1.004e-12 in one 3D volume)
    _cmp_mesh(out, ref, 1e-15 if dim == 2 else 0.0, rtol=1e-11)
    # the cells are the lattice's cubes to O(displacement)
    h = 1.0 / m
    np.testing.assert_allclose(out["volumes"][:n], h ** dim, rtol=1e-5)


def test_fallback_keeps_nondegenerate_builds_unchanged(ctx):
    """A build without degenerate cells is not displaced."""
    from proteusgpu_amd.generators import uniform_points
    pts = uniform_points(20000, 3, 31)
    a = api.Mesh(ctx, pts, boundary="reflective", k=32, restart=True)
    b = api.Mesh(ctx, pts, boundary="reflective", k=32, restart=True, perturb=1e-6)
    assert b.displacement == 0.0 and b.n_degenerate == 0
    ga, gb = a.copy_out(), b.copy_out()
    for key in ga:
        assert np.array_equal(ga[key], gb[key]), key


def test_moving_mesh_from_an_exact_lattice(ctx):
    """A moving-mesh Sod run from an exact 2D lattice: refused without the
    fallback (every cell fails), conservative with it."""
    pts = lattice(2, 48)
    rho, vel, E = sod_state(pts, 1.4)
    with pytest.raises(RuntimeError):
        api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=16, gamma=1.4, cfl=0.3)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=16, gamma=1.4, cfl=0.3, perturb=1e-6)
    t0 = sim.totals()
    for _ in range(5):
        sim.step()
    t1 = sim.totals()
    np.testing.assert_allclose(t1[0], t0[0], rtol=1e-13)  # mass
    np.testing.assert_allclose(t1[-1], t0[-1], rtol=1e-13)  # energy
    _, r, _, _ = sim.get()
    assert np.all(np.isfinite(r)) and np.all(r > 0)


@pytest.mark.parametrize("dim,n,k", [(2, 4000, 16), (3, 13000, 32)])
def test_duplicate_generators(ctx, dim, n, k):
    """Duplicate generators give a zero bisector (REF-Q10): both copies stay
    uncertified (the final certificate overwrites their needs_exact_predicates,
    voronoi.cpp:145-147) and the cells around them, which see the pair's two
    identical planes, fail with needs_exact_predicates, all as in the
    reference; the fallback separates the pair and every cell certifies,
    equal to the oracle's displaced mesh."""
    from proteusgpu_amd.generators import uniform_points
    print(f"duplicates dim={dim} n={n}", flush=True)
    pts = uniform_points(n, dim, 41)
    pts[5], pts[100], pts[n // 2] = pts[17], pts[101], pts[n // 2 + 3]
    g = api.Mesh(ctx, pts, boundary="reflective", k=k, restart=True)
    ref = O.Mesh(pts, "reflective", k=k, restart=True)
    out = g.copy_out()
    _cmp_mesh(out, ref, 1e-15 if dim == 2 else 0.0)
    st = out["status"]
    for a, b in ((5, 17), (100, 101), (n // 2, n // 2 + 3)):
        assert st[a] != 4 and st[b] != 4
    assert np.sum(st == 5) > 0
    f = api.Mesh(ctx, pts, boundary="reflective", k=k, restart=True, perturb=1e-6)
    assert f.n_failed == 0 and f.n_degenerate == np.sum(st == 5) + np.sum(st == 2)
    d = O.displace(pts, f.displacement, 0, False)
    # geometry rtol 1e-8: the separated pair lies ~1e-6 h apart, so the
    # vertices on its bisector are ~1e6 times worse conditioned than usual and
    # the last-ulp differences of evaluation order between the two sides grow
    # accordingly (largest seen: 9.5e-10 in one of 13000 3D volumes)
    _cmp_mesh(f.copy_out(), O.Mesh(d, "reflective", k=k, restart=True), 1e-15 if dim == 2 else 0.0, rtol=1e-8)
