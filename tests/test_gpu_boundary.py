"""The per-face and per-problem entry points of the drop-in boundary.

pg_hll_flux is hydro::riemann_hll (riemann.cpp:7-45) for many faces; it
runs the same device HLL as the cell kernels and must equal the oracle's
restatement bit for bit.  pg_voronoi_compute is voronoi::compute_cells
(voronoi.cpp:68-111): the cells of a solved KNN problem, no restarts, in the
problem's sorted order (neighbour ids in sorted-index space); mapped through
the problem's permutation it must equal the oracle's fixed-k mesh (statuses,
topology and normals bit-exact, geometry rtol 1e-12).
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import uniform_points

pytestmark = pytest.mark.gpu


def random_faces(dim, n, gamma, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 2.0, (2, n))
    v = rng.uniform(-2.0, 2.0, (2, n, dim))
    P = rng.uniform(0.05, 3.0, (2, n))
    E = P / (gamma - 1.0) + 0.5 * rho * np.sum(v * v, axis=2)
    states = [np.concatenate([rho[s][:, None], v[s], E[s][:, None]], axis=1) for s in (0, 1)]
    nrm = rng.standard_normal((n, 3))
    if dim == 2:
        nrm[:, 2] = 0.0
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    wall = (rng.random(n) < 0.2).astype(np.uint8)
    return states[0], states[1], nrm, wall


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", ["reference", "physical"])
def test_hll_flux_equals_oracle(ctx, dim, flux):
    gamma = 1.4
    L, R, nrm, wall = random_faces(dim, 3000, gamma, 5 + dim)
    got = api.hll_flux(ctx, gamma, L, R, nrm, wall, flux)
    sgn = 1.0 if flux == "reference" else -1.0

    def five(s):  # the oracle's (rho, vx, vy, vz, E) layout
        out = np.zeros(5)
        out[0], out[1:1 + dim], out[4] = s[0], s[1:1 + dim], s[dim + 1]
        return out

    for i in range(len(L)):
        ref = O.hll(dim, gamma, five(L[i]), five(R[i]), sgn * nrm[i], bool(wall[i]))
        exp = np.concatenate([[ref[0]], ref[1:1 + dim], [ref[4]]])
        assert np.array_equal(got[i], exp), (i, got[i], exp)
    # a wall carries no mass, whatever the state
    assert np.all(got[wall == 1, 0] == 0.0)


def test_hll_flux_consistency_and_arguments(ctx):
    """F(U, U) is the Euler flux along the normal; bad arguments are errors."""
    gamma = 5.0 / 3.0
    L, _, nrm, _ = random_faces(3, 500, gamma, 11)
    F = api.hll_flux(ctx, gamma, L, L, nrm)
    rho, v, E = L[:, 0], L[:, 1:4], L[:, 4]
    P = (gamma - 1.0) * (E - 0.5 * rho * np.sum(v * v, axis=1))
    un = np.sum(v * nrm, axis=1)
    np.testing.assert_allclose(F[:, 0], rho * un, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(F[:, 1:4], rho[:, None] * un[:, None] * v + P[:, None] * nrm, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(F[:, 4], (E + P) * un, rtol=1e-11, atol=1e-12)
    with pytest.raises(RuntimeError):
        api.hll_flux(ctx, 1.0, L, L, nrm)  # gamma must exceed 1


@pytest.mark.parametrize("dim,n,k,seed", [(2, 2000, 16, 21), (3, 13000, 16, 22), (3, 13000, 32, 23)])
def test_voronoi_compute_equals_oracle(ctx, dim, n, k, seed):
    print(f"compute_cells dim={dim} n={n} k={k}", flush=True)
    pts = uniform_points(n, dim, seed)
    kp = api.KnnProblem(ctx, pts)
    kp.solve(k)
    perm = kp.get_permutation().astype(np.int64)
    mesh, status = api.Mesh.from_knn(ctx, kp)
    ref = O.Mesh(pts, "none", k=k, restart=False)
    ost = ref.get("status")
    assert np.array_equal(status, ost[perm])
    assert mesh.n_failed == int(np.sum(ost != 4))
    g = mesh.copy_out()
    assert np.array_equal(g["status"], ost[perm])
    assert np.array_equal(g["kcert"], ref.get("kcert")[perm])
    assert np.array_equal(g["face_counts"], ref.get("face_counts")[perm])
    # seeds of extracted cells (failed cells are not extracted: zeros, as the oracle's)
    np.testing.assert_array_equal(g["seeds"], ref.get("seeds")[perm])
    ok = g["status"] == 4
    np.testing.assert_array_equal(g["seeds"][ok, :dim], kp.get_points()[ok])
    np.testing.assert_allclose(g["volumes"], ref.get("volumes")[perm], rtol=1e-12, atol=1e-15 if dim == 2 else 0)
    # faces of sorted cell s are the oracle's faces of cell perm[s], with
    # neighbour ids in sorted-index space
    ofp, ofc = ref.get("face_ptr"), ref.get("face_counts")
    idx = np.concatenate([np.arange(ofp[o], ofp[o] + ofc[o]) for o in perm])
    nb = g["neighbor_cell"].astype(np.int64)
    mapped = np.where(nb >= 0, perm[np.maximum(nb, 0)], nb)
    assert np.array_equal(mapped, ref.get("neighbor_cell")[idx])
    assert np.array_equal(g["face_normal"], ref.get("face_normal")[idx])
    np.testing.assert_allclose(g["face_area"], ref.get("face_area")[idx], rtol=1e-12, atol=1e-300)


def test_unpermute_equals_compute_mesh(ctx):
    """compute_cells + unpermute_vmesh == compute_mesh (voronoi.cpp:17-65)."""
    pts = uniform_points(13000, 3, 24)
    kp = api.KnnProblem(ctx, pts)
    kp.solve(32)
    mesh, _ = api.Mesh.from_knn(ctx, kp)
    mesh.unpermute()
    direct = api.Mesh(ctx, pts, boundary="none", k=32, restart=False)
    a, b = mesh.copy_out(), direct.copy_out()
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    with pytest.raises(RuntimeError):
        mesh.unpermute()

