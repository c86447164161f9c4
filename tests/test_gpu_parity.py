"""GPU parity: libpgpu (through the C ABI) against the FP64 oracle.

Bar: KNN lists, sorted points, permutation and Voronoi topology (face counts,
neighbour ids, statuses, certified k) bit-exact; normals bit-exact (same
arithmetic on the same planes); volumes and areas within rtol 1e-12 (the GPU
sums fan triangles in its own vertex order); cell and face centroids within
1e-13 absolute (coordinates of order one; 2D cell centroids 5e-11); hydro states within rtol 1e-11
after several steps (geometry round-off propagated; dt within rtol 1e-13
because cbrt is not pinned across libms).
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import gaussian_mixture, sod_state, uniform_points

pytestmark = pytest.mark.gpu


def cubic_lattice(m):
    g = (np.arange(m) + 0.5) / m
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


@pytest.mark.parametrize("dim,n,k,seed", [(2, 2000, 16, 1), (3, 13000, 32, 2), (2, 5000, 48, 3)])
def test_knn_bitexact(ctx, dim, n, k, seed):
    print(f"knn dim={dim} n={n} k={k} seed={seed}", flush=True)
    pts = uniform_points(n, dim, seed)
    kp = api.KnnProblem(ctx, pts)
    kp.solve(k)
    N, s, perm, kn = O.knn(pts, k)
    assert kp.N_grid == N
    assert np.array_equal(kp.get_points(), s)
    assert np.array_equal(kp.get_permutation(), perm)
    assert np.array_equal(kp.get_knearest(), kn)


@pytest.mark.parametrize("dim,n,k", [(3, 12000, 8), (2, 6000, 16), (3, 12000, 24), (3, 12000, 64), (3, 12000, 100),
                                     (2, 6000, 200), (3, 13000, 256), (2, 1000, 250)])
def test_knn_clustered_bitexact(ctx, dim, n, k):
    """Clustered seeds (dense clumps, sparse space between them): the wave
    query's first ball, guessed from the density around the query, is often
    far too small or far too large there, and the clumps' grid cells hold
    hundreds of points each (searched through their sub-grids).  Rows
    bit-exact against the oracle's ring search."""
    print(f"knn clustered dim={dim} n={n} k={k}", flush=True)
    pts = np.ascontiguousarray(gaussian_mixture(n, 8, 0.003, 0.03, 0.3, 40 + k)[:, :dim])
    kp = api.KnnProblem(ctx, pts)
    kp.solve(k)
    _, _, _, kn = O.knn(pts, k)
    assert np.array_equal(kp.get_knearest(), kn)


def test_knn_ties_lattice(ctx):
    """Cubic lattice: every distance is tied many times; order is (dist^2, original index)."""
    pts = cubic_lattice(24)
    print("knn lattice n=", len(pts), flush=True)
    kp = api.KnnProblem(ctx, pts)
    kp.solve(26)
    _, _, _, kn = O.knn(pts, 26)
    assert np.array_equal(kp.get_knearest(), kn)


def _cmp_mesh(g, ref, geom_rtol=1e-12):
    # 2D volumes: the reference's shoelace (cell.cpp:475-498) works in absolute
    # coordinates, so its round-off is absolute (~u |x|^2), not relative.
    vol_atol = 1e-15 if ref.dim == 2 else 0.0
    assert np.array_equal(g["status"], ref.get("status"))
    assert np.array_equal(g["kcert"], ref.get("kcert"))
    assert np.array_equal(g["face_counts"], ref.get("face_counts"))
    assert np.array_equal(g["face_ptr"], ref.get("face_ptr"))
    assert np.array_equal(g["neighbor_cell"], ref.get("neighbor_cell"))
    assert np.array_equal(g["face_normal"], ref.get("face_normal"))
    np.testing.assert_allclose(g["volumes"], ref.get("volumes"), rtol=geom_rtol, atol=vol_atol)
    np.testing.assert_allclose(g["face_area"], ref.get("face_area"), rtol=geom_rtol, atol=1e-300)
    np.testing.assert_array_equal(g["seeds"], ref.get("seeds"))
    # centroids (no reference counterpart; the moving-mesh face velocity and
    # the regularised motion use them): certified cells, coordinates of order
    # one, each summed over fan triangles in the kernel's own vertex order;
    # 2D cell centroids share the absolute-coordinate cancellation of the 2D
    # area (~u |x|^2 / A per coordinate)
    ok = g["status"] == 4
    c_atol = 5e-11 if ref.dim == 2 else 1e-13
    np.testing.assert_allclose(g["cell_centroid"][ok], ref.get("cell_centroid")[ok], rtol=0, atol=c_atol)
    np.testing.assert_allclose(g["face_centroid"], ref.get("face_centroid"), rtol=0, atol=1e-13)


@pytest.mark.parametrize("dim,n,k,seed", [(2, 2000, 32, 4), (3, 13000, 64, 5), (3, 13000, 32, 6)])
def test_mesh_reference_semantics(ctx, dim, n, k, seed):
    """voronoi::compute_mesh at fixed k, no restart: statuses (incl. failures) match."""
    print(f"mesh none dim={dim} n={n} k={k}", flush=True)
    pts = uniform_points(n, dim, seed)
    m = api.Mesh(ctx, pts, boundary="none", k=k, restart=False)
    ref = O.Mesh(pts, "none", k=k, restart=False)
    assert m.n_failed == ref.n_failed
    _cmp_mesh(m.copy_out(), ref)


@pytest.mark.parametrize("dim,n,k,buff", [(2, 10000, 16, 0.04), (3, 20000, 32, 0.12)])
def test_periodic_restart(ctx, dim, n, k, buff):
    """Periodic ghosts + certificate restarts: every cell certified, identical to the oracle."""
    print(f"periodic restart dim={dim} n={n} k={k}", flush=True)
    pts = uniform_points(n, dim, 7)
    m = api.Mesh(ctx, pts, boundary="periodic", buff=buff, k=k, restart=True)
    ref = O.Mesh(pts, "periodic", buff=buff, k=k, restart=True)
    assert m.n_failed == 0 and ref.n_failed == 0
    g = m.copy_out()
    _cmp_mesh(g, ref)
    assert np.array_equal(g["ghost_ids"], ref.get("ghost_ids"))
    vol = g["volumes"][: m.n_hydro].sum()
    assert abs(vol - 1.0) < 1e-12


@pytest.mark.parametrize("flux", ["reference", "physical"])
def test_hydro_steps(ctx, flux):
    dim, n, k, gamma = 2, 10000, 48, 1.4
    print(f"hydro flux={flux}", flush=True)
    pts = uniform_points(n, dim, 8)
    rho, vel, E = sod_state(pts, gamma)
    m = api.Mesh(ctx, pts, boundary="periodic", k=k, restart=True)
    ref = O.Mesh(pts, "periodic", k=k, restart=True)
    st = api.HydroState(ctx, m, rho, vel, E)
    r, v, e = rho.copy(), vel.copy(), E.copy()
    mode = 0 if flux == "reference" else 1
    for _ in range(3):
        dt_g = st.dt_cfl(gamma, 0.3)
        dt_o = ref.dt_cfl(gamma, 0.3, r, v, e)
        assert abs(dt_g - dt_o) <= 1e-13 * dt_o
        st.step(gamma, dt_o, flux)
        ref.hydro_step(gamma, dt_o, r, v, e, flux_mode=mode)
    rg, vg, eg = st.copy_out()
    np.testing.assert_allclose(rg, r, rtol=1e-11)
    np.testing.assert_allclose(eg, e, rtol=1e-11)
    np.testing.assert_allclose(vg, v, rtol=1e-9, atol=1e-12)


def test_reflective_box(ctx):
    """Box planes are walls: no ghosts, unit volume, no mass through walls."""
    dim, n, k, gamma = 3, 20000, 32, 1.4
    pts = uniform_points(n, dim, 9)
    m = api.Mesh(ctx, pts, boundary="reflective", k=k, restart=True)
    ref = O.Mesh(pts, "reflective", k=k, restart=True)
    g = m.copy_out()
    _cmp_mesh(g, ref)
    assert abs(g["volumes"].sum() - 1.0) < 1e-12
    rho, vel, E = sod_state(pts, gamma)
    st = api.HydroState(ctx, m, rho, vel, E)
    mass0 = np.sum(rho * g["volumes"])
    for _ in range(3):
        st.step(gamma, st.dt_cfl(gamma, 0.3), "physical")
    rg, _, _ = st.copy_out()
    assert abs(np.sum(rg * g["volumes"]) - mass0) < 1e-12 * mass0


@pytest.mark.parametrize("lanes,slots", [(16, 48), (8, 64), (8, 48)])
def test_clip_shapes_equal_oracle(ctx, lanes, slots):
    """Every shape of the fast clip kernel gives the oracle's mesh."""
    print(f"clip shape {lanes}x{slots}", flush=True)
    pts = uniform_points(13000, 3, 9)
    api.set_clip_config(lanes, slots)
    try:
        m = api.Mesh(ctx, pts, boundary="none", k=48, restart=False)
        g = m.copy_out()
    finally:
        api.set_clip_config(16, 64)
    _cmp_mesh(g, O.Mesh(pts, "none", k=48, restart=False))


@pytest.mark.parametrize("dim,n,k,lanes,slots", [(3, 20000, 16, 16, 64), (3, 20000, 24, 8, 48), (2, 10000, 8, 16, 48)])
def test_restart_resume_equals_scratch(ctx, dim, n, k, lanes, slots):
    """Certificate restarts that continue the saved cell give the mesh of
    restarts that clip from the first neighbour, and the oracle's."""
    print(f"resume dim={dim} n={n} k={k} shape {lanes}x{slots}", flush=True)
    pts = uniform_points(n, dim, 12)
    api.set_clip_config(lanes, slots)
    try:
        assert api.get_clip_resume()
        a = api.Mesh(ctx, pts, boundary="periodic", buff=0.12, k=k, restart=True)
        api.set_clip_resume(False)
        b = api.Mesh(ctx, pts, boundary="periodic", buff=0.12, k=k, restart=True)
    finally:
        api.set_clip_resume(True)
        api.set_clip_config(16, 64)
    assert a.n_restarted > 0.5 * a.n_seeds and a.n_restarted == b.n_restarted
    ga, gb = a.copy_out(), b.copy_out()
    for key in ga:
        assert np.array_equal(ga[key], gb[key]), key
    _cmp_mesh(ga, O.Mesh(pts, "periodic", buff=0.12, k=k, restart=True))


@pytest.mark.parametrize("boundary", ["none", "periodic"])
def test_mesh_knn_problem_equals_oracle(ctx, boundary):
    """pg_mesh_get_knn (the KNN dump of a mesh build): sorted points, rows and
    permutation equal the oracle's knn problem of the same build."""
    pts = uniform_points(13000, 3, 13)
    m = api.Mesh(ctx, pts, boundary=boundary, k=32, restart=True)
    P, rows, perm = m.knn()
    ref = O.Mesh(pts, boundary, k=32, restart=False)
    assert np.array_equal(P, ref.get("sorted_points"))
    assert np.array_equal(rows, ref.get("knearest"))
    assert np.array_equal(perm, ref.get("permutation"))
    km = api.KModel(ctx)
    from proteusgpu_amd._native import PgError
    with pytest.raises(PgError):
        api.Mesh(ctx, pts, boundary=boundary, k=32, kmodel=km).knn()
