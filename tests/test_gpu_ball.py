"""The vertex-ball completion on the GPU (k_voronoi_ball).

Seeds still uncertified when the certificate restarts reach k_max are
finished by the empty-sphere test (no reference counterpart; the reference
leaves them with security_radius_not_reached, voronoi.cpp:145-147).  A low
k_max sends most seeds of a clustered set through it; the mesh must equal
the oracle's (the same completion, pinned against exhaustive rows in
test_oracle_ball.py): statuses, kcert, topology and face normals bit-exact,
volumes and areas to the clip's usual round-off.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import gaussian_mixture

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n,boundary,k,k_max", [(3, 20000, "periodic", 16, 32), (3, 20000, "reflective", 16, 24),
                                                    (2, 12000, "reflective", 8, 16)])
def test_ball_completion_equals_oracle(ctx, dim, n, boundary, k, k_max):
    print(f"ball dim={dim} n={n} {boundary} k={k} k_max={k_max}", flush=True)
    pts = np.ascontiguousarray(gaussian_mixture(n, 8, 0.004, 0.04, 0.3, 21 + dim)[:, :dim])
    g = api.Mesh(ctx, pts, boundary=boundary, buff=0.12, k=k, k_max=k_max, restart=True)
    ref = O.Mesh(pts, boundary, buff=0.12, k=k, restart=True, k_max=k_max)
    out = g.copy_out()
    assert g.n_ball > 0.05 * n  # the completion is exercised
    assert np.sum(out["kcert"] == k_max + 1) > 0.05 * n
    for key in ("status", "kcert", "face_counts", "face_ptr", "neighbor_cell", "face_normal"):
        assert np.array_equal(out[key], ref.get(key)), key
    np.testing.assert_array_equal(out["seeds"], ref.get("seeds"))
    np.testing.assert_allclose(out["volumes"], ref.get("volumes"), rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(out["face_area"], ref.get("face_area"), rtol=1e-11, atol=1e-300)
    print(f"ball: {g.n_ball} seeds completed, {g.n_failed} left failed", flush=True)
