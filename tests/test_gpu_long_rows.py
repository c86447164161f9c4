"""Certificate restarts beyond 256 neighbours.

A seed beside a dense cluster has a security radius that holds the whole
cluster: its certificate needs a row longer than 256.  The restart chain
continues at 384, 512, 768 and 1024 (the wave KNN keeps up to 1024 entries;
a resumed seed's row skips its known prefix), and the cells equal the
oracle's at the same cap bit for bit.  Seeds the cap still leaves
uncertified go to the vertex-ball completion (kcert = k_max + 1) in both.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import uniform_points

pytestmark = pytest.mark.gpu


def clustered(ncl, sig, n, seed):
    rng = np.random.default_rng(seed)
    c = np.clip(0.5 + sig * rng.standard_normal((ncl, 3)), 0.0, 0.999999)
    return np.concatenate([c, uniform_points(n - ncl, 3, 77)])


def test_knn_rows_of_256_equal_oracle(ctx):
    """The longest first-pass row (longer rows only come from restarts)."""
    pts = uniform_points(13000, 3, 14)
    print("knn k=256", flush=True)
    kp = api.KnnProblem(ctx, pts)
    kp.solve(256)
    _, _, _, kn = O.knn(pts, 256)
    assert np.array_equal(kp.get_knearest(), kn)


@pytest.mark.parametrize("k_max", [256, 1024])
def test_cluster_neighbours_past_256(ctx, k_max):
    pts = clustered(15000, 0.004, 20000, 5)
    print(f"clustered 20k, k_max={k_max}", flush=True)
    m = api.Mesh(ctx, pts, boundary="reflective", k=32, restart=True, k_max=k_max)
    ref = O.Mesh(pts, "reflective", k=32, restart=True, k_max=k_max)
    g = m.copy_out()
    assert m.n_failed == ref.n_failed
    for key in ("status", "kcert", "face_counts", "neighbor_cell", "face_normal"):
        assert np.array_equal(g[key], ref.get(key)), key
    np.testing.assert_allclose(g["volumes"], ref.get("volumes"), rtol=1e-10)
    kc = g["kcert"]
    rows = kc[(kc > 0) & (kc <= k_max)]
    if k_max == 256:
        # the cap leaves seeds to the completion
        assert 0 < np.count_nonzero(kc == 257) <= m.n_ball
    else:
        # rows past 256 certify cells the 256 cap leaves to the completion
        assert np.count_nonzero(rows > 256) > 100 and m.k_used > 256
