"""The oracle's vertex-ball completion, pinned by exhaustive rows.

A seed still uncertified at k_max is finished by the empty-sphere test (no
reference counterpart; the reference leaves it failed with
security_radius_not_reached, voronoi.cpp:145-147).  The Voronoi cell of a
seed is unique, so a completed cell must equal the cell a row holding every
other seed gives (certified by exhaustion; periodic: rows of up to 1024,
which certify nearly every cell): the same faces in the same order ((distance, index)
order), volumes and areas to round-off of the different clip order.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd.generators import gaussian_mixture


@pytest.mark.parametrize("dim,n,boundary", [(2, 1500, "reflective"), (3, 2500, "reflective"), (3, 2000, "periodic")])
def test_ball_completion_equals_exhaustive_rows(dim, n, boundary):
    pts = gaussian_mixture(n, 6, 0.004, 0.03, 0.3, 11 + dim)[:, :dim].copy()
    if dim == 2:
        pts = np.ascontiguousarray(pts)
    buff = 0.12
    small = O.Mesh(pts, boundary, buff=buff, k=8, restart=True, k_max=16)
    st = small.get("status")
    kc = small.get("kcert")
    ball = kc == 17
    assert ball.sum() > 20  # the completion is exercised
    assert np.all(st == 4)
    ns = len(st)
    if boundary == "periodic":
        # ghosts make rows longer than the input: rows of up to 1024
        # neighbours certify nearly every cell here without the completion
        full = O.Mesh(pts, boundary, buff=buff, k=8, restart=True, k_max=1024)
        assert np.sum(full.get("kcert") == 1025) < 0.1 * ball.sum()
    else:
        full = O.Mesh(pts, boundary, buff=buff, k=ns - 1, restart=False)
    # cells the long rows certify (a few run into the plane capacity on the
    # way: the rows add every neighbour that clips in distance order, the
    # completion only those its vertex balls meet)
    ok = full.get("status") == 4
    assert np.sum(~ok) <= 0.002 * ns
    fc_a, fp_a, fc_b, fp_b = small.get("face_counts"), small.get("face_ptr"), full.get("face_counts"), full.get("face_ptr")
    nb_a, nb_b = small.get("neighbor_cell"), full.get("neighbor_cell")
    nr_a, nr_b = small.get("face_normal"), full.get("face_normal")
    ar_a, ar_b = small.get("face_area"), full.get("face_area")
    for i in np.nonzero(ok)[0]:
        sa = slice(fp_a[i], fp_a[i] + fc_a[i])
        sb = slice(fp_b[i], fp_b[i] + fc_b[i])
        assert np.array_equal(nb_a[sa], nb_b[sb]), i
        assert np.array_equal(nr_a[sa], nr_b[sb]), i
        np.testing.assert_allclose(ar_a[sa], ar_b[sb], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(small.get("volumes")[ok], full.get("volumes")[ok], rtol=1e-10, atol=1e-15)
