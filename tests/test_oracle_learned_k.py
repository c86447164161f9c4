"""Pins of the learned-k oracle (features, network, labels, decision rule).

The reference has no learned k, so nothing here can be a golden vector of
the reference.  Instead: central finite differences for the FP64 backward
pass, known answers of bfloat16 rounding, closed forms of the features on
lattices and brute-force nearest-neighbour distances, and the certificate
semantics of the labels checked against restarts of the oracle mesh.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import learned_k as L
from oracle import oracle as O
from proteusgpu_amd.generators import uniform_points


def rand_params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(L.NPARAM) * scale).astype(np.float32)


# ------------------------------------------------------------------ network
def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = rand_params(4).astype(np.float64)
    X = rng.standard_normal((37, 16))
    y = rng.integers(0, 6, 37)
    loss, g = L.loss_grad(P, X, y, bf16_ops=False)
    idx = np.concatenate([rng.choice(np.arange(0, 1024), 25, replace=False),      # W1
                          np.arange(1024, 1088, 9),                                # b1
                          rng.choice(np.arange(1088, 5184), 25, replace=False),    # W2
                          np.arange(5184, 5248, 9),                                # b2
                          rng.choice(np.arange(5248, 5632), 25, replace=False),    # W3
                          np.arange(5632, 5638)])                                  # b3
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (L.loss_grad(Pp, X, y, bf16_ops=False)[0] - L.loss_grad(Pm, X, y, bf16_ops=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_loss_of_uniform_logits_is_log6():
    P = np.zeros(L.NPARAM)
    X = np.random.default_rng(0).standard_normal((10, 16))
    loss, g = L.loss_grad(P, X, np.arange(10) % 6, bf16_ops=False)
    assert abs(loss - np.log(6.0)) < 1e-15
    # with zero weights only b3 receives a gradient: mean(p - onehot)
    gb3 = g[5632:]
    expect = np.full(6, 1 / 6) - np.bincount(np.arange(10) % 6, minlength=6) / 10
    np.testing.assert_allclose(gb3, expect, atol=1e-15)
    assert np.all(g[:5632] == 0)


def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 3.0e38,
                     1e-40], dtype=np.float32)
    got = L.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                       # tie -> even (1.0)
    assert got[2] == float(one + 2 * ulp)      # tie -> even (1 + 2^-6)
    assert got[3] == float(one + ulp)          # above the tie rounds up
    assert got[4] == -2.5
    assert np.isfinite(got[5]) or got[5] == np.inf
    assert got[6] >= 0
    assert np.isnan(L.bf16(np.array([np.nan], dtype=np.float32))[0])


def test_bf16_forward_close_to_fp64():
    rng = np.random.default_rng(5)
    P = rand_params(6)
    X = rng.standard_normal((200, 16)).astype(np.float32)
    z_b, _ = L.forward(P, X, True)
    z_f, _ = L.forward(P, X, False)
    # bfloat16 operands: 8 significant bits, errors ~ 2^-8 relative per operand
    assert np.max(np.abs(z_b - z_f)) < 0.05 * np.max(np.abs(z_f))


def test_labels_and_decision():
    assert list(L.labels([1, 16, 17, 24, 25, 33, 48, 49, 64, 65, 96, 97, 300, 0])) == \
        [0, 0, 1, 1, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5]
    big = 40.0
    for c in range(6):
        z = np.full((1, 6), -big)
        z[0, c] = big
        for r in (0.5, 2.8, 10.0):
            assert L.decide(z, r)[0] == c  # a certain class is chosen as is
    # uncertain between 32 and 64 at even odds: the restart penalty favours 64
    z = np.full((1, 6), -big)
    z[0, 2] = z[0, 4] = 0.0
    assert L.decide(z, 2.8)[0] == 4
    # chain costs by hand: start 16, certified k 56 -> passes 16, 32 fail, 64 passes
    assert L.chain_cost(16, 56, 2.0) == (16 + 32) + (32 + 64) + (64 + 112)


def test_adam_first_step_is_signed_lr():
    P = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 1e-3, -1e-6])
    P1, m, v = L.adam(P, g, np.zeros(5), np.zeros(5), 1, 0.01)
    np.testing.assert_allclose(P1, -0.01 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-12)


# ----------------------------------------------------------------- features
def lattice(dim, m, offset=0.5):
    a = 1.0 / m
    ax = (np.arange(m) + offset) * a
    g = np.meshgrid(*([ax] * dim), indexing="ij")
    return np.stack([x.ravel() for x in g], axis=1)


@pytest.mark.parametrize("dim", [2, 3])
def test_features_on_a_lattice_closed_forms(dim):
    # a cubic lattice of spacing 1/m, m^dim points, on the reference grid
    m = 40 if dim == 2 else 22
    pts = lattice(dim, m)
    N = O.ngrid(dim, len(pts))
    s, perm, cnt, start = O.grid_sort(pts, N)
    F = O.kfeatures(s, N)
    cells = np.clip(np.floor(s * N).astype(int), 0, N - 1)
    inner = np.all((cells > 0) & (cells < N - 1), axis=1)
    lam = len(pts) / N ** dim
    # f4: interior cubes are fully inside the grid; corner cells see 2^dim cells
    assert np.all(F[inner, 4] == 1.0)
    corner = np.all((cells == 0), axis=1)
    np.testing.assert_allclose(F[corner, 4], 2 ** dim / 3 ** dim, rtol=1e-7)
    # f0 = log2(1 + own cell count / lam)
    own = cnt[(cells[:, 0] + cells[:, 1] * N + (cells[:, 2] * N * N if dim == 3 else 0))]
    np.testing.assert_allclose(F[:, 0], np.log2(1.0 + own / lam).astype(np.float32), rtol=1e-6)
    # nearest lattice neighbour is at spacing a for every point
    a = 1.0 / m
    np.testing.assert_allclose(F[:, 10], a * N, rtol=1e-6)
    np.testing.assert_allclose(F[:, 15], F[:, 9] / F[:, 8], rtol=1e-6)
    # f12 from the wall distance
    wall = np.min(np.minimum(s, 1 - s), axis=1)
    np.testing.assert_allclose(F[:, 12], np.minimum(wall * N, 3) / 3, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("dim,n", [(2, 20000), (3, 30000)])
def test_feature_distances_against_ckdtree(dim, n):
    pts = uniform_points(n, dim, 71)
    N = O.ngrid(dim, n)
    s, perm, cnt, start = O.grid_sort(pts, N)
    F = O.kfeatures(s, N)
    d, _ = cKDTree(s).query(s, k=9)
    # wherever the 8th neighbour lies within the cube's reach (distance to
    # the cube's boundary), the cube search finds the true 4th / 8th / 1st
    # neighbour distances
    cells = np.clip(np.floor(s * N).astype(int), 0, N - 1)
    reach = (np.min(np.minimum(s * N - cells, cells + 1 - s * N), axis=1) + 1.0) / N
    ok = (d[:, 8] < reach) & np.all((cells > 0) & (cells < N - 1), axis=1)
    assert ok.sum() > n // 20
    np.testing.assert_allclose(F[ok, 8], (d[ok, 4] * N).astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(F[ok, 9], (d[ok, 8] * N).astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(F[ok, 10], (d[ok, 1] * N).astype(np.float32), rtol=1e-6)
    # class counts: the m = 1 class of an interior cell is the mean of its 2*dim face cells
    c = cells[ok][0]
    offs = [np.eye(dim, dtype=int)[i] * sgn for i in range(dim) for sgn in (-1, 1)]
    cid = lambda v: v[0] + v[1] * N + (v[2] * N * N if dim == 3 else 0)
    face_mean = np.mean([cnt[cid(c + o)] for o in offs])
    lam = n / N ** dim
    np.testing.assert_allclose(F[ok][0, 1], np.float32(np.log2(1.0 + face_mean / lam)), rtol=1e-6)


def test_labels_are_the_minimal_certified_k():
    # a seed labelled with bucket b certifies with a row of k_b neighbours but
    # not with the bucket below (when kcert > 16)
    pts = uniform_points(3000, 3, 12)
    full = O.Mesh(pts, "reflective", k=32, restart=True)
    kc = full.get("kcert")
    y = L.labels(kc)
    for c, k in enumerate(L.BUCKETS[:-1]):
        m = O.Mesh(pts, "reflective", k=k, restart=False)
        st = m.get("status")
        ok = st == 4
        assert np.all(ok[y <= c]), f"bucket {k}: a seed labelled at or below it failed"
        assert not np.any(ok[(y > c) & (kc > k)]), f"bucket {k}: a seed labelled above it passed"
