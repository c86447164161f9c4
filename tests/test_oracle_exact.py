"""The oracle's exact-predicate fallback, pinned by rational arithmetic and
by closed forms.

The reference marks a cell needs_exact_predicates when its filtered conflict
determinant is below eps (cell.cpp:216-218) and leaves it failed
(voronoi.cpp:104-110).  With exact=True the oracle decides such tests
exactly, ties broken by Simulation of Simplicity on power weights.  Here:

* or_exact_conflict against the same determinant over Python Fractions,
  built in global coordinates from the plane definitions (the oracle works
  on scaled integers; the GPU on seed-relative ones), on generic, degenerate
  (lattice) and wide-exponent inputs, with box planes;
* the symbolic perturbation against its definition: the leading
  coefficient of det(planes with w += (w_s - w_x) / 2) for weights
  w = eps^rank(key);
* meshes of exact lattices: every cell certified, cells are the lattice's
  squares / cubes (volume h^dim), faces come in (i, j) / (j, i) pairs, and
  every face of positive area is a lattice face of area h^(dim-1).
"""
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import oracle as O

LO, HI = -1e-14, 1.0 + 1e-14


def frac_det(M):
    """Determinant by fraction-exact Gaussian elimination."""
    M = [list(r) for r in M]
    n = len(M)
    det = F(1)
    for c in range(n):
        p = next((r for r in range(c, n) if M[r][c] != 0), None)
        if p is None:
            return F(0)
        if p != c:
            M[c], M[p] = M[p], M[c]
            det = -det
        det *= M[c][c]
        for r in range(c + 1, n):
            f = M[r][c] / M[c][c]
            for j in range(c, n):
                M[r][j] -= f * M[c][j]
    return det


def plane_rows(seed, rows):
    dim = len(seed)
    s = [F(x) for x in seed]
    out = []
    for row in rows:
        if isinstance(row[0], str):
            _, ax, sg = row
            c = F(LO) if sg > 0 else F(HI)
            out.append([F(sg) if d == ax else F(0) for d in range(dim)] + [-sg * c])
        else:
            b = [F(x) for x in row[0]]
            n = [s[d] - b[d] for d in range(dim)]
            w = (sum(x * x for x in b) - sum(x * x for x in s)) / 2
            out.append(n + [w])
    return out


def rational_conflict(seed, skey, rows):
    """Sign of the determinant, else of its leading symbolic term."""
    dim = len(seed)
    M = plane_rows(seed, rows)
    d = frac_det(M)
    if d != 0:
        return 1 if d > 0 else -1
    cand = []
    tot = F(0)
    for k, row in enumerate(rows):
        if isinstance(row[0], str):
            continue
        minor = [[M[i][j] for j in range(dim)] for i in range(dim + 1) if i != k]
        ck = (-1) ** (k + dim) * frac_det(minor)
        tot += ck
        cand.append((row[1], -ck))  # the w entry of row k gains -w_k / 2
    cand.append((skey, tot))  # every bisector's w gains +w_s / 2
    for _, c in sorted(cand):
        if c != 0:
            return 1 if c > 0 else -1
    return 0


def random_case(rng, dim, kind):
    seed = rng.random(dim)
    rows = []
    keys = rng.permutation(1000)[: dim + 2]
    for r in range(dim + 1):
        if kind == "box" and r < dim and rng.random() < 0.5:
            rows.append(("box", r, 1 if rng.random() < 0.5 else -1))
            continue
        if kind == "lattice":
            p = seed + rng.integers(-2, 3, dim) / 16.0
        elif kind == "wide":
            p = seed + rng.standard_normal(dim) * 10.0 ** rng.integers(-25, 0)
        else:
            p = seed + rng.standard_normal(dim) * 0.05
        rows.append((p, int(keys[r])))
    if kind == "lattice":
        seed = np.round(seed * 16) / 16
        rows = [r if isinstance(r[0], str) else (np.round(np.array(r[0]) * 16) / 16, r[1]) for r in rows]
    return seed, int(keys[-1]), rows


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["generic", "lattice", "wide", "box"])
def test_exact_conflict_equals_rational_arithmetic(dim, kind):
    rng = np.random.default_rng(7 + dim + len(kind))
    n_zero = 0
    for _ in range(300):
        seed, skey, rows = random_case(rng, dim, kind)
        # a duplicate of the seed has no bisector
        if any(not isinstance(r[0], str) and np.array_equal(r[0], seed) for r in rows):
            continue
        want = rational_conflict(seed, skey, rows)
        got = O.exact_conflict(seed, skey, rows)
        assert got == want, (seed, skey, rows)
        n_zero += frac_det(plane_rows(seed, rows)) == 0
    if kind == "lattice":
        assert n_zero > 10  # the symbolic perturbation is exercised


def test_exact_conflict_orientation_and_ties():
    """2D: the vertex (0.625, 0.625) of the bisectors of (0.5, 0.5) with
    (0.5, 0.75) and (0.75, 0.5), in the orientation the cell stores (Cramer
    denominator < 0, cell.cpp:346-364), against points inside, outside and on
    the circle through the four points; on the circle the smaller key wins."""
    s = [0.5, 0.5]
    a, b = ((0.5, 0.75), 2), ((0.75, 0.5), 1)
    assert O.exact_conflict(s, 0, [a, b, ((0.7, 0.7), 3)]) == 1
    assert O.exact_conflict(s, 0, [a, b, ((0.76, 0.76), 3)]) == -1
    on = ((0.75, 0.75), 3)
    r0, r5 = O.exact_conflict(s, 0, [a, b, on]), O.exact_conflict(s, 5, [a, b, on])
    assert {r0, r5} == {1, -1}
    assert r0 == rational_conflict(s, 0, [a, b, on]) and r5 == rational_conflict(s, 5, [a, b, on])
    # subnormal input: undecided
    assert O.exact_conflict([5e-324, 0.5], 0, [a, b, on]) == 0


def lattice(dim, m):
    g = (np.arange(m) + 0.5) / m
    pts = np.stack([x.ravel() for x in np.meshgrid(*([g] * dim), indexing="ij")], 1)
    return pts


def _faces(mesh, n):
    fc, fp = mesh.get("face_counts"), mesh.get("face_ptr")
    nb, ar, nr = mesh.get("neighbor_cell"), mesh.get("face_area"), mesh.get("face_normal")
    return [(nb[fp[i]:fp[i] + fc[i]], ar[fp[i]:fp[i] + fc[i]], nr[fp[i]:fp[i] + fc[i]]) for i in range(n)]


@pytest.mark.parametrize("dim,m,k,boundary", [(2, 12, 16, "reflective"), (3, 6, 32, "reflective"),
                                               (2, 16, 16, "periodic"), (3, 10, 32, "periodic")])
def test_exact_lattice_mesh_closed_forms(dim, m, k, boundary):
    """Periodic lattices are scaled into the ghost box, which rounds them off
    the lattice by ~1e-17: there the reference's eps filter decides some
    signs wrongly (|det| a little above its eps, opposite exact sign), and
    only the exact mode's bound (XTOL) sends them to the exact test."""
    pts = lattice(dim, m)
    n = len(pts)
    buff = 0.12 if dim == 3 else 0.06
    plain = O.Mesh(pts, boundary, buff=buff, k=k, restart=True)
    assert plain.n_failed > 0.5 * n  # the reference's predicate fails here
    mesh = O.Mesh(pts, boundary, buff=buff, k=k, restart=True, exact=True)
    st = mesh.get("status")
    assert np.all(st == 4)
    h = 1.0 / m
    np.testing.assert_allclose(mesh.get("volumes")[:n], h ** dim, rtol=2e-12)
    seeds = mesh.get("seeds")[:, :dim]
    faces = _faces(mesh, len(st))
    pairs = set()
    for i, (nb, ar, nr) in enumerate(faces):
        for j, a, nv in zip(nb, ar, nr):
            if j < 0:
                continue
            pairs.add((i, int(j)))
            if i < n and a > 1e-9 * h ** (dim - 1):
                # a lattice face: axis-aligned neighbour at distance h
                d = seeds[j] - seeds[i]
                assert np.isclose(np.linalg.norm(d), h, rtol=1e-12)
                np.testing.assert_allclose(a, h ** (dim - 1), rtol=2e-12)
    # every face is listed by both of its cells
    assert all((j, i) in pairs for (i, j) in pairs)
    # keys break ties the same way in every cell: a build at a longer row
    # gives the same mesh
    longer = O.Mesh(pts, boundary, buff=buff, k=2 * k, restart=True, exact=True)
    for key in ("status", "face_counts", "neighbor_cell"):
        assert np.array_equal(mesh.get(key), longer.get(key)), key


def test_exact_is_inert_on_generic_input():
    from proteusgpu_amd.generators import uniform_points
    pts = uniform_points(3000, 3, 5)
    a = O.Mesh(pts, "periodic", buff=0.12, k=32, restart=True)
    b = O.Mesh(pts, "periodic", buff=0.12, k=32, restart=True, exact=True)
    for key in ("status", "kcert", "face_counts", "neighbor_cell", "face_normal", "face_area", "volumes"):
        assert np.array_equal(a.get(key), b.get(key)), key
