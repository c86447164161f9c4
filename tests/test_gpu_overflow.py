"""Seeds whose cell outgrows the fast clip kernel's vertex slots.

A seed wrapped by a close shell of 55 points (radii jittered by 10 %) has a
cell of up to about 48 faces, hence up to 2F - 4 = 92 vertices (a simple 3D
polytope): more than the 64 (or 48) slots of the fast kernel and within the
128 of the large kernel and the reference's capacity of 96.  Such seeds take the overflow path (one
128-slot pass per build, at a row of at least 256); the mesh must still
equal the oracle's, which has no slot limit, and the batched build must
equal the host-driven one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import sod_state, uniform_points

from test_gpu_parity import _cmp_mesh

pytestmark = pytest.mark.gpu


def shell_points(n_bg=13000, n_shell=12, per_shell=55, r=0.015, seed=31):
    rng = np.random.default_rng(seed)
    parts = [uniform_points(n_bg, 3, seed)]
    for c in rng.uniform(0.1, 0.9, size=(n_shell, 3)):
        u = rng.normal(size=(per_shell, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        # radii jittered: points on one sphere would make ties of the
        # in-sphere tests (degenerate cells) instead of large ones
        rr = r * rng.uniform(0.9, 1.1, size=(per_shell, 1))
        parts.append(c[None, :])
        parts.append(c + rr * u)
    return np.ascontiguousarray(np.concatenate(parts))


@pytest.mark.parametrize("lanes,slots", [(16, 64), (16, 48)])
def test_overflowed_cells_equal_oracle(ctx, lanes, slots):
    pts = shell_points()
    print(f"overflow shells n={len(pts)} shape {lanes}x{slots}", flush=True)
    ref = O.Mesh(pts, "periodic", k=32, restart=True)
    # the shell centres' cells hold more vertices than the fast kernel's
    # slots (V = 2F - 4), so they can only come from the 128-slot pass
    assert 2 * int(np.max(ref.get("face_counts"))) - 4 > 64
    api.set_clip_config(lanes, slots)
    try:
        m = api.Mesh(ctx, pts, boundary="periodic", k=32, restart=True)
        g = m.copy_out()
    finally:
        api.set_clip_config(16, 64)
    assert m.n_failed == ref.n_failed == 0
    _cmp_mesh(g, ref)


def test_overflowed_cells_batched_equal_host(ctx):
    pts = shell_points(seed=32)
    print(f"overflow shells batched n={len(pts)}", flush=True)
    rho, vel, E = sod_state(pts, 1.4)
    kw = dict(boundary="periodic", buff=0.12, k=32, gamma=1.4, cfl=0.3, flux="physical", move_mesh=True)
    B = api.Batch(ctx, np.array([len(pts)]), pts, rho, vel, E, **kw)
    B.step()
    bx, br, bv, bE = B.get()
    B.free()
    sim = api.Simulation(ctx, pts, rho, vel, E, **kw)
    sim.step()
    ax, ar, av, aE = sim.get()
    sim.free()
    assert np.array_equal(bx, ax) and np.array_equal(br, ar)
    assert np.array_equal(bv, av) and np.array_equal(bE, aE)
