"""Device-resident simulation loop (pg_sim_*) on the GPU.

The reference never moves its mesh, so the moving-mesh path is pinned by
(i) move_mesh=0 reducing to the oracle's static HLL step, and a moving step
with zero generator velocity reducing to the static step bit for bit, (ii)
Galilean invariance (uniform flow on a moving mesh stays uniform to
round-off) and (iii) conservation of mass, momentum (periodic) and energy to
round-off.
"""
import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import sod_state, uniform_points

pytestmark = pytest.mark.gpu


def test_static_sim_matches_oracle(ctx):
    dim, n, gamma = 2, 10000, 1.4
    pts = uniform_points(n, dim, 21)
    rho, vel, E = sod_state(pts, gamma)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="periodic", k=16, gamma=gamma, cfl=0.3,
                         flux="physical", move_mesh=False)
    ref = O.Mesh(pts, "periodic", k=16, restart=True)
    r, v, e = rho.copy(), vel.copy(), E.copy()
    for _ in range(3):
        dt_o = ref.dt_cfl(gamma, 0.3, r, v, e)
        dt_g = sim.step()
        # 2D cell areas of the reference carry ~1e-16 absolute round-off from
        # its absolute-coordinate shoelace (cell.cpp:475-498): ~1e-12 relative
        # on cells of area 1e-4, which dt = CFL*sqrt(V/pi)/(c+|v|) inherits.
        assert abs(dt_g - dt_o) <= 1e-11 * dt_o
        ref.hydro_step(gamma, dt_g, r, v, e, flux_mode=1)
    _, rg, vg, eg = sim.get()
    np.testing.assert_allclose(rg, r, rtol=1e-11)
    np.testing.assert_allclose(eg, e, rtol=1e-11)
    np.testing.assert_allclose(vg, v, rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("dim,k", [(2, 16), (3, 32)])
def test_galilean_invariance(ctx, dim, k):
    n = 10000 if dim == 2 else 30000
    pts = uniform_points(n, dim, 22)
    u = np.array([0.31, -0.17, 0.11])[:dim]
    rho = np.full(n, 1.3)
    vel = np.tile(u, (n, 1))
    P = 0.7
    gamma = 5.0 / 3.0
    E = P / (gamma - 1) + 0.5 * rho * np.sum(vel * vel, 1)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="periodic", buff=0.12 if dim == 3 else 0.04,
                         k=k, gamma=gamma, cfl=0.3, flux="physical", move_mesh=True)
    for _ in range(3):
        sim.step()
    p2, r2, v2, e2 = sim.get()
    # The mass flux is exactly zero (identical states at rest in the face
    # frame); what remains is the round-off of the rebuilt cell volumes: the
    # reference's planes use absolute coordinates (cell.cpp:164-168), so
    # vertices carry ~u*|x|^2/h error, ~1e-12 relative volume per rebuild.
    assert np.max(np.abs(r2 - 1.3)) < 1e-10
    assert np.max(np.abs(v2 - u)) < 1e-10
    # generators moved with the flow (periodic wrap)
    assert not np.array_equal(p2, pts)


@pytest.mark.parametrize("boundary", ["periodic", "reflective"])
def test_conservation_moving(ctx, boundary):
    dim, n, gamma = 3, 30000, 1.4
    pts = uniform_points(n, dim, 23)
    rho, vel, E = sod_state(pts, gamma)
    vel[:, 1] = 0.05 * np.sin(2 * np.pi * pts[:, 0])
    E = E + 0.5 * rho * np.sum(vel * vel, 1)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary=boundary, buff=0.12, k=32, gamma=gamma,
                         cfl=0.3, flux="physical", move_mesh=True)
    t0 = sim.totals()
    for _ in range(5):
        sim.step()
    t1 = sim.totals()
    assert abs(t1[0] - t0[0]) < 1e-12 * t0[0]        # mass
    assert abs(t1[4] - t0[4]) < 1e-12 * t0[4]        # energy
    if boundary == "periodic":
        assert np.all(np.abs(t1[1:4] - t0[1:4]) < 1e-12 * (t0[0] + 1.0))  # momentum


def test_thin_ghost_layer_is_an_error(ctx):
    """A periodic hydro cell that reaches the box plane is reported, never
    silently treated as a wall (the reference reads out of bounds there)."""
    from proteusgpu_amd._native import PgError
    pts = uniform_points(20000, 3, 24)
    rho, vel, E = sod_state(pts, 1.4)
    with pytest.raises(PgError, match="ghost layer"):
        api.Simulation(ctx, pts, rho, vel, E, boundary="periodic", buff=0.01, k=32, gamma=1.4)


@pytest.mark.parametrize("dim,n,k", [(2, 10000, 16), (3, 30000, 32)])
def test_zero_velocity_moving_step_equals_static(ctx, dim, n, k):
    """w = 0 reduces the ALE path to the static one: from a state at rest
    (Sod), one step with the generators moving with the fluid (whose
    velocity is still 0) gives the static step's state and time step bit for
    bit, and the generators stay where they are."""
    pts = uniform_points(n, dim, 31)
    rho, vel, E = sod_state(pts, 1.4)
    out = []
    for mode in ("static", "fluid"):
        sim = api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=k, gamma=1.4, cfl=0.3,
                             flux="physical", move_mesh=mode)
        dt = sim.step()
        out.append((dt,) + tuple(sim.get()))
        sim.free()
    (dt_s, x_s, r_s, v_s, e_s), (dt_m, x_m, r_m, v_m, e_m) = out
    assert dt_s == dt_m
    assert np.array_equal(x_m, pts) and np.array_equal(x_s, pts)
    for a, b in ((r_s, r_m), (v_s, v_m), (e_s, e_m)):
        assert np.array_equal(a, b)


def test_failed_step_makes_the_sim_unusable(ctx):
    """A step whose rebuild fails leaves the simulation unusable: two
    generators moving head-on meet exactly (velocities +-0.5, 2^-10 apart,
    the step capped at 2^-10, below every cell's CFL step), the rebuild
    reports the duplicate pair's cells uncertified (PG_ERR_CELLS), and every
    later call that would launch on the simulation returns PG_ERR_FAILED."""
    from proteusgpu_amd._native import PgError
    n, gamma = 1000, 1.4
    pts = uniform_points(n, 2, 71)
    pts[0] = [0.25, 0.5]
    pts[1] = [0.25 + 2.0 ** -10, 0.5]
    rho = np.ones(n)
    vel = np.zeros((n, 2))
    vel[0], vel[1] = [0.5, 0.0], [-0.5, 0.0]
    E = 1e-3 / (gamma - 1) + 0.5 * rho * np.sum(vel * vel, axis=1)
    sim = api.Simulation(ctx, pts, rho, vel, E, boundary="reflective", k=16, gamma=gamma, cfl=0.3,
                         flux="physical", move_mesh="fluid")
    with pytest.raises(PgError) as e:
        sim.step(dt_max=2.0 ** -10)
    assert e.value.code == -5
    for call in (lambda: sim.step(), lambda: sim.hydro_steps(1), sim.get, sim.totals, sim.mesh_handle):
        with pytest.raises(PgError) as e:
            call()
        assert e.value.code == -7 and "earlier step" in str(e.value)
