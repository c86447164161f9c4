"""Seeded, counter-based generators for initial conditions.

Every random number is SplitMix64 of (seed, counter): value ``i`` of stream
``seed`` is ``mix(seed + (i + 1) * 0x9E3779B97F4A7C15)`` mapped to a 53-bit
double in [0, 1), so tests, fixture scripts and benchmarks draw identical
points from a (seed, size) pair.  The reference draws its ICs with unseeded
``np.random`` (create.py:29, create_sod.py:83); these generators replace
those scripts.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed: int, start: int, count: int) -> np.ndarray:
    """Doubles in [0,1) for counters start .. start+count-1 of stream ``seed``."""
    with np.errstate(over="ignore"):
        ctr = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + ctr * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_points(n: int, dim: int, seed: int) -> np.ndarray:
    """n points uniform in [0,1)^dim, row-major (n, dim)."""
    return splitmix_uniform(seed, 0, n * dim).reshape(n, dim)


def sod_state(pts: np.ndarray, gamma: float, x0: float = 0.5):
    """Sod tube along x: (rho,P)=(1,1) for x<x0 else (0.125,0.1), v=0.

    Returns rho[n], vel[n,dim], E[n] with E the total energy per volume
    (create_sod.py:96-121 layout).
    """
    n, dim = pts.shape
    left = pts[:, 0] < x0
    rho = np.where(left, 1.0, 0.125)
    P = np.where(left, 1.0, 0.1)
    vel = np.zeros((n, dim))
    E = P / (gamma - 1.0) + 0.5 * rho * np.sum(vel * vel, axis=1)
    return rho, vel, E


def kelvin_helmholtz_state(pts: np.ndarray, gamma: float = 5.0 / 3.0):
    """Kelvin-Helmholtz shear layer: rho=2, vx=+0.5 for |y-0.5|<0.25, else
    rho=1, vx=-0.5; P=2.5; vy = 0.01 sin(4 pi x)."""
    n, dim = pts.shape
    inner = np.abs(pts[:, 1] - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vel = np.zeros((n, dim))
    vel[:, 0] = np.where(inner, 0.5, -0.5)
    vel[:, 1] = 0.01 * np.sin(4.0 * np.pi * pts[:, 0])
    P = np.full(n, 2.5)
    E = P / (gamma - 1.0) + 0.5 * rho * np.sum(vel * vel, axis=1)
    return rho, vel, E


def box_sizes(nbox: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Box sizes floor(exp(U[ln lo, ln hi])) (log-uniform)."""
    u = splitmix_uniform(seed, 0, nbox)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch_points(sizes, dim: int, seed: int) -> np.ndarray:
    """Concatenated uniform points of boxes of the given sizes (box b uses
    stream seed + b)."""
    return np.concatenate([uniform_points(int(n), dim, seed + b) for b, n in enumerate(sizes)])


def jittered_lattice(m: int, dim: int, seed: int, amp: float = 0.25) -> np.ndarray:
    """m^dim lattice of spacing h = 1/m at cell centres, each point moved by
    U(-amp h, amp h) per axis (a glass-like start)."""
    h = 1.0 / m
    ax = (np.arange(m) + 0.5) * h
    g = np.meshgrid(*([ax] * dim), indexing="ij")
    pts = np.stack([x.ravel() for x in g], axis=1)
    u = splitmix_uniform(seed, 0, pts.size).reshape(pts.shape)
    return pts + (2.0 * u - 1.0) * amp * h


def gaussian_mixture(n: int, ncent: int, sig_lo: float, sig_hi: float, frac_uniform: float, seed: int):
    """n points in [0,1)^3: a fraction frac_uniform uniform, the rest from
    ncent isotropic Gaussians of equal weight with centres U[0,1)^3 and sigma
    log-uniform in [sig_lo, sig_hi].  Draws outside the box are rejected and
    redrawn (Box-Muller normals from the same SplitMix64 streams)."""
    c = splitmix_uniform(seed, 0, 3 * ncent).reshape(ncent, 3)
    su = splitmix_uniform(seed + 1, 0, ncent)
    sig = np.exp(np.log(sig_lo) + su * (np.log(sig_hi) - np.log(sig_lo)))
    nu = int(round(frac_uniform * n))
    out = [uniform_points(nu, 3, seed + 2)]
    ng = n - nu
    which = np.arange(ng) % ncent
    pts = np.empty((ng, 3))
    todo = np.arange(ng)
    ctr = 0
    while todo.size:
        m = todo.size
        u1 = splitmix_uniform(seed + 3, ctr, 3 * m).reshape(m, 3)
        u2 = splitmix_uniform(seed + 4, ctr, 3 * m).reshape(m, 3)
        ctr += 3 * m
        z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
        p = c[which[todo]] + sig[which[todo], None] * z
        ok = np.all((p >= 0.0) & (p < 1.0), axis=1)
        pts[todo[ok]] = p[ok]
        todo = todo[~ok]
    out.append(pts)
    return np.concatenate(out)
