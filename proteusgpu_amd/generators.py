"""Seeded, counter-based generators for initial conditions.

Every random number is SplitMix64 of (seed, counter): value ``i`` of stream
``seed`` is ``mix(seed + (i + 1) * 0x9E3779B97F4A7C15)`` mapped to a 53-bit
double in [0, 1).  The same mapping is implemented in C++
(``proteusgpu_amd/csrc/gen.h``) so host, device and fixture scripts produce
identical points.  The reference draws its ICs with unseeded ``np.random``
(create.py:29, create_sod.py:83); these generators replace those scripts.
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
