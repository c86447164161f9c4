"""ORACLE (test infrastructure): ctypes binding of liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The library is the CPU FP64
restatement of the reference path (see oracle.cpp for what pins each part).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        dp, i64p, u32p, i32p = (np.ctypeslib.ndpointer(dtype=t, flags="C") for t in
                                (np.float64, np.int64, np.uint32, np.int32))
        L.or_ngrid.argtypes = [C.c_int, C.c_int64]
        L.or_ngrid.restype = C.c_int
        L.or_periodic_ghosts.argtypes = [C.c_int, dp, C.c_int64, C.c_double, dp, i64p]
        L.or_periodic_ghosts.restype = C.c_int64
        L.or_grid_sort.argtypes = [C.c_int, dp, C.c_int64, C.c_int, dp, u32p, i32p, i32p]
        L.or_knn.argtypes = [C.c_int, dp, C.c_int64, C.c_int, i32p, i32p, u32p, C.c_int, u32p]
        L.or_kfeatures.argtypes = [C.c_int, dp, C.c_int64, C.c_int, i32p, i32p,
                                   np.ctypeslib.ndpointer(dtype=np.float32, flags="C")]
        L.or_mesh_build.argtypes = [C.c_int, dp, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_exact_conflict.argtypes = [C.c_int, dp, C.c_uint32, i32p, dp, u32p]
        L.or_exact_conflict.restype = C.c_int
        L.or_mesh_build.restype = C.c_void_p
        L.or_mesh_free.argtypes = [C.c_void_p]
        L.or_mesh_sizes.argtypes = [C.c_void_p, i64p]
        L.or_mesh_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.or_dt_cfl.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, dp, dp]
        L.or_dt_cfl.restype = C.c_double
        L.or_hydro_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, dp, dp, dp]
        L.or_hll.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_int, dp]
        L.or_hll.restype = None
        L.or_mesh_times.argtypes = [C.c_void_p, dp]
        L.or_mesh_times.restype = None
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = None
        L.or_get_threads.argtypes = []
        L.or_get_threads.restype = C.c_int
        L.or_face_velocity.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp]
        L.or_face_velocity.restype = None
        L.or_cell_fluxes.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, dp, dp, dp, C.c_void_p, dp]
        L.or_sim_create.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int64, C.c_int, C.c_double, C.c_int,
                                    C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.or_sim_create.restype = C.c_void_p
        L.or_sim_step.argtypes = [C.c_void_p, C.c_double]
        L.or_sim_step.restype = C.c_double
        L.or_sim_get.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_sim_get.restype = None
        L.or_sim_totals.argtypes = [C.c_void_p, dp]
        L.or_sim_totals.restype = None
        L.or_sim_mesh.argtypes = [C.c_void_p]
        L.or_sim_mesh.restype = C.c_void_p
        L.or_sim_free.argtypes = [C.c_void_p]
        L.or_sim_free.restype = None
        _LIB = L
    return _LIB


def ngrid(dim, n):
    return lib().or_ngrid(dim, n)


def periodic_ghosts(pts, buff):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    cap = n * 3 ** dim
    out = np.empty((cap, dim))
    gid = np.empty(cap, dtype=np.int64)
    ntot = lib().or_periodic_ghosts(dim, pts, n, buff, out, gid)
    return out[:ntot].copy(), gid[: ntot - n].copy()


def grid_sort(pts, N):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    sorted_ = np.empty_like(pts)
    perm = np.empty(n, dtype=np.uint32)
    ncell = N ** dim
    cnt = np.empty(ncell, dtype=np.int32)
    start = np.empty(ncell, dtype=np.int32)
    lib().or_grid_sort(dim, pts, n, N, sorted_, perm, cnt, start)
    return sorted_, perm, cnt, start


def knn(pts, k):
    """Reference knn::init + knn::solve on pts: (N, sorted, perm, knearest[n,k])."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    N = ngrid(dim, n)
    s, perm, cnt, start = grid_sort(pts, N)
    out = np.empty(n * k, dtype=np.uint32)
    lib().or_knn(dim, s, n, N, cnt, start, perm, k, out)
    return N, s, perm, out.reshape(n, k)


def kfeatures(sorted_pts, N):
    """Learned-k features (n, 16) float32 of every seed of a sorted point set
    (rows in grid order, as returned by grid_sort / Mesh.get('sorted_points'))."""
    s = np.ascontiguousarray(sorted_pts, dtype=np.float64)
    n, dim = s.shape
    # the grid tables of an already sorted set: sorting again keeps the order
    s2, perm, cnt, start = grid_sort(s, N)
    assert np.array_equal(s2, s), "points are not in grid order"
    out = np.empty((n, 16), dtype=np.float32)
    lib().or_kfeatures(dim, s, n, N, cnt, start, out)
    return out


# Degeneracy fallback of a mesh build (no reference counterpart): the
# generators of retry t are displaced with the SplitMix64 stream seeded
# DISPLACE_SALT * (t + 1); coordinate d of point i uses counter i * dim + d.
DISPLACE_SALT = 0xD1B54A32D192ED03


def splitmix64(seed, counters):
    """SplitMix64 output at the given counters (uint64 array): mix(seed + (c + 1) * golden)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed % 2 ** 64) + (np.asarray(counters, dtype=np.uint64) + np.uint64(1)) * \
            np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def displace(pts, amp, retry=0, periodic=False):
    """The fallback's displaced generators: x + amp * (2u - 1) per coordinate,
    u = top 53 bits / 2^53; wrapped into [0, 1) (periodic) or clamped."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n, dim = pts.shape
    z = splitmix64(DISPLACE_SALT * (retry + 1), np.arange(n * dim, dtype=np.uint64))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    t = 2.0 * u - 1.0
    x = pts.ravel() + amp * t
    if periodic:
        x = np.where(x < 0.0, x + 1.0, x)
        x = np.where(x >= 1.0, x - 1.0, x)
    else:
        x = np.minimum(np.maximum(x, 0.0), np.nextafter(1.0, 0.0))
    return x.reshape(n, dim)


def set_threads(n):
    """OpenMP threads of the following oracle calls."""
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def face_velocity(ri, rj, wi, wj, fc):
    """Velocity of the face between generators ri, rj moving with wi, wj, at face centroid fc."""
    dim = len(ri)
    out = np.zeros(3)
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in (ri, rj, wi, wj, fc)]
    lib().or_face_velocity(dim, *f, out)
    return out[:dim]


def hll(dim, gamma, UL, UR, n, wall=False):
    """HLL flux (riemann.cpp:7-45) of states (rho, vx, vy, vz, E) across normal n."""
    out = np.zeros(5)
    lib().or_hll(dim, gamma, np.asarray(UL, dtype=np.float64), np.asarray(UR, dtype=np.float64),
                 np.asarray(n, dtype=np.float64), int(wall), out)
    return out


_I8 = ("face_counts", "face_ptr", "ghost_ids")
_I4 = ("neighbor_cell", "status", "kcert")


def exact_conflict(seed, skey, rows):
    """The exact-predicate fallback's conflict test (or_exact_conflict).
    rows: dim + 1 entries, each ("box", axis, +1 | -1) or (point, key); the
    first dim rows are the vertex's planes, the last the new plane.  Returns
    +1 (conflict), -1 (no conflict), 0 (undecided)."""
    seed = np.ascontiguousarray(seed, dtype=np.float64)
    dim = len(seed)
    box = np.zeros(dim + 1, dtype=np.int32)
    pts = np.zeros((dim + 1) * 3, dtype=np.float64)
    keys = np.zeros(dim + 1, dtype=np.uint32)
    for r, row in enumerate(rows):
        if isinstance(row[0], str):
            box[r] = (row[1] + 1) * row[2]
        else:
            pts[3 * r:3 * r + dim] = row[0]
            keys[r] = row[1]
    return lib().or_exact_conflict(dim, seed, int(skey), box, pts, keys)


class Mesh:
    """Oracle mesh in the reference VMesh layout (original order)."""

    BOUNDARY = {"periodic": 0, "reflective": 1, "none": 2}

    def __init__(self, pts, boundary="periodic", buff=0.04, k=32, max_p=64, max_t=96, restart=False,
                 k_max=1024, exact=False, _borrow=None):
        self._owner = None
        if _borrow is not None:
            # a mesh owned by an oracle Sim (kept alive through _owner)
            self._owner, self.h, self.dim = _borrow
        else:
            pts = np.ascontiguousarray(pts, dtype=np.float64)
            n, dim = pts.shape
            self.dim = dim
            self.h = lib().or_mesh_build(dim, pts, n, self.BOUNDARY[boundary], buff, k, max_p, max_t,
                                         int(restart), int(k_max), int(exact))
            if not self.h:
                raise ValueError("oracle mesh build rejected its arguments")
        s = np.zeros(6, dtype=np.int64)
        lib().or_mesh_sizes(self.h, s)
        self.n_seeds, self.n_hydro, self.num_faces, self.n_failed, self.N_grid, self.k = map(int, s)

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_owner", None) is None:
            lib().or_mesh_free(self.h)
        self.h = None

    def times(self):
        """Build stage seconds: ghosts + grid sort, KNN, first clip pass, restarts."""
        t = np.zeros(4)
        lib().or_mesh_times(self.h, t)
        return {"grid_s": t[0], "knn_s": t[1], "clip_s": t[2], "restart_s": t[3]}

    def get(self, name):
        n, F, d = self.n_seeds, self.num_faces, self.dim
        shapes = {"seeds": (n, 3), "volumes": (n,), "face_counts": (n,), "face_ptr": (n,),
                  "neighbor_cell": (F,), "face_normal": (F, 3), "face_area": (F,),
                  "ghost_ids": (n - self.n_hydro if self.n_hydro else 0,), "status": (n,),
                  "kcert": (n,), "cell_centroid": (n, 3), "face_centroid": (F, 3),
                  "sorted_points": (n, d), "permutation": (n,), "knearest": (n, self.k)}
        dt = (np.int64 if name in _I8 else np.int32 if name in _I4 else
              np.uint32 if name in ("permutation", "knearest") else np.float64)
        out = np.zeros(shapes[name], dtype=dt)
        if lib().or_mesh_get(self.h, name.encode(), out.ctypes.data) != 0:
            raise KeyError(name)
        return out

    def dt_cfl(self, gamma, cfl, rho, vel, E):
        return lib().or_dt_cfl(self.h, gamma, cfl, np.ascontiguousarray(rho), np.ascontiguousarray(vel),
                               np.ascontiguousarray(E))

    def hydro_step(self, gamma, dt, rho, vel, E, flux_mode=0):
        """In-place static-mesh HLL step; returns the number of faceless hydro cells."""
        return lib().or_hydro_step(self.h, gamma, dt, flux_mode, rho, vel, E)

    def cell_fluxes(self, gamma, rho, vel, E, flux_mode=1, move=0, wgen=None):
        """Area-integrated face-flux sums of every hydro cell, (n_hydro, 5):
        mass, momentum[3], energy; move=1: faces moving with wgen (or the
        fluid velocity), flux solved in the face frame."""
        out = np.zeros(self.n_hydro * 5)
        w = None if wgen is None else np.ascontiguousarray(wgen, dtype=np.float64)
        bad = lib().or_cell_fluxes(self.h, gamma, flux_mode, move, np.ascontiguousarray(rho, dtype=np.float64),
                                   np.ascontiguousarray(vel, dtype=np.float64),
                                   np.ascontiguousarray(E, dtype=np.float64), None if w is None else w.ctypes.data,
                                   out)
        if bad:
            raise ValueError(f"{bad} box faces on a periodic mesh: the ghost layer is thinner than the cells")
        return out.reshape(-1, 5)


class Sim:
    """Oracle moving-mesh simulation (FP64, extensive conserved state)."""

    MOVE = {"static": 0, "fluid": 1, "regularized": 2}

    def __init__(self, pts, rho, vel, E, boundary="reflective", buff=0.04, k=32, gamma=1.4, cfl=0.3,
                 flux="physical", move="fluid", exact=False):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.n, self.dim = pts.shape
        self.h = lib().or_sim_create(self.dim, pts, np.ascontiguousarray(rho, dtype=np.float64),
                                     np.ascontiguousarray(vel, dtype=np.float64),
                                     np.ascontiguousarray(E, dtype=np.float64), self.n,
                                     Mesh.BOUNDARY[boundary], buff, k, gamma, cfl,
                                     1 if flux == "physical" else 0, self.MOVE[move], int(exact))
        if not self.h:
            raise ValueError("oracle simulation rejected its arguments or its first mesh")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_free(self.h)
            self.h = None

    def step(self, dt_max=0.0):
        dt = lib().or_sim_step(self.h, dt_max)
        if dt < 0:
            raise RuntimeError("oracle rebuild left cells uncertified")
        return dt

    def get(self):
        pts = np.empty((self.n, self.dim))
        rho = np.empty(self.n)
        vel = np.empty((self.n, self.dim))
        E = np.empty(self.n)
        lib().or_sim_get(self.h, pts, rho, vel, E)
        return pts, rho, vel, E

    def totals(self):
        t = np.zeros(5)
        lib().or_sim_totals(self.h, t)
        return t

    def mesh(self):
        """The current mesh (valid until the next step)."""
        return Mesh(None, _borrow=(self, lib().or_sim_mesh(self.h), self.dim))
