"""Python view of the C ABI: context, KNN problem, Voronoi mesh, hydro state.

Names and argument meaning follow the reference's C++ interfaces
(src/knn/knn.h:36-62, src/voronoi/voronoi.h:64-81, src/voronoi/periodic_mesh.h:22,
src/hydro/finite_volume_solver.h:20-27); every call goes through libpgpu.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import (BOUNDARY, FLUX, KMODEL_NCLASS, KMODEL_NFEAT, KMODEL_NPARAM, MeshInfo, MeshParams,
                      VMeshView, check, lib)


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _state_arrays(rho, vel, E, n, dim):
    """rho (n,), vel (n, dim), E (n,) as contiguous float64; ValueError on a
    shape mismatch (the C side reads exactly n cells from each array)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    E = np.ascontiguousarray(E, dtype=np.float64)
    if rho.shape != (n,) or E.shape != (n,) or vel.shape != (n, dim):
        raise ValueError(f"state shapes rho {rho.shape}, vel {vel.shape}, E {E.shape} do not match "
                         f"n = {n} cells in {dim}D: expected ({n},), ({n}, {dim}), ({n},)")
    return rho, vel, E


def set_clip_config(lanes: int = 16, slots: int = 64):
    """Shape of the fast clip kernel for the following builds (timing only)."""
    check(lib().pg_set_clip_config(lanes, slots))


def get_clip_config():
    lanes, slots = C.c_int(), C.c_int()
    check(lib().pg_get_clip_config(C.byref(lanes), C.byref(slots)))
    return lanes.value, slots.value


def set_clip_resume(on: bool = True):
    """Certificate restarts continue the saved cell (True) or clip again from
    the first neighbour (False), for the following builds; same results."""
    check(lib().pg_set_clip_resume(int(bool(on))))


def get_clip_resume() -> bool:
    on = C.c_int()
    check(lib().pg_get_clip_resume(C.byref(on)))
    return bool(on.value)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().pg_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU and one stream (replaces the reference's gpu* shim, allvars.h:92-129)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().pg_ctx_create(device, C.byref(h)))
        self.h = h

    def sync(self):
        check(lib().pg_ctx_sync(self.h))

    def stream_copy_probe(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Measured device copy bandwidth in GB/s (reads + writes)."""
        ms = C.c_double()
        check(lib().pg_stream_copy_probe(self.h, nbytes, reps, C.byref(ms)))
        return 2.0 * nbytes / (ms.value * 1e-3) / 1e9

    def close(self):
        if getattr(self, "h", None):
            lib().pg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class KnnProblem:
    """knn::init / knn::solve / get_points / get_knearest / get_permutation."""

    def __init__(self, ctx: Context, pts: np.ndarray):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx, self.n, self.dim = ctx, pts.shape[0], pts.shape[1]
        h = C.c_void_p()
        check(lib().pg_knn_init(ctx.h, self.dim, _ptr(pts), self.n, C.byref(h)))
        self.h = h
        self.k = 0
        n64, N, k = C.c_int64(), C.c_int(), C.c_int()
        check(lib().pg_knn_info(h, C.byref(n64), C.byref(N), C.byref(k)))
        self.N_grid = N.value

    def solve(self, k: int):
        check(lib().pg_knn_solve(self.ctx.h, self.h, k))
        self.k = k

    def get_points(self):
        out = np.empty((self.n, self.dim))
        check(lib().pg_knn_get_points(self.ctx.h, self.h, _ptr(out)))
        return out

    def get_knearest(self):
        out = np.empty((self.n, self.k), dtype=np.uint32)
        check(lib().pg_knn_get_knearest(self.ctx.h, self.h, _ptr(out)))
        return out

    def get_permutation(self):
        out = np.empty(self.n, dtype=np.uint32)
        check(lib().pg_knn_get_permutation(self.ctx.h, self.h, _ptr(out)))
        return out

    def free(self):
        if getattr(self, "h", None):
            lib().pg_knn_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()


def mesh_params(dim=3, boundary="periodic", buff=0.04, k=32, max_p=64, max_t=96, restart=True,
                k_max=1024, perturb=0.0, exact=False) -> MeshParams:
    p = MeshParams()
    lib().pg_mesh_params_default(C.byref(p))
    p.dim, p.boundary, p.buff, p.k = dim, BOUNDARY[boundary], buff, k
    p.max_p, p.max_t, p.restart, p.k_max = max_p, max_t, int(restart), k_max
    p.perturb = perturb
    p.exact = int(exact)
    return p


class KModel:
    """Learned neighbour count (no reference counterpart): a 16-64-64-6
    classifier of each seed's k bucket, trained and evaluated on the GPU."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().pg_kmodel_create(ctx.h, seed, C.byref(h)))
        self.h = h

    @property
    def params(self) -> np.ndarray:
        out = np.empty(KMODEL_NPARAM, dtype=np.float32)
        check(lib().pg_kmodel_get_params(self.ctx.h, self.h, _ptr(out)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (KMODEL_NPARAM,)
        check(lib().pg_kmodel_set_params(self.ctx.h, self.h, _ptr(p)))

    def set_cost(self, a0: float):
        check(lib().pg_kmodel_set_cost(self.h, a0))

    def predict(self, feat, logits=False):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        n = feat.shape[0]
        cls = np.empty(n, dtype=np.int32)
        lg = np.empty((n, KMODEL_NCLASS), dtype=np.float32) if logits else None
        check(lib().pg_kmodel_predict(self.ctx.h, self.h, _ptr(feat), n, _ptr(cls), _ptr(lg)))
        return (cls, lg) if logits else cls

    def grad(self, feat, cls):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        g = np.empty(KMODEL_NPARAM, dtype=np.float32)
        loss = C.c_double()
        check(lib().pg_kmodel_grad(self.ctx.h, self.h, _ptr(feat), _ptr(cls), feat.shape[0], _ptr(g),
                                   C.byref(loss)))
        return loss.value, g

    def train_step(self, feat, cls, lr):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        loss = C.c_double()
        check(lib().pg_kmodel_train_step(self.ctx.h, self.h, _ptr(feat), _ptr(cls), feat.shape[0], lr,
                                         C.byref(loss)))
        return loss.value

    def train_mesh(self, mesh, steps=400, batch=4096, lr=3e-3, seed=7):
        """Adam steps on the seeds of a built mesh (Mesh or Simulation);
        returns the last step's mean loss."""
        h = mesh.mesh_handle() if hasattr(mesh, "mesh_handle") else mesh.h
        loss = C.c_double()
        check(lib().pg_kmodel_train_mesh(self.ctx.h, self.h, h, steps, batch, lr, seed, C.byref(loss)))
        return loss.value

    def free(self):
        if getattr(self, "h", None):
            lib().pg_kmodel_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()


class Mesh:
    """voronoi::compute_mesh / compute_periodic_mesh and the VMesh copy-out.
    kmodel: a trained KModel picks each seed's first k (learned k).
    perturb > 0: degeneracy fallback (builds that leave cells with
    degenerate predicates are repeated on deterministically displaced
    generators; `displacement` reports the amplitude used).
    exact: cells with degenerate predicates are clipped again with the
    exact predicate and symbolic tie-breaking (`n_exact` counts them)."""

    def __init__(self, ctx: Context, pts: np.ndarray, boundary="periodic", buff=0.04, k=32, max_p=64,
                 max_t=96, restart=True, k_max=1024, kmodel: KModel | None = None, perturb=0.0, exact=False):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx, self.dim, self.k = ctx, pts.shape[1], k
        prm = mesh_params(self.dim, boundary, buff, k, max_p, max_t, restart, k_max, perturb, exact)
        h = C.c_void_p()
        km = kmodel.h if kmodel is not None else None
        check(lib().pg_mesh_compute_k(ctx.h, C.byref(prm), km, _ptr(pts), pts.shape[0], C.byref(h)))
        self.h = h
        self._load_info()

    @classmethod
    def from_knn(cls, ctx: Context, knn: "KnnProblem", max_p=64, max_t=96):
        """voronoi::compute_cells: the cells of a solved KNN problem (no
        restarts, box [0,1]^dim), in the problem's sorted order.  Returns
        (mesh, status)."""
        m = cls.__new__(cls)
        m.ctx, m.dim, m.k = ctx, knn.dim, knn.k
        status = np.empty(knn.n, np.int32)
        h = C.c_void_p()
        check(lib().pg_voronoi_compute(ctx.h, knn.h, max_p, max_t, C.byref(h), _ptr(status)))
        m.h = h
        m._load_info()
        return m, status

    def unpermute(self):
        """voronoi::unpermute_vmesh for a mesh of from_knn: later copy-outs
        are in input order."""
        check(lib().pg_mesh_unpermute(self.ctx.h, self.h))

    def _load_info(self):
        info = MeshInfo()
        check(lib().pg_mesh_info(self.h, C.byref(info)))
        self.info = info
        self.n_seeds, self.n_hydro, self.num_faces = info.n_seeds, info.n_hydro, info.num_faces
        self.n_failed, self.n_restarted, self.N_grid, self.k_used = (
            info.n_failed, info.n_restarted, info.N_grid, info.k_used)
        self.row_entries = info.row_entries
        self.n_class = np.array(info.n_class[:], dtype=np.int64)
        self.n_degenerate, self.displacement = info.n_degenerate, info.displacement
        self.n_exact = info.n_exact
        self.n_ball = info.n_ball

    def features(self):
        """Learned-k features of every seed, original order (n_seeds, 16)."""
        out = np.empty((self.n_seeds, KMODEL_NFEAT), dtype=np.float32)
        check(lib().pg_mesh_features(self.ctx.h, self.h, _ptr(out)))
        return out

    def knn(self):
        """The build's KNN problem as knn::get_points / get_knearest /
        get_permutation return it (fixed-k builds): sorted points (scaled
        into the unit box for periodic meshes), rows of length k in
        sorted-index space, sorted -> input permutation."""
        n = self.n_seeds
        pts = np.empty((n, self.dim))
        rows = np.empty((n, self.k), np.uint32)
        perm = np.empty(n, np.uint32)
        check(lib().pg_mesh_get_knn(self.ctx.h, self.h, _ptr(pts), _ptr(rows), _ptr(perm)))
        return pts, rows, perm

    def copy_out(self):
        """Reference VMesh arrays in original order (dict of numpy arrays)."""
        n, F = self.n_seeds, self.num_faces
        ng = n - self.n_hydro if self.n_hydro else 0
        out = {
            "cell_ids": np.empty(n, np.int64), "seeds": np.empty((n, 3)), "volumes": np.empty(n),
            "face_counts": np.empty(n, np.int64), "face_ptr": np.empty(n, np.int64),
            "neighbor_cell": np.empty(F, np.int32), "face_normal": np.empty((F, 3)),
            "face_area": np.empty(F), "ghost_ids": np.empty(max(ng, 0), np.int64),
            "status": np.empty(n, np.int32), "kcert": np.empty(n, np.int32),
            "cell_centroid": np.empty((n, 3)), "face_centroid": np.empty((F, 3)),
        }
        v = VMeshView()
        for name, arr in out.items():
            setattr(v, name, arr.ctypes.data if arr.size else None)
        check(lib().pg_mesh_copy_out(self.ctx.h, self.h, C.byref(v)))
        return out

    def free(self):
        if getattr(self, "h", None):
            lib().pg_mesh_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()


def hll_flux(ctx: Context, gamma, left, right, normal, wall=None, flux="reference"):
    """hydro::riemann_hll for many faces: states (n, dim+2) as (rho, v, E),
    stored normals (n, 3), optional wall flags; fluxes (n, dim+2)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    normal = np.ascontiguousarray(normal, dtype=np.float64)
    if left.ndim != 2 or left.shape[1] not in (4, 5):
        raise ValueError(f"left states must be (n, dim + 2) with dim 2 or 3, got {left.shape}")
    n, w = left.shape
    if right.shape != (n, w) or normal.shape != (n, 3):
        raise ValueError(f"right {right.shape} and normal {normal.shape} must be ({n}, {w}) and ({n}, 3)")
    wl = None if wall is None else np.ascontiguousarray(wall, dtype=np.uint8)
    if wl is not None and wl.shape != (n,):
        raise ValueError(f"wall flags must be ({n},), got {wl.shape}")
    out = np.empty((n, w))
    check(lib().pg_hll_flux(ctx.h, w - 2, gamma, n, _ptr(left), _ptr(right), _ptr(normal), _ptr(wl),
                            FLUX[flux], _ptr(out)))
    return out


class HydroState:
    """hydro::init / dt_CFL / hydro_step / free_prim on a static mesh."""

    def __init__(self, ctx: Context, mesh: Mesh, rho, vel, E):
        self.ctx, self.mesh = ctx, mesh
        self.n, self.dim = mesh.n_hydro, mesh.dim
        rho, vel, E = _state_arrays(rho, vel, E, self.n, self.dim)
        h = C.c_void_p()
        check(lib().pg_state_create(ctx.h, mesh.h, _ptr(rho), _ptr(vel), _ptr(E), C.byref(h)))
        self.h = h

    def dt_cfl(self, gamma, cfl):
        dt = C.c_double()
        check(lib().pg_dt_cfl(self.ctx.h, self.mesh.h, self.h, gamma, cfl, C.byref(dt)))
        return dt.value

    def step(self, gamma, dt, flux="reference"):
        check(lib().pg_hydro_step(self.ctx.h, self.mesh.h, self.h, gamma, dt, FLUX[flux]))

    def copy_out(self):
        rho = np.empty(self.n)
        vel = np.empty((self.n, self.dim))
        E = np.empty(self.n)
        check(lib().pg_state_copy_out(self.ctx.h, self.h, _ptr(rho), _ptr(vel), _ptr(E)))
        return rho, vel, E

    def free(self):
        if getattr(self, "h", None):
            lib().pg_state_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()


MOVE = {"static": 0, "fluid": 1, "regularized": 2}


def move_mode(move_mesh):
    """PG_MOVE_* from False/True (static / with the fluid) or a name:
    "static", "fluid", "regularized" (fluid velocity plus a drift towards the
    cell centroid that keeps cells round)."""
    if isinstance(move_mesh, str):
        return MOVE[move_mesh]
    return int(bool(move_mesh))


class Simulation:
    """Device-resident moving-mesh loop (main.cpp:50-63 kept in HBM).

    step(): primitives, CFL dt, HLL update, drift (move_mesh) and one full
    Voronoi rebuild.  hydro_steps(n): n static-mesh HLL updates.
    """

    def __init__(self, ctx: Context, pts, rho, vel, E, boundary="reflective", buff=0.04, k=32,
                 gamma=5.0 / 3.0, cfl=0.4, flux="physical", move_mesh=True, max_p=64, max_t=96,
                 restart=True, k_max=1024, perturb=0.0, exact=False):
        from ._native import SimParams
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx, self.n, self.dim = ctx, pts.shape[0], pts.shape[1]
        p = SimParams()
        p.mesh = mesh_params(self.dim, boundary, buff, k, max_p, max_t, restart, k_max, perturb, exact)
        p.gamma, p.cfl, p.flux_mode, p.move_mesh = gamma, cfl, FLUX[flux], move_mode(move_mesh)
        arrs = _state_arrays(rho, vel, E, self.n, self.dim)
        h = C.c_void_p()
        check(lib().pg_sim_create(ctx.h, C.byref(p), _ptr(pts), _ptr(arrs[0]), _ptr(arrs[1]),
                                  _ptr(arrs[2]), self.n, C.byref(h)))
        self.h = h

    def step(self, dt_max: float = 0.0) -> float:
        dt = C.c_double()
        check(lib().pg_sim_step(self.ctx.h, self.h, dt_max, C.byref(dt)))
        return dt.value

    def hydro_steps(self, n: int, dt_max: float = 0.0) -> float:
        t = C.c_double()
        check(lib().pg_sim_hydro_steps(self.ctx.h, self.h, n, dt_max, C.byref(t)))
        return t.value

    def times(self) -> dict:
        from ._native import SimTimes
        t = SimTimes()
        check(lib().pg_sim_times_last(self.h, C.byref(t)))
        return t.as_dict()

    def get(self):
        pts = np.empty((self.n, self.dim))
        rho = np.empty(self.n)
        vel = np.empty((self.n, self.dim))
        E = np.empty(self.n)
        check(lib().pg_sim_get(self.ctx.h, self.h, _ptr(pts), _ptr(rho), _ptr(vel), _ptr(E)))
        return pts, rho, vel, E

    def totals(self):
        t = np.zeros(5)
        check(lib().pg_sim_totals(self.ctx.h, self.h, _ptr(t)))
        return t

    def set_kmodel(self, km: KModel | None):
        """Learned k for the following rebuilds (None: fixed k)."""
        check(lib().pg_sim_set_kmodel(self.h, km.h if km is not None else None))
        self._km = km

    def mesh_info(self) -> dict:
        """Counters of the current mesh (pg_mesh_info)."""
        info = MeshInfo()
        check(lib().pg_mesh_info(self.mesh_handle(), C.byref(info)))
        d = {name: getattr(info, name) for name, _ in MeshInfo._fields_}
        d["n_class"] = [int(x) for x in info.n_class]
        return d

    def mesh_handle(self):
        hm = C.c_void_p()
        check(lib().pg_sim_mesh(self.h, C.byref(hm)))
        return hm

    def mesh_copy_out(self):
        """Reference-layout VMesh of the current mesh."""
        hm = C.c_void_p()
        check(lib().pg_sim_mesh(self.h, C.byref(hm)))
        m = Mesh.__new__(Mesh)
        m.ctx, m.dim, m.h = self.ctx, self.dim, hm
        info = MeshInfo()
        check(lib().pg_mesh_info(hm, C.byref(info)))
        m.n_seeds, m.n_hydro, m.num_faces = info.n_seeds, info.n_hydro, info.num_faces
        try:
            return m.copy_out()
        finally:
            m.h = None  # owned by the simulation

    def free(self):
        if getattr(self, "h", None):
            lib().pg_sim_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()


class Batch:
    """Many independent boxes advanced by one captured step (pg_batch_*).

    sizes: points per box; pts/rho/vel/E concatenated box-major, each box in
    its own unit box.  step() advances every box by its own CFL step."""

    def __init__(self, ctx: Context, sizes, pts, rho, vel, E, boundary="periodic", buff=0.04, k=16,
                 gamma=5.0 / 3.0, cfl=0.3, flux="physical", move_mesh=True, max_p=64, max_t=96, restart=True,
                 k_max=256):
        from ._native import SimParams
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx, self.dim = ctx, pts.shape[1]
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self.n = int(self.sizes.sum())
        if pts.ndim != 2 or pts.shape[0] != self.n:
            raise ValueError(f"points {pts.shape} do not match the {self.n} cells of the boxes")
        p = SimParams()
        p.mesh = mesh_params(self.dim, boundary, buff, k, max_p, max_t, restart, k_max)
        p.gamma, p.cfl, p.flux_mode, p.move_mesh = gamma, cfl, FLUX[flux], move_mode(move_mesh)
        arrs = _state_arrays(rho, vel, E, self.n, self.dim)
        h = C.c_void_p()
        check(lib().pg_batch_create(ctx.h, C.byref(p), len(self.sizes), _ptr(self.sizes), _ptr(pts),
                                    _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), C.byref(h)))
        self.h = h

    def step(self) -> float:
        """One step of every box; returns its GPU time in ms."""
        ms = C.c_double()
        check(lib().pg_batch_step(self.ctx.h, self.h, C.byref(ms)))
        return ms.value

    @property
    def tuning(self) -> dict:
        from ._native import BatchTuning
        t = BatchTuning()
        check(lib().pg_batch_get_tuning(self.h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def set_tuning(self, clip_slots=64, knn_wave_min_k=33, blocks_per_cu=64, use_graph=True, clip_lanes=16):
        from ._native import BatchTuning
        t = BatchTuning(clip_lanes, clip_slots, knn_wave_min_k, blocks_per_cu, int(use_graph))
        check(lib().pg_batch_set_tuning(self.h, C.byref(t)))

    def set_kmodel(self, km: "KModel | None"):
        """Learned k for the following steps (None: fixed k)."""
        check(lib().pg_batch_set_kmodel(self.h, km.h if km is not None else None))
        self._km = km

    def autotune(self, reps=2):
        from ._native import BatchTuning
        t, ms = BatchTuning(), C.c_double()
        check(lib().pg_batch_autotune(self.ctx.h, self.h, reps, C.byref(t), C.byref(ms)))
        return {name: getattr(t, name) for name, _ in t._fields_}, ms.value

    def get(self):
        pts = np.empty((self.n, self.dim))
        rho = np.empty(self.n)
        vel = np.empty((self.n, self.dim))
        E = np.empty(self.n)
        check(lib().pg_batch_get(self.ctx.h, self.h, _ptr(pts), _ptr(rho), _ptr(vel), _ptr(E)))
        return pts, rho, vel, E

    def dt(self):
        out = np.empty(len(self.sizes))
        check(lib().pg_batch_dt(self.ctx.h, self.h, _ptr(out)))
        return out

    def totals(self):
        out = np.empty((len(self.sizes), 5))
        check(lib().pg_batch_totals(self.ctx.h, self.h, _ptr(out)))
        return out

    def info(self) -> dict:
        from ._native import BatchInfo
        i = BatchInfo()
        check(lib().pg_batch_info(self.ctx.h, self.h, C.byref(i)))
        return {name: getattr(i, name) for name, _ in i._fields_}

    def free(self):
        if getattr(self, "h", None):
            lib().pg_batch_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.free()
