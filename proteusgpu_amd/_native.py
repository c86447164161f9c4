"""ctypes binding of libpgpu.so (the C ABI declared in include/pgpu.h).

The library is built in-tree (proteusgpu_amd/libpgpu.so) by
``make -C proteusgpu_amd/csrc``.  There is no other implementation of these
entry points: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgpu.so")

PG_OK = 0
ERRORS = {-1: "PG_ERR_ARG", -2: "PG_ERR_HIP", -3: "PG_ERR_NOMEM", -4: "PG_ERR_GRID",
          -5: "PG_ERR_CELLS", -6: "PG_ERR_NODEV", -7: "PG_ERR_FAILED"}

BOUNDARY = {"periodic": 0, "reflective": 1, "none": 2}
FLUX = {"reference": 0, "physical": 1}


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class MeshParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("boundary", C.c_int), ("buff", C.c_double), ("k", C.c_int),
                ("max_p", C.c_int), ("max_t", C.c_int), ("restart", C.c_int), ("k_max", C.c_int),
                ("perturb", C.c_double), ("exact", C.c_int)]


class MeshInfo(C.Structure):
    _fields_ = [("n_seeds", C.c_int64), ("n_hydro", C.c_int64), ("num_faces", C.c_int64),
                ("n_failed", C.c_int64), ("n_restarted", C.c_int64), ("N_grid", C.c_int),
                ("k_used", C.c_int), ("row_entries", C.c_int64), ("n_class", C.c_int64 * 6),
                ("n_degenerate", C.c_int64), ("displacement", C.c_double), ("n_exact", C.c_int64),
                ("n_ball", C.c_int64)]


class VMeshView(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "cell_ids", "seeds", "volumes", "face_counts", "face_ptr", "neighbor_cell", "face_normal",
        "face_area", "ghost_ids", "status", "kcert", "cell_centroid", "face_centroid")]


class SimParams(C.Structure):
    _fields_ = [("mesh", MeshParams), ("gamma", C.c_double), ("cfl", C.c_double),
                ("flux_mode", C.c_int), ("move_mesh", C.c_int)]


class SimTimes(C.Structure):
    _fields_ = [("hydro_ms", C.c_double), ("ghosts_ms", C.c_double), ("grid_ms", C.c_double),
                ("knn_ms", C.c_double), ("clip_ms", C.c_double), ("restart_ms", C.c_double),
                ("final_ms", C.c_double), ("total_ms", C.c_double), ("n_restarted", C.c_int64),
                ("k_used", C.c_int), ("predict_ms", C.c_double), ("row_entries", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchTuning(C.Structure):
    _fields_ = [("clip_lanes", C.c_int), ("clip_slots", C.c_int), ("knn_wave_min_k", C.c_int), ("blocks_per_cu", C.c_int),
                ("use_graph", C.c_int)]


class BatchInfo(C.Structure):
    _fields_ = [("nbox", C.c_int), ("n_hydro", C.c_int64), ("n_seeds", C.c_int64), ("seed_capacity", C.c_int64),
                ("last_step_ms", C.c_double)]


KMODEL_NFEAT, KMODEL_NCLASS, KMODEL_NPARAM = 16, 6, 5638
KMODEL_BUCKETS = (16, 24, 32, 48, 64, 96)


_lib = None


def _signatures():
    """ctypes argument and result types of every C ABI entry point."""
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    pp = C.POINTER(C.c_void_p)
    return {
        "pg_version": ([], C.c_char_p), "pg_last_error": ([], C.c_char_p),
        "pg_device_count": ([C.POINTER(C.c_int)], i32),
        "pg_set_clip_config": ([i32, i32], i32),
        "pg_get_clip_config": ([C.POINTER(C.c_int), C.POINTER(C.c_int)], i32),
        "pg_set_clip_resume": ([i32], i32),
        "pg_get_clip_resume": ([C.POINTER(C.c_int)], i32),
        "pg_ctx_create": ([i32, pp], i32), "pg_ctx_destroy": ([vp], i32), "pg_ctx_sync": ([vp], i32),
        "pg_stream_copy_probe": ([vp, i64, i32, C.POINTER(C.c_double)], i32),
        "pg_knn_init": ([vp, i32, vp, i64, pp], i32), "pg_knn_solve": ([vp, vp, i32], i32),
        "pg_knn_info": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int)], i32),
        "pg_knn_get_points": ([vp, vp, vp], i32), "pg_knn_get_knearest": ([vp, vp, vp], i32),
        "pg_knn_get_permutation": ([vp, vp, vp], i32), "pg_knn_free": ([pp], i32),
        "pg_knn_solve_noalloc": ([vp, i32, vp, i64, i32, vp, vp, vp, C.POINTER(C.c_int)], i32),
        "pg_mesh_params_default": ([C.POINTER(MeshParams)], None),
        "pg_mesh_compute": ([vp, C.POINTER(MeshParams), vp, i64, pp], i32),
        "pg_mesh_info": ([vp, C.POINTER(MeshInfo)], i32),
        "pg_mesh_copy_out": ([vp, vp, C.POINTER(VMeshView)], i32), "pg_mesh_free": ([pp], i32),
        "pg_mesh_get_knn": ([vp, vp, vp, vp, vp], i32),
        "pg_state_create": ([vp, vp, vp, vp, vp, pp], i32),
        "pg_dt_cfl": ([vp, vp, vp, dbl, dbl, C.POINTER(C.c_double)], i32),
        "pg_hydro_step": ([vp, vp, vp, dbl, dbl, i32], i32),
        "pg_state_copy_out": ([vp, vp, vp, vp, vp], i32), "pg_state_free": ([pp], i32),
        "pg_hll_flux": ([vp, i32, dbl, i64, vp, vp, vp, vp, i32, vp], i32),
        "pg_voronoi_compute": ([vp, vp, i32, i32, pp, vp], i32),
        "pg_mesh_unpermute": ([vp, vp], i32),
        "pg_sim_create": ([vp, C.POINTER(SimParams), vp, vp, vp, vp, i64, pp], i32),
        "pg_sim_step": ([vp, vp, dbl, C.POINTER(C.c_double)], i32),
        "pg_sim_hydro_steps": ([vp, vp, i32, dbl, C.POINTER(C.c_double)], i32),
        "pg_sim_times_last": ([vp, C.POINTER(SimTimes)], i32),
        "pg_sim_mesh": ([vp, pp], i32),
        "pg_sim_get": ([vp, vp, vp, vp, vp, vp], i32),
        "pg_sim_totals": ([vp, vp, vp], i32),
        "pg_sim_free": ([pp], i32),
        "pg_sim_set_kmodel": ([vp, vp], i32),
        "pg_kmodel_create": ([vp, C.c_uint64, pp], i32),
        "pg_kmodel_set_params": ([vp, vp, vp], i32), "pg_kmodel_get_params": ([vp, vp, vp], i32),
        "pg_kmodel_set_cost": ([vp, dbl], i32),
        "pg_kmodel_predict": ([vp, vp, vp, i64, vp, vp], i32),
        "pg_kmodel_grad": ([vp, vp, vp, vp, i64, vp, C.POINTER(C.c_double)], i32),
        "pg_kmodel_train_step": ([vp, vp, vp, vp, i64, dbl, C.POINTER(C.c_double)], i32),
        "pg_kmodel_train_mesh": ([vp, vp, vp, i32, i64, dbl, C.c_uint64, C.POINTER(C.c_double)], i32),
        "pg_kmodel_free": ([pp], i32),
        "pg_mesh_features": ([vp, vp, vp], i32),
        "pg_mesh_compute_k": ([vp, C.POINTER(MeshParams), vp, vp, i64, pp], i32),
        "pg_batch_create": ([vp, C.POINTER(SimParams), i32, vp, vp, vp, vp, vp, pp], i32),
        "pg_batch_step": ([vp, vp, C.POINTER(C.c_double)], i32),
        "pg_batch_set_tuning": ([vp, C.POINTER(BatchTuning)], i32),
        "pg_batch_set_kmodel": ([vp, vp], i32),
        "pg_batch_get_tuning": ([vp, C.POINTER(BatchTuning)], i32),
        "pg_batch_autotune": ([vp, vp, i32, C.POINTER(BatchTuning), C.POINTER(C.c_double)], i32),
        "pg_batch_get": ([vp, vp, vp, vp, vp, vp], i32),
        "pg_batch_dt": ([vp, vp, vp], i32),
        "pg_batch_totals": ([vp, vp, vp], i32),
        "pg_batch_info": ([vp, vp, C.POINTER(BatchInfo)], i32),
        "pg_batch_free": ([pp], i32),
    }


def lib():
    """Load libpgpu.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libpgpu.so is not built ({LIB_PATH}); run `make -C proteusgpu_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    for name, (args, res) in _signatures().items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc):
    if rc != PG_OK:
        raise PgError(rc, lib().pg_last_error().decode(errors="replace"))


def exported_symbols():
    """Names of the C ABI entry points this binding knows (for the export test)."""
    return sorted(_signatures())
