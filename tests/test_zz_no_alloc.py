"""pg_knn_solve_noalloc: the KNN build whose scratch lives in a static device
array (for callers that cannot allocate) equals the allocating path
(pg_knn_init + pg_knn_solve + get_*) bit for bit, on two calls in a row --
the second smaller, so it runs over the first call's leftovers in the
array -- and refuses sizes beyond the array before launching anything."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd._native import PgError, check, lib
from proteusgpu_amd.generators import uniform_points

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime libpgpu.so itself loads (the
    caller-side plumbing a C user would write with hipMalloc/hipMemcpy)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(int(nbytes), 1)) == 0
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ps):
        for p in ps:
            assert self.lib.hipFree(p) == 0


def _noalloc(ctx, hip, pts, k):
    n, dim = pts.shape
    d_pts = hip.upload(pts.astype(np.float64))
    # outputs start as garbage: every entry must be written by the call
    d_sorted = hip.upload(np.full((n, dim), np.nan))
    d_perm = hip.upload(np.full(n, 0xFFFFFFFF, np.uint32))
    d_rows = hip.upload(np.full((n, k), 0xFFFFFFF9, np.uint32))
    assert hip.lib.hipDeviceSynchronize() == 0  # uploads done before the context's stream reads them
    N = C.c_int()
    check(lib().pg_knn_solve_noalloc(ctx.h, dim, d_pts, n, k, d_sorted, d_perm, d_rows, C.byref(N)))
    ctx.sync()
    out = (N.value, hip.download(d_sorted, (n, dim), np.float64), hip.download(d_perm, n, np.uint32),
           hip.download(d_rows, (n, k), np.uint32))
    return out, (d_pts, d_sorted, d_perm, d_rows)


def _allocating(ctx, pts, k):
    kp = api.KnnProblem(ctx, pts)
    kp.solve(k)
    return kp.N_grid, kp.get_points(), kp.get_permutation(), kp.get_knearest()


def test_noalloc_knn_equals_allocating_path(ctx):
    hip = _Hip()
    cases = [(uniform_points(200_000, 3, 61), 32), (uniform_points(5000, 2, 62), 16)]
    for i, (pts, k) in enumerate(cases):
        print(f"noalloc call {i}: n={len(pts)} dim={pts.shape[1]} k={k}", flush=True)
        got, bufs = _noalloc(ctx, hip, pts, k)
        want = _allocating(ctx, pts, k)
        assert got[0] == want[0]
        for a, b in zip(got[1:], want[1:]):
            assert np.array_equal(a, b)
        if i == 1:
            _, _, _, kn = O.knn(pts, k)
            assert np.array_equal(got[3], kn)
            # beyond the scratch array: refused before any launch
            with pytest.raises(PgError) as e:
                check(lib().pg_knn_solve_noalloc(ctx.h, 3, bufs[0], 6_000_000, 32, bufs[1], bufs[2], bufs[3],
                                                 None))
            assert e.value.code == -1 and "scratch" in str(e.value)
        hip.free(*bufs)
