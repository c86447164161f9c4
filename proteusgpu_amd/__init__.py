"""MI355X-native moving-mesh hot path (grid KNN, Voronoi clipping, HLL update).

The compute path is libpgpu.so (HIP kernels for gfx950 behind include/pgpu.h);
this package is its Python binding plus seeded IC generators.
"""
from .generators import sod_state, splitmix_uniform, uniform_points  # noqa: F401


def __getattr__(name):
    # the binding loads libpgpu.so on first use only
    if name in ("Context", "KnnProblem", "Mesh", "HydroState", "device_count", "mesh_params"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
