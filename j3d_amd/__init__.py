"""j3d_amd -- j3d's per-frame ray-casting path on MI355X (gfx950).

The compute path is libj3dg.so (hand-written HIP kernels behind the C ABI in
include/j3dg.h).  This package holds its sources (csrc/), the C++ canvas layer
that mirrors j3d's `canvas` class (host/), the build script and a ctypes
binding used by the tests and the benchmark.
"""
from . import ffi  # noqa: F401
from .api import (Frame, J3dgError, Renderer, bbox, camera_with_vertical_fov, default_camera,  # noqa: F401
                  icosphere, invert_orthonormal, make_db_id, make_frame, make_matcap, orbit_camera, projection,
                  random_hemisphere, read_proposal, scene_framing, torus, translation_camera, write_proposal)
