"""ctypes binding of libj3dg.so (include/j3dg.h, j3dg_host.h, j3dg_scenegen.h).

The library must have been built (``python -m j3d_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise.  There is no
CPU fallback: rendering calls go to the HIP kernels on the GPU.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libj3dg.so")

J3DG_OK = 0
ERRORS = {1: "invalid argument", 2: "no device", 3: "out of memory", 4: "launch failure",
          5: "not found", 6: "unsupported", 7: "internal"}

ONE_BIT = 1 << 0
SHADOW = 1 << 1
EDGES = 1 << 2
WIREFRAME = 1 << 3
SHADING = 1 << 4
TEXTURED = 1 << 5
VERTEXCOLORS = 1 << 6
DIRECTIONAL_LIGHT = 1 << 7
NO_SCENE = 1 << 8
LEARNED = 1 << 9
KEEP_PROPOSAL = 1 << 10
MLP_PARAMS = 13506

STAGE_TRACE = 1
STAGE_SHADE = 2
STAGE_ALL = 3
BUILD_HOST = 0
BUILD_DEVICE = 1
MAX_OBJECTS = 16

# j3d/pixel.h:11-27
PIXEL_DTYPE = np.dtype([("mark", "u1"), ("r", "u1"), ("g", "u1"), ("b", "u1"), ("u", "<f4"), ("v", "<f4"),
                        ("depth", "<f4"), ("object_id", "<u4"), ("barycentric_u", "<f4"),
                        ("barycentric_v", "<f4"), ("db_id", "<u4")])
assert PIXEL_DTYPE.itemsize == 32


class FrameDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("x0", C.c_int32), ("y0", C.c_int32), ("x1", C.c_int32), ("y1", C.c_int32),
                ("proj_inv", C.c_float * 16), ("near_plane", C.c_float),
                ("cs", C.c_float * 16), ("cs_inv", C.c_float * 16),
                ("pivot", C.c_float * 3), ("diagonal", C.c_float),
                ("flags", C.c_uint32), ("light_dir", C.c_float * 3),
                ("bg_top", C.c_uint32), ("bg_bottom", C.c_uint32),
                ("n_objects", C.c_uint32), ("objects", C.c_uint32 * MAX_OBJECTS),
                ("margin", C.c_float), ("object_list", C.c_void_p)]


class BuildOpts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("leaf_size", C.c_uint32), ("threads", C.c_uint32),
                ("builder", C.c_uint32)]


class MeshInfo(C.Structure):
    _fields_ = [("n_vertices", C.c_uint64), ("n_triangles", C.c_uint64), ("n_nodes", C.c_uint64),
                ("width", C.c_uint32), ("max_depth", C.c_uint32),
                ("bb_min", C.c_float * 3), ("bb_max", C.c_float * 3),
                ("build_seconds", C.c_double), ("upload_seconds", C.c_double),
                ("device_bytes", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("stack_depth", C.c_uint32), ("rays_per_item", C.c_uint32),
                ("coop_tail", C.c_uint32), ("block_size", C.c_uint32), ("refill_min", C.c_uint32),
                ("split_passes", C.c_uint32), ("cost_order", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("hits", C.c_uint64), ("shadowed", C.c_uint64),
                ("learned_resolved", C.c_uint64), ("learned_fallback", C.c_uint64),
                ("stack_overflow", C.c_uint64), ("wave_trips", C.c_uint64), ("lane_steps", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("shadow_wave_trips", C.c_uint64), ("shadow_lane_steps", C.c_uint64),
                ("max_ray_steps", C.c_uint64), ("max_shadow_ray_steps", C.c_uint64),
                ("ray_steps_log2", C.c_uint64 * 7), ("shadow_ray_steps_log2", C.c_uint64 * 7),
                ("trace_ms", C.c_float), ("shade_ms", C.c_float), ("mlp_ms", C.c_float), ("shadow_ms", C.c_float),
                ("splat_ms", C.c_float)]


class AutotuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("best_margin", C.c_float), ("best_ms", C.c_float), ("baseline_ms", C.c_float),
                ("n_tried", C.c_uint32), ("width", C.c_uint32)]


class TrainOpts(C.Structure):
    _fields_ = [("n_rays", C.c_uint64), ("epochs", C.c_uint32), ("batch", C.c_uint32),
                ("lr", C.c_float), ("lr_final", C.c_float), ("seed", C.c_uint64),
                ("rays_per_camera", C.c_uint32), ("dist_min", C.c_float), ("dist_max", C.c_float),
                ("fov_deg", C.c_float), ("aspect", C.c_float), ("n_validation", C.c_uint32),
                ("warm_start", C.c_uint32)]


class TrainReport(C.Structure):
    _fields_ = [("loss_first", C.c_float), ("loss_last", C.c_float), ("steps", C.c_uint64),
                ("hit_fraction", C.c_float), ("label_ms", C.c_float), ("train_ms", C.c_float),
                ("val_hit_accuracy", C.c_float), ("val_mean_abs_error", C.c_float),
                ("val_resolved", C.c_float * 4), ("norm", C.c_float * 4)]


class Camera(C.Structure):
    _fields_ = [("focal_length", C.c_float), ("film_w_inch", C.c_float), ("film_h_inch", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("zoom", C.c_float),
                ("fit_film", C.c_int32)]


_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_I = C.c_int
_F = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    # j3dg.h
    "j3dg_create": (_I, [_I, C.POINTER(_P)]),
    "j3dg_destroy": (None, [_P]),
    "j3dg_last_error": (C.c_char_p, [_P]),
    "j3dg_abi_version": (_I, []),
    "j3dg_device_count": (_I, [C.POINTER(_I)]),
    "j3dg_synchronize": (_I, [_P]),
    "j3dg_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "j3dg_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "j3dg_set_counting": (_I, [_P, _I]),
    "j3dg_get_stats": (_I, [_P, C.POINTER(Stats)]),
    "j3dg_reset_timings": (_I, [_P]),
    "j3dg_get_timings": (_I, [_P, _P, _P, _P, _P, _U32, C.POINTER(_U32)]),
    "j3dg_upload_mesh": (_I, [_P, _U32, _P, _U64, _P, _U64, _P, C.POINTER(BuildOpts), C.POINTER(_U32)]),
    "j3dg_all_hits": (_I, [_P, _U32, _P, _U64, _U32, _P, _P, _P, _P, _P]),
    "j3dg_voxelize": (_I, [_P, _U32, _U32, _P, _P, _U64]),
    "j3dg_upload_points": (_I, [_P, _U32, _P, _U64, _P, _P, _P, _P]),
    "j3dg_remove_points": (_I, [_P, _U32]),
    "j3dg_set_points_cs": (_I, [_P, _U32, _P]),
    "j3dg_splat_points": (_I, [_P, _U32, _P, _P, _U32, _P, _U32]),
    "j3dg_remove_mesh": (_I, [_P, _U32]),
    "j3dg_set_mesh_cs": (_I, [_P, _U32, _P]),
    "j3dg_set_mesh_colors": (_I, [_P, _U32, _P, _P, _P, _U32, _U32]),
    "j3dg_get_mesh_info": (_I, [_P, _U32, C.POINTER(MeshInfo)]),
    "j3dg_mesh_normals": (_I, [_P, _U32, _P]),
    "j3dg_set_matcap": (_I, [_P, _P, _U32, _U32, _U32]),
    "j3dg_create_target": (_I, [_P, _U32, _U32, C.POINTER(_U32)]),
    "j3dg_destroy_target": (_I, [_P, _U32]),
    "j3dg_write_target_gbuffer": (_I, [_P, _U32, _P]),
    "j3dg_target_device_ptrs": (_I, [_P, _U32, C.POINTER(_P), C.POINTER(_P)]),
    "j3dg_render": (_I, [_P, C.POINTER(FrameDesc), _U32, _U32, _P, _P]),
    "j3dg_read_target": (_I, [_P, _U32, _P, _P]),
    "j3dg_read_target_strided": (_I, [_P, _U32, _P, _U32, _P, _U32]),
    "j3dg_render_batch": (_I, [_P, C.POINTER(FrameDesc), _P, _U32]),
    "j3dg_trace_rays": (_I, [_P, _P, _U32, _P, _U64, _P, _P, _P, _P, _P]),
    "j3dg_any_hit_rays": (_I, [_P, _P, _U32, _P, _U64, _P]),
    "j3dg_trace_rays_device": (_I, [_P, _P, _U32, _P, _U64, _P, _P, _P, _P, _P]),
    "j3dg_calibration_copy": (_I, [_P, _U64, _U32, C.POINTER(_F)]),
    "j3dg_debug_timeline": (_I, [_P, _P, _U32, C.POINTER(_U32)]),
    "j3dg_autotune": (_I, [_P, C.POINTER(FrameDesc), _U32, _U32, C.POINTER(AutotuneReport)]),
    "j3dg_autotune_widths": (_I, [_P, C.POINTER(FrameDesc), C.POINTER(FrameDesc), _U32, _U32,
                                  C.POINTER(AutotuneReport)]),
    "j3dg_autotune_batch": (_I, [_P, _P, _P, _P, _U32, _U32, C.POINTER(AutotuneReport)]),
    "j3dg_default_train_opts": (_I, [C.POINTER(TrainOpts)]),
    "j3dg_train_proposal": (_I, [_P, _U32, C.POINTER(TrainOpts), C.POINTER(TrainReport)]),
    "j3dg_get_proposal": (_I, [_P, _U32, _P, _P]),
    "j3dg_set_proposal": (_I, [_P, _U32, _P, _P]),
    "j3dg_write_target_proposal": (_I, [_P, _U32, _P]),
    "j3dg_read_target_proposal": (_I, [_P, _U32, _P]),
    "j3dg_mlp_forward": (_I, [_P, _P, _P, _P, _U64, _P]),
    "j3dg_mlp_gradients": (_I, [_P, _P, _P, _P, _P, _U64, _P, C.POINTER(_F)]),
    "j3dg_mlp_adam_step": (_I, [_P, _P, _P, _P, _P, _U32, _F]),
    # j3dg_host.h
    "j3dg_default_camera": (None, [C.POINTER(Camera)]),
    "j3dg_frustum": (None, [_F, _F, _F, _F, _F, _F, _P]),
    "j3dg_invert_projection": (None, [_P, _P]),
    "j3dg_projection": (None, [C.POINTER(Camera), _I, _I, _P]),
    "j3dg_invert_orthonormal": (None, [_P, _P]),
    "j3dg_make_matcap": (_I, [_I, _P, C.POINTER(_U32)]),
    # j3dg_scenegen.h
    "j3dg_icosphere_vertex_count": (_U64, [_U32]),
    "j3dg_icosphere_triangle_count": (_U64, [_U32]),
    "j3dg_gen_icosphere": (_I, [_U32, _U64, _F, _P, _P]),
    "j3dg_gen_torus": (_I, [_U32, _U32, _F, _F, _U64, _F, _U32, _F, _F, _F, _I, _P, _P]),
    "j3dg_orbit_camera": (_I, [_F, _F, _F, _P, _P, _P]),
    "j3dg_splitmix64": (_U64, [C.POINTER(_U64)]),
    "j3dg_uniform01": (C.c_double, [C.POINTER(_U64)]),
    "j3dg_random_hemisphere": (None, [_U64, _P]),
}

_lib = None


def load(path=None):
    """Loads libj3dg.so (once) and declares every exported function.  The
    J3DG_LIBRARY environment variable names another build of it (an older one
    for A/B timing may lack newer entry points: those stay undeclared)."""
    global _lib
    if _lib is not None:
        return _lib
    other = os.environ.get("J3DG_LIBRARY")
    path = path or other or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("libj3dg.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        if other and not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Address of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)
