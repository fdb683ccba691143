"""PLY files through the canvas layer's C ABI (include/j3dg_io.h, in
libj3d_canvas.so): read_ply / write_ply of j3d's io layer (j3d/io.cpp:733-741)."""
import ctypes as C
import os

import numpy as np

CANVAS_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libj3d_canvas.so")


class MeshArrays(C.Structure):
    _fields_ = [("n_vertices", C.c_uint64), ("n_triangles", C.c_uint64),
                ("vertices", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("colors", C.POINTER(C.c_uint32)), ("triangles", C.POINTER(C.c_uint32)),
                ("uv", C.POINTER(C.c_float))]


class PlyError(RuntimeError):
    pass


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(CANVAS_LIB):
            raise RuntimeError("libj3d_canvas.so is not built (%s); run __graft_entry__.build()" % CANVAS_LIB)
        lib = C.CDLL(CANVAS_LIB)
        lib.j3dg_read_ply.argtypes = [C.c_char_p, C.POINTER(MeshArrays)]
        lib.j3dg_read_ply.restype = C.c_int
        lib.j3dg_free_mesh_arrays.argtypes = [C.POINTER(MeshArrays)]
        lib.j3dg_free_mesh_arrays.restype = None
        lib.j3dg_write_ply.argtypes = [C.c_char_p, C.POINTER(MeshArrays), C.c_int]
        lib.j3dg_write_ply.restype = C.c_int
        lib.j3dg_ply_error.argtypes = []
        lib.j3dg_ply_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _copy(ptr, n, dtype, shape):
    if not ptr or n == 0:
        return None if not ptr else np.zeros(shape, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def read_ply(path):
    """{"vertices": n x 3 float32, "triangles": m x 3 uint32, "normals": n x 3 or
    None, "colors": n uint32 (0xAABBGGRR) or None, "uv": m x 6 float32 or None}."""
    lib = _load()
    m = MeshArrays()
    if lib.j3dg_read_ply(os.fsencode(path), C.byref(m)):
        raise PlyError(lib.j3dg_ply_error().decode())
    try:
        nv, nt = int(m.n_vertices), int(m.n_triangles)
        return {"vertices": _copy(m.vertices, 3 * nv, np.float32, (nv, 3)) if nv else np.zeros((0, 3), np.float32),
                "triangles": _copy(m.triangles, 3 * nt, np.uint32, (nt, 3)) if nt else np.zeros((0, 3), np.uint32),
                "normals": _copy(m.normals, 3 * nv, np.float32, (nv, 3)),
                "colors": _copy(m.colors, nv, np.uint32, (nv,)),
                "uv": _copy(m.uv, 6 * nt, np.float32, (nt, 6))}
    finally:
        lib.j3dg_free_mesh_arrays(C.byref(m))


def write_ply(path, vertices, triangles, normals=None, colors=None, uv=None, ascii=False):
    lib = _load()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    keep = [v, t]
    m = MeshArrays()
    m.n_vertices, m.n_triangles = v.shape[0], t.shape[0]
    m.vertices = v.ctypes.data_as(C.POINTER(C.c_float))
    m.triangles = t.ctypes.data_as(C.POINTER(C.c_uint32))
    if normals is not None:
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        if n.shape[0] != v.shape[0]:
            raise ValueError("normals: one per vertex")
        keep.append(n)
        m.normals = n.ctypes.data_as(C.POINTER(C.c_float))
    if colors is not None:
        c = np.ascontiguousarray(colors, np.uint32).reshape(-1)
        if c.shape[0] != v.shape[0]:
            raise ValueError("colors: one per vertex")
        keep.append(c)
        m.colors = c.ctypes.data_as(C.POINTER(C.c_uint32))
    if uv is not None:
        u = np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
        if u.shape[0] != t.shape[0]:
            raise ValueError("uv: 6 floats per triangle")
        keep.append(u)
        m.uv = u.ctypes.data_as(C.POINTER(C.c_float))
    if lib.j3dg_write_ply(os.fsencode(path), C.byref(m), 1 if ascii else 0):
        raise PlyError(lib.j3dg_ply_error().decode())
