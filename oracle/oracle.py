"""ctypes wrapper of oracle/liboracle.so -- the CPU restatement of j3d's
ray-casting path (see oracle.h for what it restates and how it is pinned).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ONE_BIT = 1 << 0
SHADOW = 1 << 1
EDGES = 1 << 2
WIREFRAME = 1 << 3
SHADING = 1 << 4
TEXTURED = 1 << 5
VERTEXCOLORS = 1 << 6
DIRECTIONAL_LIGHT = 1 << 7

PIXEL_DTYPE = np.dtype([("mark", "u1"), ("r", "u1"), ("g", "u1"), ("b", "u1"), ("u", "<f4"), ("v", "<f4"),
                        ("depth", "<f4"), ("object_id", "<u4"), ("barycentric_u", "<f4"),
                        ("barycentric_v", "<f4"), ("db_id", "<u4")])


class Camera(C.Structure):
    _fields_ = [("focal_length", C.c_float), ("film_w_inch", C.c_float), ("film_h_inch", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("zoom", C.c_float), ("fit_film", C.c_int32)]


class Object(C.Structure):
    _fields_ = [("bvh", C.c_void_p), ("vertices", C.c_void_p), ("triangles", C.c_void_p), ("normals", C.c_void_p),
                ("n_triangles", C.c_uint64), ("cs", C.c_float * 16), ("db_id", C.c_uint32),
                ("vertex_colors", C.c_void_p), ("uv", C.c_void_p), ("texture", C.c_void_p),
                ("tex_w", C.c_uint32), ("tex_h", C.c_uint32)]


class FrameC(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("x0", C.c_int32), ("y0", C.c_int32), ("x1", C.c_int32),
                ("y1", C.c_int32), ("proj_inv", C.c_float * 16), ("near_plane", C.c_float),
                ("cs", C.c_float * 16), ("cs_inv", C.c_float * 16), ("pivot", C.c_float * 3),
                ("diagonal", C.c_float), ("flags", C.c_uint32), ("light_dir", C.c_float * 3)]


class Stats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("hits", C.c_uint64),
                ("shadowed", C.c_uint64), ("near_ties", C.c_uint64)]


_P = C.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_default_camera": (None, [C.POINTER(Camera)]),
        "orc_projection": (None, [C.POINTER(Camera), C.c_int, C.c_int, _P]),
        "orc_invert_projection": (None, [_P, _P]),
        "orc_invert_orthonormal": (None, [_P, _P]),
        "orc_mat_vec": (None, [_P, _P, _P]),
        "orc_matcap": (None, [C.c_int, _P, C.POINTER(C.c_uint32)]),
        "orc_fill_background": (None, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_triangle_normals": (None, [_P, _P, C.c_uint64, _P]),
        "orc_bbox": (None, [_P, C.c_uint64, _P, _P]),
        "orc_prepare_and_unzoom": (None, [C.c_uint32, _P, _P, _P, _P, _P, _P, _P]),
        "orc_bvh_build": (_P, [_P, C.c_uint64, _P, C.c_uint64, C.c_int]),
        "orc_bvh_free": (None, [_P]),
        "orc_bvh_node_count": (C.c_uint64, [_P]),
        "orc_update_canvas": (C.c_int, [_P, C.c_uint32, C.POINTER(FrameC), _P, C.c_int, C.c_int, C.POINTER(Stats)]),
        "orc_canvas_to_image": (None, [_P, C.c_uint32, C.c_uint32, _P, C.c_float, C.c_uint32, _P, C.c_uint32,
                                       C.c_uint32, C.c_uint32, _P]),
        "orc_closest_hits": (None, [_P, C.c_uint32, _P, C.c_uint64, C.c_int, _P, _P, _P, _P, _P, _P]),
        "orc_any_hits": (None, [_P, C.c_uint32, _P, C.c_uint64, C.c_int, _P]),
        "orc_all_hits": (None, [_P, _P, C.c_uint64, C.c_uint32, C.c_int, _P, _P, _P, _P, _P]),
        "orc_vox_dims": (None, [_P, _P, C.c_uint32, _P]),
        "orc_voxelize": (None, [_P, _P, _P, _P, _P]),
        "orc_vox_color_index": (C.c_uint8, [C.c_uint8, C.c_uint8, C.c_uint8]),
        "orc_splat_zbuffer": (None, [_P, C.c_uint64, _P]),
        "orc_splat_cloud": (None, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, _P, C.c_uint32, _P, C.c_uint64, _P,
                                   _P, C.c_int, C.c_int]),
        "orc_intersect_triangle": (C.c_int, [_P, _P, _P, _P, _P, C.c_float, C.c_float, _P, _P, _P]),
        "orc_color_byte": (C.c_uint8, [C.c_float]),
        "orc_mlp_param_count": (C.c_int, []),
        "orc_mlp_line": (None, [_P, C.c_uint64, _P, _P]),
        "orc_mlp_features": (None, [_P, C.c_uint64, _P, _P]),
        "orc_mlp_forward": (None, [_P, _P, _P, C.c_uint64, C.c_int, _P]),
        "orc_mlp_gradients": (C.c_float, [_P, _P, _P, _P, C.c_uint64, C.c_int, _P]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def default_camera():
    c = Camera()
    load().orc_default_camera(C.byref(c))
    return c


def projection(camera, w, h):
    P = np.zeros(16, np.float32)
    Pi = np.zeros(16, np.float32)
    lib = load()
    lib.orc_projection(C.byref(camera), int(w), int(h), _p(P))
    lib.orc_invert_projection(_p(P), _p(Pi))
    return P, Pi


def invert_orthonormal(m):
    m = np.ascontiguousarray(m, np.float32)
    o = np.zeros(16, np.float32)
    load().orc_invert_orthonormal(_p(m), _p(o))
    return o


def mat_vec(m, v):
    m = np.ascontiguousarray(m, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    o = np.zeros(4, np.float32)
    load().orc_mat_vec(_p(m), _p(v), _p(o))
    return o


def matcap(kind=0):
    im = np.zeros(512 * 512, np.uint32)
    cav = C.c_uint32()
    load().orc_matcap(int(kind), _p(im), C.byref(cav))
    return im.reshape(512, 512), int(cav.value)


def background(w, h, top=0xff000000, bottom=0xff404040):
    im = np.zeros((h, w), np.uint32)
    load().orc_fill_background(_p(im), w, h, top, bottom)
    return im


def prepare_and_unzoom(boxes):
    """prepare_scene + unzoom (j3d/scene.cpp:50-110) over a list of (cs, bb_min,
    bb_max): the scene's objects, then its point clouds.  Returns (diagonal,
    pivot, cs, cs_inv)."""
    n = len(boxes)
    cs16 = np.ascontiguousarray(np.concatenate([np.asarray(b[0], np.float32).reshape(16) for b in boxes]), np.float32)
    lo = np.ascontiguousarray(np.concatenate([np.asarray(b[1], np.float32).reshape(3) for b in boxes]), np.float32)
    hi = np.ascontiguousarray(np.concatenate([np.asarray(b[2], np.float32).reshape(3) for b in boxes]), np.float32)
    d = np.zeros(1, np.float32)
    pivot = np.zeros(3, np.float32)
    cs = np.zeros(16, np.float32)
    ci = np.zeros(16, np.float32)
    load().orc_prepare_and_unzoom(n, _p(cs16), _p(lo), _p(hi), _p(d), _p(pivot), _p(cs), _p(ci))
    return float(d[0]), pivot, cs, ci


def triangle_normals(v, t):
    v = np.ascontiguousarray(v, np.float32)
    t = np.ascontiguousarray(t, np.uint32)
    n = np.zeros((t.shape[0], 3), np.float32)
    load().orc_triangle_normals(_p(v), _p(t), t.shape[0], _p(n))
    return n


class Mesh:
    """One scene object for the oracle: arrays, normals and (optionally) a BVH."""

    def __init__(self, verts, tris, cs=None, db_id=(1 << 29), use_bvh=True, threads=0, vertex_colors=None,
                 uv=None, texture=None):
        self.v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self.t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        # optional surface colour: rgb per vertex, (u, v) per triangle corner, texture (h, w) uint32
        self.vcol = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, np.float32).reshape(-1, 3)
        self.uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
        self.tex = None if texture is None else np.ascontiguousarray(texture, np.uint32)
        self.n = triangle_normals(self.v, self.t)
        self.cs = np.eye(4, dtype=np.float32).reshape(16) if cs is None else np.asarray(cs, np.float32)
        self.db_id = int(db_id)
        self.bvh = None
        if use_bvh:
            self.bvh = load().orc_bvh_build(_p(self.v), self.v.shape[0], _p(self.t), self.t.shape[0], int(threads))

    def obj(self, brute=False):
        o = Object()
        o.bvh = None if brute else self.bvh
        o.vertices = self.v.ctypes.data
        o.triangles = self.t.ctypes.data
        o.normals = self.n.ctypes.data
        o.n_triangles = self.t.shape[0]
        o.cs[:] = [float(x) for x in self.cs]
        o.db_id = self.db_id
        o.vertex_colors = None if self.vcol is None else self.vcol.ctypes.data
        o.uv = None if self.uv is None else self.uv.ctypes.data
        if self.tex is not None:
            o.texture = self.tex.ctypes.data
            o.tex_h, o.tex_w = self.tex.shape
        return o

    def __del__(self):
        if getattr(self, "bvh", None):
            load().orc_bvh_free(self.bvh)
            self.bvh = None


def _objs(meshes, brute=False):
    arr = (Object * max(1, len(meshes)))(*[m.obj(brute) for m in meshes])
    return arr, len(meshes)


def color_byte(c):
    """(uint8_t)(c * 255.f) as the oracle defines it (see oracle.h)."""
    return int(load().orc_color_byte(C.c_float(float(c))))


def update_canvas(meshes, frame, gbuf=None, threads=0, count_ties=False, brute=False):
    """canvas::update_canvas on a G-buffer (created with pixel() defaults if None).
    `frame` is a j3d_amd.Frame-like object (width, height, rect, proj_inv, ...)."""
    w, h = frame.width, frame.height
    if gbuf is None:
        gbuf = np.zeros((h, w), PIXEL_DTYPE)
        gbuf["object_id"] = 0xFFFFFFFF
    f = FrameC()
    f.w, f.h = w, h
    f.x0, f.y0, f.x1, f.y1 = [int(x) for x in frame.rect]
    f.proj_inv[:] = [float(x) for x in frame.proj_inv]
    f.near_plane = float(frame.near_plane)
    f.cs[:] = [float(x) for x in frame.cs]
    f.cs_inv[:] = [float(x) for x in frame.cs_inv]
    f.pivot[:] = [float(x) for x in frame.pivot]
    f.diagonal = float(frame.diagonal)
    f.flags = int(frame.flags) & 0xFF
    f.light_dir[:] = [float(x) for x in frame.light_dir]
    arr, n = _objs(meshes, brute)
    st = Stats()
    load().orc_update_canvas(arr, n, C.byref(f), _p(gbuf), int(threads), 1 if count_ties else 0, C.byref(st))
    return gbuf, {k: getattr(st, k) for k, _ in st._fields_}


def canvas_to_image(gbuf, frame, matcap_im, cavity, im=None):
    h, w = gbuf.shape
    if im is None:
        im = background(w, h, frame.bg_top, frame.bg_bottom)
    pinv = np.ascontiguousarray(frame.proj_inv, np.float32)
    mi = np.ascontiguousarray(matcap_im, np.uint32)
    load().orc_canvas_to_image(_p(np.ascontiguousarray(gbuf)), w, h, _p(pinv), float(frame.near_plane),
                               int(frame.flags) & 0xFF, _p(mi), mi.shape[1], mi.shape[0], int(cavity), _p(im))
    return im


def render(meshes, frame, matcap_im, cavity, threads=0, count_ties=False, gbuf=None):
    g, st = update_canvas(meshes, frame, gbuf=gbuf, threads=threads, count_ties=count_ties)
    im = canvas_to_image(g, frame, matcap_im, cavity)
    return g, im, st


def closest_hits(meshes, rays, threads=0, brute=False, second=False):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = r.shape[0]
    out = {"obj": np.zeros(n, np.uint32), "tri": np.zeros(n, np.uint32), "t": np.zeros(n, np.float32),
           "u": np.zeros(n, np.float32), "v": np.zeros(n, np.float32)}
    t2 = np.zeros(n, np.float32) if second else None
    arr, no = _objs(meshes, brute)
    load().orc_closest_hits(arr, no, _p(r), n, int(threads), _p(out["obj"]), _p(out["tri"]), _p(out["t"]),
                            _p(out["u"]), _p(out["v"]), _p(t2))
    if second:
        out["t2"] = t2
    return out


def all_hits(mesh, rays, max_hits, threads=0, brute=False):
    """Every hit of each ray in one mesh, (t, triangle id) order (see oracle.h):
    {"count": n, "tri"/"t"/"u"/"v": n x max_hits}, unused slots tri = MISS."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = r.shape[0]
    out = {"count": np.zeros(n, np.uint32), "tri": np.zeros((n, max_hits), np.uint32),
           "t": np.zeros((n, max_hits), np.float32), "u": np.zeros((n, max_hits), np.float32),
           "v": np.zeros((n, max_hits), np.float32)}
    arr, _ = _objs([mesh], brute)
    load().orc_all_hits(arr, _p(r), n, int(max_hits), int(threads), _p(out["count"]), _p(out["tri"]),
                        _p(out["t"]), _p(out["u"]), _p(out["v"]))
    return out


def vox_dims(bb_min, bb_max, max_dim):
    lo = np.ascontiguousarray(bb_min, np.float32)
    hi = np.ascontiguousarray(bb_max, np.float32)
    d = np.zeros(3, np.uint32)
    load().orc_vox_dims(_p(lo), _p(hi), int(max_dim), _p(d))
    return tuple(int(x) for x in d)


def voxelize(mesh, max_dim, brute=False):
    """write_vox's voxel grid of one mesh (see oracle.h): (dims, uint8 array [z, y, x])."""
    # the mesh box as std::min / std::max folded from the first vertex
    # (j3d/mesh.cpp:39-58): NaN coordinates are skipped, except in the first vertex
    lo = np.fmin.reduce(mesh.v, axis=0).astype(np.float32)
    hi = np.fmax.reduce(mesh.v, axis=0).astype(np.float32)
    first_nan = np.isnan(mesh.v[0])
    lo[first_nan] = np.nan
    hi[first_nan] = np.nan
    dims = vox_dims(lo, hi, max_dim)
    data = np.zeros(dims[0] * dims[1] * dims[2], np.uint8)
    arr, _ = _objs([mesh], brute)
    d = np.array(dims, np.uint32)
    load().orc_voxelize(arr, _p(lo), _p(hi), _p(d), _p(data))
    return dims, data.reshape(dims[2], dims[1], dims[0])


def splat_points(gbuf, rgba, proj, cs_inv, flags, clouds):
    """Point clouds drawn over a rendered G-buffer and image
    (canvas::render_pointclouds_on_image): clouds is a list of dicts with
    xyz (n x 3), optional normals (n x 3), colors (n uint32), db_id and cs.
    Returns new (gbuf, rgba) arrays."""
    g = np.ascontiguousarray(gbuf, PIXEL_DTYPE).copy()
    im = np.ascontiguousarray(rgba, np.uint32).copy()
    h, w = g.shape
    if not clouds:
        return g, im
    z = np.zeros(h * w, np.float32)
    lib = load()
    lib.orc_splat_zbuffer(_p(g), h * w, _p(z))
    P = np.ascontiguousarray(proj, np.float32)
    ci = np.ascontiguousarray(cs_inv, np.float32)
    for c in clouds:
        xyz = np.ascontiguousarray(c["xyz"], np.float32).reshape(-1, 3)
        nrm = c.get("normals")
        nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        col = c.get("colors")
        col = None if col is None else np.ascontiguousarray(col, np.uint32).reshape(-1)
        cs = c.get("cs")
        cs = np.eye(4, dtype=np.float32).reshape(16) if cs is None else np.ascontiguousarray(cs, np.float32)
        lib.orc_splat_cloud(_p(g), _p(im), w, h, _p(z), _p(P), _p(ci), _p(cs), int(c.get("db_id", 0)), _p(xyz),
                            xyz.shape[0], _p(nrm), _p(col), int(bool(flags & SHADING)), int(bool(flags & ONE_BIT)))
    return g, im


def vox_color_index(r, g, b):
    return int(load().orc_vox_color_index(int(r), int(g), int(b)))


def any_hits(meshes, rays, threads=0, brute=False):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hit = np.zeros(r.shape[0], np.uint8)
    arr, no = _objs(meshes, brute)
    load().orc_any_hits(arr, no, _p(r), r.shape[0], int(threads), _p(hit))
    return hit


def intersect_triangle(o, d, v0, v1, v2, t_near=0.0, t_far=3.4028234663852886e38):
    a = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = load().orc_intersect_triangle(*[_p(x) for x in a], float(t_near), float(t_far), C.byref(t), C.byref(u),
                                        C.byref(v))
    return bool(hit), t.value, u.value, v.value


# ---- learned proposal network (mlp_oracle.cpp) ----------------------------

def mlp_param_count():
    return int(load().orc_mlp_param_count())


def mlp_line(rays, norm):
    """Per ray: (|d|, a, x0..x5), the network's line input (mlp_oracle.cpp)."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    out = np.zeros((r.shape[0], 8), np.float32)
    load().orc_mlp_line(_p(r), r.shape[0], _p(np.ascontiguousarray(norm, np.float32)), _p(out))
    return out


def mlp_features(rays, norm):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    out = np.zeros((r.shape[0], 78), np.float32)
    load().orc_mlp_features(_p(r), r.shape[0], _p(np.ascontiguousarray(norm, np.float32)), _p(out))
    return out


def mlp_forward(params, norm, rays, bf16=True):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    out = np.zeros((r.shape[0], 2), np.float32)
    load().orc_mlp_forward(_p(np.ascontiguousarray(params, np.float32)), _p(np.ascontiguousarray(norm, np.float32)),
                           _p(r), r.shape[0], int(bool(bf16)), _p(out))
    return out


def mlp_gradients(params, norm, rays, labels, bf16=True):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    lab = np.ascontiguousarray(labels, np.float32).reshape(-1, 2)
    g = np.zeros(mlp_param_count(), np.float32)
    loss = load().orc_mlp_gradients(_p(np.ascontiguousarray(params, np.float32)),
                                    _p(np.ascontiguousarray(norm, np.float32)), _p(r), _p(lab), r.shape[0],
                                    int(bool(bf16)), _p(g))
    return g, float(loss)

