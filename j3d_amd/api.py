"""Python host API over libj3dg.so: meshes, frames, targets, ray queries.

Mirrors the reference's render sequence (j3d/view.cpp:421-430): a Renderer
holds the device context (canvas + scene state on the GPU), `render` runs
canvas::render_scene and canvas::canvas_to_image as HIP kernels and, on
request, reads the G-buffer (`pixel`, j3d/pixel.h) and RGBA image back.
"""
import ctypes as C
import os
import math

import numpy as np

from . import ffi

MESH_KEY = 1


def make_db_id(key, index):
    """j3d/db.h:25-28"""
    return ((key << 29) | index) & 0xFFFFFFFF


def _matrix16(m, what):
    """A 4x4 matrix as 16 contiguous float32 values (the C side reads 16)."""
    a = np.ascontiguousarray(m, np.float32).reshape(-1)
    if a.size != 16:
        raise ValueError("%s: a 4x4 matrix has 16 values, got %d" % (what, a.size))
    return a


class J3dgError(RuntimeError):
    pass


def _check(lib, ctx, rc, what):
    if rc != ffi.J3DG_OK:
        msg = lib.j3dg_last_error(ctx) if ctx else b""
        raise J3dgError("%s failed: %s (%s)" % (what, ffi.ERRORS.get(rc, rc), (msg or b"").decode()))


# ---------------------------------------------------------------- camera ----

def default_camera():
    lib = ffi.load()
    c = ffi.Camera()
    lib.j3dg_default_camera(C.byref(c))
    return c


def camera_with_vertical_fov(fov_deg):
    """Reference camera model (j3d/camera.cpp:5-16) with the focal length that
    gives the requested vertical field of view on the 0.768-inch film."""
    c = default_camera()
    half_h_mm = c.film_h_inch * 25.4 / 2.0
    c.focal_length = float(np.float32(half_h_mm / math.tan(math.radians(fov_deg) / 2.0)))
    return c


def projection(camera, w, h):
    """(P, P^-1) as float32[16] (j3d/camera.cpp:26-73)."""
    lib = ffi.load()
    P = np.zeros(16, np.float32)
    Pi = np.zeros(16, np.float32)
    lib.j3dg_projection(C.byref(camera), int(w), int(h), ffi.ptr(P))
    lib.j3dg_invert_projection(ffi.ptr(P), ffi.ptr(Pi))
    return P, Pi


def invert_orthonormal(m):
    lib = ffi.load()
    m = np.ascontiguousarray(m, np.float32)
    o = np.zeros(16, np.float32)
    lib.j3dg_invert_orthonormal(ffi.ptr(m), ffi.ptr(o))
    return o


def orbit_camera(dist, azimuth, elevation, target=(0.0, 0.0, 0.0)):
    lib = ffi.load()
    t = np.asarray(target, np.float32)
    cs = np.zeros(16, np.float32)
    ci = np.zeros(16, np.float32)
    lib.j3dg_orbit_camera(float(dist), float(azimuth), float(elevation), ffi.ptr(t), ffi.ptr(cs), ffi.ptr(ci))
    return cs, ci


def translation_camera(x, y, z):
    """Camera at (x, y, z) looking down -z (the unzoom convention, j3d/scene.cpp:91-110)."""
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[12], cs[13], cs[14] = x, y, z
    return cs, invert_orthonormal(cs)


def make_matcap(kind=0):
    lib = ffi.load()
    im = np.zeros(512 * 512, np.uint32)
    cav = C.c_uint32(0)
    rc = lib.j3dg_make_matcap(int(kind), ffi.ptr(im), C.byref(cav))
    if rc:
        raise J3dgError("unknown matcap type %r" % kind)
    return im.reshape(512, 512), int(cav.value)


# -------------------------------------------------------------- scenegen ----

def icosphere(level, seed=1, amplitude=0.05):
    lib = ffi.load()
    nv = lib.j3dg_icosphere_vertex_count(level)
    nt = lib.j3dg_icosphere_triangle_count(level)
    v = np.zeros((nv, 3), np.float32)
    t = np.zeros((nt, 3), np.uint32)
    rc = lib.j3dg_gen_icosphere(int(level), int(seed), float(amplitude), ffi.ptr(v), ffi.ptr(t))
    if rc:
        raise J3dgError("icosphere generation failed")
    return v, t


def torus(nu, nv, R=1.0, r=0.35, seed=1, amplitude=0.02, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4.0,
          threads=0):
    lib = ffi.load()
    v = np.zeros((nu * nv, 3), np.float32)
    t = np.zeros((2 * nu * nv, 3), np.uint32)
    rc = lib.j3dg_gen_torus(int(nu), int(nv), float(R), float(r), int(seed), float(amplitude), int(octaves),
                            float(lacunarity), float(gain), float(base_freq), int(threads), ffi.ptr(v), ffi.ptr(t))
    if rc:
        raise J3dgError("torus generation failed")
    return v, t


def random_hemisphere(seed):
    lib = ffi.load()
    o = np.zeros(3, np.float32)
    lib.j3dg_random_hemisphere(int(seed), ffi.ptr(o))
    return o


def bbox(v):
    """j3d/mesh.cpp:39-58"""
    return v.min(axis=0).astype(np.float32), v.max(axis=0).astype(np.float32)


def scene_framing(v, cs=None):
    """prepare_scene (j3d/scene.cpp:50-89) for one object: (diagonal, pivot)."""
    lo, hi = bbox(v)
    if cs is not None:
        m = np.asarray(cs, np.float32).reshape(4, 4).T
        a = (m[:3, :3] @ lo + m[:3, 3]).astype(np.float32)
        b = (m[:3, :3] @ hi + m[:3, 3]).astype(np.float32)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
    ext = hi - lo
    diag = np.float32(max(ext[0], ext[1], ext[2]))
    pivot = ((lo + hi) * np.float32(0.5)).astype(np.float32)
    return float(diag), pivot


# ------------------------------------------------------------------ frame ----

class Frame:
    """Python view of j3dg_frame_desc."""

    def __init__(self, width, height, proj_inv, cs, cs_inv, pivot, diagonal, objects, flags=ffi.EDGES | ffi.SHADING,
                 near_plane=0.1, light_dir=(0.0, 0.0, 1.0), bg_top=0xff000000, bg_bottom=0xff404040, rect=None,
                 margin=0.0):
        self.width, self.height = int(width), int(height)
        self.proj_inv = np.asarray(proj_inv, np.float32)
        self.cs = np.asarray(cs, np.float32)
        self.cs_inv = np.asarray(cs_inv, np.float32)
        self.pivot = np.asarray(pivot, np.float32)
        self.diagonal = float(np.float32(diagonal))
        self.objects = list(objects)
        self.flags = int(flags)
        self.near_plane = float(np.float32(near_plane))
        self.light_dir = np.asarray(light_dir, np.float32)
        self.bg_top, self.bg_bottom = int(bg_top), int(bg_bottom)
        self.rect = rect if rect is not None else (0, 0, self.width - 1, self.height - 1)
        self.margin = float(margin)

    def desc(self):
        d = ffi.FrameDesc()
        d.width, d.height = self.width, self.height
        d.x0, d.y0, d.x1, d.y1 = [int(x) for x in self.rect]
        d.proj_inv[:] = [float(x) for x in self.proj_inv]
        d.near_plane = self.near_plane
        d.cs[:] = [float(x) for x in self.cs]
        d.cs_inv[:] = [float(x) for x in self.cs_inv]
        d.pivot[:] = [float(x) for x in self.pivot]
        d.diagonal = self.diagonal
        d.flags = self.flags
        d.light_dir[:] = [float(x) for x in self.light_dir]
        d.bg_top, d.bg_bottom = self.bg_top, self.bg_bottom
        d.n_objects = len(self.objects)
        if len(self.objects) > ffi.MAX_OBJECTS:
            # more objects than objects[] holds: the slots go by object_list (kept alive with the struct)
            lst = np.ascontiguousarray(self.objects, np.uint32)
            d.object_list = lst.ctypes.data
            d._object_list = lst
        else:
            for i, s in enumerate(self.objects):
                d.objects[i] = int(s)
        d.margin = self.margin
        return d


def make_frame(width, height, camera, cs, cs_inv, pivot, diagonal, objects, **kw):
    _, Pi = projection(camera, width, height)
    return Frame(width, height, Pi, cs, cs_inv, pivot, diagonal, objects, near_plane=camera.near_plane, **kw)


# --------------------------------------------------------------- renderer ----

PROPOSAL_MAGIC = b"J3DGMLP\x00"
PROPOSAL_VERSION = 1
PROPOSAL_LAYERS = (78, 64, 64, 64, 2)


def write_proposal(path, params, norm):
    """Versioned binary checkpoint of a proposal network: 8-byte magic, uint32
    version, uint32 layer count + sizes, 4 float32 normalisation values, then
    the float32 parameters (W1, b1, ..., W4, b4), all little-endian."""
    p = np.ascontiguousarray(params, "<f4")
    nrm = np.ascontiguousarray(norm, "<f4")
    if p.size != ffi.MLP_PARAMS or nrm.size != 4:
        raise ValueError("proposal network: %d parameters and 4 normalisation values expected, got %d and %d"
                         % (ffi.MLP_PARAMS, p.size, nrm.size))
    with open(path, "wb") as f:
        f.write(PROPOSAL_MAGIC)
        f.write(np.array([PROPOSAL_VERSION, len(PROPOSAL_LAYERS)] + list(PROPOSAL_LAYERS), "<u4").tobytes())
        f.write(nrm.tobytes())
        f.write(p.tobytes())


def read_proposal(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != PROPOSAL_MAGIC:
        raise ValueError("not a proposal checkpoint: %s" % path)
    head = np.frombuffer(data, "<u4", 2, 8)
    if int(head[0]) != PROPOSAL_VERSION:
        raise ValueError("unsupported proposal checkpoint version %d" % int(head[0]))
    nl = int(head[1])
    layers = tuple(int(x) for x in np.frombuffer(data, "<u4", nl, 16))
    if layers != PROPOSAL_LAYERS:
        raise ValueError("checkpoint network %s differs from %s" % (layers, PROPOSAL_LAYERS))
    off = 16 + 4 * nl
    nrm = np.frombuffer(data, "<f4", 4, off).copy()
    p = np.frombuffer(data, "<f4", ffi.MLP_PARAMS, off + 16).copy()
    if len(data) != off + 16 + 4 * ffi.MLP_PARAMS:
        raise ValueError("truncated or oversized proposal checkpoint")
    return p, nrm


class Renderer:
    """One j3dg context on one GPU."""

    def __init__(self, device=0):
        self.lib = ffi.load()
        h = C.c_void_p()
        rc = self.lib.j3dg_create(int(device), C.byref(h))
        _check(self.lib, None, rc, "j3dg_create")
        self.ctx = h
        self._targets = {}

    def close(self):
        if self.ctx:
            self.lib.j3dg_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        _check(self.lib, self.ctx, rc, what)

    def upload_mesh(self, verts, tris, db_id=None, cs=None, width=8, leaf_size=0, threads=0, builder=0):
        """builder: ffi.BUILD_HOST (binned SAH on the host) or ffi.BUILD_DEVICE
        (normals and a linear BVH built on the GPU from the raw arrays)."""
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        o = ffi.BuildOpts(int(width), int(leaf_size), int(threads), int(builder))
        slot = C.c_uint32(0)
        csa = None if cs is None else np.ascontiguousarray(cs, np.float32)
        if db_id is None:
            db_id = make_db_id(MESH_KEY, 0)
        self._c(self.lib.j3dg_upload_mesh(self.ctx, int(db_id), ffi.ptr(v), v.shape[0], ffi.ptr(t), t.shape[0],
                                          ffi.ptr(csa), C.byref(o), C.byref(slot)), "j3dg_upload_mesh")
        return int(slot.value)

    def remove_mesh(self, slot):
        self._c(self.lib.j3dg_remove_mesh(self.ctx, int(slot)), "j3dg_remove_mesh")

    def set_mesh_cs(self, slot, cs):
        a = np.ascontiguousarray(cs, np.float32)
        self._c(self.lib.j3dg_set_mesh_cs(self.ctx, int(slot), ffi.ptr(a)), "j3dg_set_mesh_cs")

    def set_mesh_colors(self, slot, vertex_colors=None, uv=None, texture=None):
        """Surface colour of a mesh (j3dg_set_mesh_colors): rgb per vertex
        (n_vertices x 3 floats), (u, v) per triangle corner (n_triangles x 6
        floats) with a (h, w) uint32 texture; None removes a source."""
        vc = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, np.float32).reshape(-1, 3)
        uvs = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
        tex = None if texture is None else np.ascontiguousarray(texture, np.uint32)
        th, tw = (0, 0) if tex is None else tex.shape
        self._c(self.lib.j3dg_set_mesh_colors(self.ctx, int(slot), ffi.ptr(vc), ffi.ptr(uvs), ffi.ptr(tex), int(tw),
                                              int(th)), "j3dg_set_mesh_colors")

    def mesh_info(self, slot):
        mi = ffi.MeshInfo()
        self._c(self.lib.j3dg_get_mesh_info(self.ctx, int(slot), C.byref(mi)), "j3dg_get_mesh_info")
        return {k: (list(getattr(mi, k)) if k.startswith("bb_") else getattr(mi, k)) for k, _ in mi._fields_}

    def mesh_normals(self, slot, n_tris):
        out = np.zeros((n_tris, 3), np.float32)
        self._c(self.lib.j3dg_mesh_normals(self.ctx, int(slot), ffi.ptr(out)), "j3dg_mesh_normals")
        return out

    def set_matcap(self, im, cavity):
        a = np.ascontiguousarray(im, np.uint32)
        h, w = a.shape
        self._c(self.lib.j3dg_set_matcap(self.ctx, ffi.ptr(a), w, h, int(cavity)), "j3dg_set_matcap")

    def create_target(self, w, h):
        t = C.c_uint32(0)
        self._c(self.lib.j3dg_create_target(self.ctx, int(w), int(h), C.byref(t)), "j3dg_create_target")
        self._targets[int(t.value)] = (int(w), int(h))
        return int(t.value)

    def destroy_target(self, t):
        self._c(self.lib.j3dg_destroy_target(self.ctx, int(t)), "j3dg_destroy_target")
        self._targets.pop(int(t), None)

    def write_gbuffer(self, t, gbuf):
        a = np.ascontiguousarray(gbuf, ffi.PIXEL_DTYPE)
        self._c(self.lib.j3dg_write_target_gbuffer(self.ctx, int(t), ffi.ptr(a)), "j3dg_write_target_gbuffer")

    def render(self, frame, target, stages=ffi.STAGE_ALL, readback=True):
        d = frame.desc()
        w, h = self._targets[int(target)]
        if readback:
            g = np.zeros((h, w), ffi.PIXEL_DTYPE)
            im = np.zeros((h, w), np.uint32)
            self._c(self.lib.j3dg_render(self.ctx, C.byref(d), int(target), int(stages), ffi.ptr(g), ffi.ptr(im)),
                    "j3dg_render")
            return g, im
        self._c(self.lib.j3dg_render(self.ctx, C.byref(d), int(target), int(stages), None, None), "j3dg_render")
        return None

    def read_target(self, t):
        w, h = self._targets[int(t)]
        g = np.zeros((h, w), ffi.PIXEL_DTYPE)
        im = np.zeros((h, w), np.uint32)
        self._c(self.lib.j3dg_read_target(self.ctx, int(t), ffi.ptr(g), ffi.ptr(im)), "j3dg_read_target")
        return g, im

    def read_target_strided(self, t, gbuf_stride, rgba_stride):
        """j3dg_read_target_strided into images of h rows of the given strides
        (in elements); returns the padded arrays (padding left as zeros)."""
        w, h = self._targets[int(t)]
        g = np.zeros((h, gbuf_stride), ffi.PIXEL_DTYPE)
        im = np.zeros((h, rgba_stride), np.uint32)
        self._c(self.lib.j3dg_read_target_strided(self.ctx, int(t), ffi.ptr(g), int(gbuf_stride), ffi.ptr(im),
                                                  int(rgba_stride)), "j3dg_read_target_strided")
        return g, im

    def render_batch(self, frames, targets):
        arr = (ffi.FrameDesc * len(frames))(*[f.desc() for f in frames])
        tg = np.ascontiguousarray(targets, np.uint32)
        self._c(self.lib.j3dg_render_batch(self.ctx, arr, ffi.ptr(tg), len(frames)), "j3dg_render_batch")

    def synchronize(self):
        self._c(self.lib.j3dg_synchronize(self.ctx), "j3dg_synchronize")

    def set_tuning(self, stack_depth=16, rays_per_item=1, coop_tail=4, block_size=256, refill_min=16, split_passes=0,
                   cost_order=1):
        t = ffi.Tuning(stack_depth, rays_per_item, coop_tail, block_size, refill_min, split_passes, cost_order)
        self._c(self.lib.j3dg_set_tuning(self.ctx, C.byref(t)), "j3dg_set_tuning")

    def get_tuning(self):
        t = ffi.Tuning()
        self._c(self.lib.j3dg_get_tuning(self.ctx, C.byref(t)), "j3dg_get_tuning")
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_counting(self, on):
        self._c(self.lib.j3dg_set_counting(self.ctx, 1 if on else 0), "j3dg_set_counting")

    def stats(self):
        s = ffi.Stats()
        self._c(self.lib.j3dg_get_stats(self.ctx, C.byref(s)), "j3dg_get_stats")
        out = {}
        for k, _ in s._fields_:
            v = getattr(s, k)
            out[k] = list(v) if k.endswith("_log2") else v
        return out

    def reset_timings(self):
        self._c(self.lib.j3dg_reset_timings(self.ctx), "j3dg_reset_timings")

    def timings(self, max_n=1024):
        """Per-render kernel durations (ms) since the last reset: dict of arrays."""
        tr = np.zeros(max_n, np.float32)
        sh = np.zeros(max_n, np.float32)
        ml = np.zeros(max_n, np.float32)
        sd = np.zeros(max_n, np.float32)
        n = C.c_uint32(0)
        self._c(self.lib.j3dg_get_timings(self.ctx, ffi.ptr(tr), ffi.ptr(sh), ffi.ptr(ml), ffi.ptr(sd), max_n,
                                          C.byref(n)), "j3dg_get_timings")
        k = n.value
        return {"trace_ms": tr[:k], "shade_ms": sh[:k], "mlp_ms": ml[:k], "shadow_ms": sd[:k]}

    def debug_timeline(self, max_waves=1 << 16):
        """Per-wave timeline of the last counting render (see j3dg_debug_timeline)."""
        out = np.zeros((max_waves, 8), np.uint64)
        n = C.c_uint32(0)
        self._c(self.lib.j3dg_debug_timeline(self.ctx, ffi.ptr(out), max_waves, C.byref(n)), "j3dg_debug_timeline")
        return out[:n.value]

    def calibration_copy(self, nbytes, repeats=1):
        ms = C.c_float(0)
        self._c(self.lib.j3dg_calibration_copy(self.ctx, int(nbytes), int(repeats), C.byref(ms)),
                "j3dg_calibration_copy")
        return ms.value

    def trace_rays(self, slots, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        s = np.ascontiguousarray(slots, np.uint32)
        out = {"obj": np.zeros(n, np.uint32), "tri": np.zeros(n, np.uint32), "t": np.zeros(n, np.float32),
               "u": np.zeros(n, np.float32), "v": np.zeros(n, np.float32)}
        self._c(self.lib.j3dg_trace_rays(self.ctx, ffi.ptr(s), len(s), ffi.ptr(r), n, ffi.ptr(out["obj"]),
                                         ffi.ptr(out["tri"]), ffi.ptr(out["t"]), ffi.ptr(out["u"]),
                                         ffi.ptr(out["v"])), "j3dg_trace_rays")
        return out

    def autotune(self, frame, target, repeats=5, cache=None, key=None):
        """Autotunes stack depth x rays per item, then the refill threshold (and
        the margin of a learned frame) on `frame` (j3dg_autotune) and applies
        the result.  With `cache`
        (a JSON file path) results are stored and reused under `key`, e.g.
        "<triangles>@<width>x<height>"."""
        import json
        if cache and key and os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
            if key in db:
                e = db[key]
                self.set_tuning(**e["tuning"])
                return dict(e, cached=True)
        rep = ffi.AutotuneReport()
        d = frame.desc()
        self._c(self.lib.j3dg_autotune(self.ctx, C.byref(d), int(target), int(repeats), C.byref(rep)), "j3dg_autotune")
        out = self._report(rep)
        if cache and key:
            db = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    db = json.load(f)
            db[key] = {k: v for k, v in out.items() if k != "cached"}
            with open(cache, "w") as f:
                json.dump(db, f, indent=1, sort_keys=True)
        return out

    @staticmethod
    def _report(rep):
        return {"tuning": {k: getattr(rep.best, k) for k, _ in ffi.Tuning._fields_}, "margin": rep.best_margin,
                "best_ms": rep.best_ms, "baseline_ms": rep.baseline_ms, "n_tried": rep.n_tried,
                "width": rep.width, "cached": False}

    def autotune_widths(self, frame_w4, frame_w8, target, repeats=5):
        """j3dg_autotune_widths: the same frame over width-4 and width-8 uploads
        of its meshes (either may be None); the faster tuning is applied."""
        rep = ffi.AutotuneReport()
        d4 = frame_w4.desc() if frame_w4 is not None else None
        d8 = frame_w8.desc() if frame_w8 is not None else None
        self._c(self.lib.j3dg_autotune_widths(self.ctx, C.byref(d4) if d4 is not None else None,
                                              C.byref(d8) if d8 is not None else None, int(target), int(repeats),
                                              C.byref(rep)), "j3dg_autotune_widths")
        return self._report(rep)

    def autotune_batch(self, frames_w4, frames_w8, targets, repeats=3):
        """j3dg_autotune_batch: BVH width x stack depth x rays per work item
        on one graph launch of the batch; applies the fastest tuning."""
        n = len(targets)
        a4 = (ffi.FrameDesc * n)(*[f.desc() for f in frames_w4]) if frames_w4 is not None else None
        a8 = (ffi.FrameDesc * n)(*[f.desc() for f in frames_w8]) if frames_w8 is not None else None
        tg = np.ascontiguousarray(targets, np.uint32)
        rep = ffi.AutotuneReport()
        self._c(self.lib.j3dg_autotune_batch(self.ctx, a4, a8, ffi.ptr(tg), n, int(repeats), C.byref(rep)),
                "j3dg_autotune_batch")
        return self._report(rep)

    # ---- learned proposal ------------------------------------------------
    def train_proposal(self, slot, **opts):
        """Trains the mesh's tmax proposal network on the GPU (j3dg_train_proposal);
        keyword arguments override j3dg_default_train_opts fields.  Returns the report."""
        o = ffi.TrainOpts()
        self._c(self.lib.j3dg_default_train_opts(C.byref(o)), "j3dg_default_train_opts")
        for k, v in opts.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        rep = ffi.TrainReport()
        self._c(self.lib.j3dg_train_proposal(self.ctx, int(slot), C.byref(o), C.byref(rep)), "j3dg_train_proposal")
        out = {k: getattr(rep, k) for k, _ in ffi.TrainReport._fields_}
        out["val_resolved"] = list(rep.val_resolved)
        out["norm"] = list(rep.norm)
        return out

    def get_proposal(self, slot):
        p = np.zeros(ffi.MLP_PARAMS, np.float32)
        nrm = np.zeros(4, np.float32)
        self._c(self.lib.j3dg_get_proposal(self.ctx, int(slot), ffi.ptr(p), ffi.ptr(nrm)), "j3dg_get_proposal")
        return p, nrm

    def set_proposal(self, slot, params, norm):
        if params is None:
            self._c(self.lib.j3dg_set_proposal(self.ctx, int(slot), None, None), "j3dg_set_proposal")
            return
        p = np.ascontiguousarray(params, np.float32)
        nrm = np.ascontiguousarray(norm, np.float32)
        if p.size != ffi.MLP_PARAMS or nrm.size != 4:
            raise ValueError("set_proposal: %d parameters and 4 normalisation values expected, got %d and %d"
                             % (ffi.MLP_PARAMS, p.size, nrm.size))
        self._c(self.lib.j3dg_set_proposal(self.ctx, int(slot), ffi.ptr(p), ffi.ptr(nrm)), "j3dg_set_proposal")

    def save_proposal(self, slot, path):
        """Checkpoint: magic, version, layer sizes, normalisation, parameters (little-endian)."""
        p, nrm = self.get_proposal(slot)
        write_proposal(path, p, nrm)

    def load_proposal(self, slot, path):
        p, nrm = read_proposal(path)
        self.set_proposal(slot, p, nrm)

    def write_proposal(self, target, pairs):
        w, h = self._targets[int(target)]
        a = np.ascontiguousarray(pairs, np.float32).reshape(h, w, 2)
        self._c(self.lib.j3dg_write_target_proposal(self.ctx, int(target), ffi.ptr(a)), "j3dg_write_target_proposal")

    def read_proposal(self, target):
        w, h = self._targets[int(target)]
        a = np.zeros((h, w, 2), np.float32)
        self._c(self.lib.j3dg_read_target_proposal(self.ctx, int(target), ffi.ptr(a)), "j3dg_read_target_proposal")
        return a

    def mlp_forward(self, params, norm, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((r.shape[0], 2), np.float32)
        p = np.ascontiguousarray(params, np.float32)
        nrm = np.ascontiguousarray(norm, np.float32)
        self._c(self.lib.j3dg_mlp_forward(self.ctx, ffi.ptr(p), ffi.ptr(nrm), ffi.ptr(r), r.shape[0], ffi.ptr(out)),
                "j3dg_mlp_forward")
        return out

    def mlp_gradients(self, params, norm, rays, labels):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        lab = np.ascontiguousarray(labels, np.float32).reshape(-1, 2)
        p = np.ascontiguousarray(params, np.float32)
        nrm = np.ascontiguousarray(norm, np.float32)
        g = np.zeros(ffi.MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._c(self.lib.j3dg_mlp_gradients(self.ctx, ffi.ptr(p), ffi.ptr(nrm), ffi.ptr(r), ffi.ptr(lab), r.shape[0],
                                            ffi.ptr(g), C.byref(loss)), "j3dg_mlp_gradients")
        return g, loss.value

    def mlp_adam_step(self, params, m, v, grads, step, lr):
        p = np.array(params, np.float32)
        mm = np.array(m, np.float32)
        vv = np.array(v, np.float32)
        g = np.ascontiguousarray(grads, np.float32)
        self._c(self.lib.j3dg_mlp_adam_step(self.ctx, ffi.ptr(p), ffi.ptr(mm), ffi.ptr(vv), ffi.ptr(g), int(step),
                                            float(lr)), "j3dg_mlp_adam_step")
        return p, mm, vv

    def any_hit(self, slots, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        s = np.ascontiguousarray(slots, np.uint32)
        hit = np.zeros(r.shape[0], np.uint8)
        self._c(self.lib.j3dg_any_hit_rays(self.ctx, ffi.ptr(s), len(s), ffi.ptr(r), r.shape[0], ffi.ptr(hit)),
                "j3dg_any_hit_rays")
        return hit

    def all_hits(self, slot, rays, max_hits):
        """Every hit of each ray in one mesh, (t, triangle id) order
        (j3dg_all_hits): {"count": n, "tri"/"t"/"u"/"v": n x max_hits}."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        out = {"count": np.zeros(n, np.uint32), "tri": np.zeros((n, max_hits), np.uint32),
               "t": np.zeros((n, max_hits), np.float32), "u": np.zeros((n, max_hits), np.float32),
               "v": np.zeros((n, max_hits), np.float32)}
        self._c(self.lib.j3dg_all_hits(self.ctx, int(slot), ffi.ptr(r), n, int(max_hits), ffi.ptr(out["count"]),
                                       ffi.ptr(out["tri"]), ffi.ptr(out["t"]), ffi.ptr(out["u"]),
                                       ffi.ptr(out["v"])), "j3dg_all_hits")
        return out

    def voxelize(self, slot, max_dim=100):
        """write_vox's voxel grid of a mesh (j3dg_voxelize): (dims, uint8 array [z, y, x])."""
        dim = np.zeros(3, np.uint32)
        self._c(self.lib.j3dg_voxelize(self.ctx, int(slot), int(max_dim), ffi.ptr(dim), None, 0), "j3dg_voxelize")
        n = int(dim[0]) * int(dim[1]) * int(dim[2])
        data = np.zeros(n, np.uint8)
        self._c(self.lib.j3dg_voxelize(self.ctx, int(slot), int(max_dim), ffi.ptr(dim), ffi.ptr(data), n),
                "j3dg_voxelize")
        dims = tuple(int(x) for x in dim)
        return dims, data.reshape(dims[2], dims[1], dims[0])

    def upload_points(self, xyz, normals=None, colors=None, db_id=0, cs=None):
        """A point cloud (j3dg_upload_points): n x 3 points, optional n x 3
        normals and n uint32 colours (0xAABBGGRR); returns its cloud id."""
        p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        col = None if colors is None else np.ascontiguousarray(colors, np.uint32).reshape(-1)
        if nrm is not None and nrm.shape[0] != p.shape[0]:
            raise ValueError("upload_points: %d normals for %d points" % (nrm.shape[0], p.shape[0]))
        if col is not None and col.shape[0] != p.shape[0]:
            raise ValueError("upload_points: %d colours for %d points" % (col.shape[0], p.shape[0]))
        csa = None if cs is None else _matrix16(cs, "upload_points: cs")
        cloud = C.c_uint32(0)
        self._c(self.lib.j3dg_upload_points(self.ctx, int(db_id), ffi.ptr(p), p.shape[0], ffi.ptr(nrm), ffi.ptr(col),
                                            ffi.ptr(csa), C.byref(cloud)), "j3dg_upload_points")
        return int(cloud.value)

    def remove_points(self, cloud):
        self._c(self.lib.j3dg_remove_points(self.ctx, int(cloud)), "j3dg_remove_points")

    def set_points_cs(self, cloud, cs):
        a = _matrix16(cs, "set_points_cs: cs")
        self._c(self.lib.j3dg_set_points_cs(self.ctx, int(cloud), ffi.ptr(a)), "j3dg_set_points_cs")

    def splat_points(self, target, proj, cs_inv, flags, clouds):
        """Draws point clouds over a rendered target (j3dg_splat_points,
        canvas::render_pointclouds_on_image); read the result with read_target."""
        P = _matrix16(proj, "splat_points: proj")
        ci = _matrix16(cs_inv, "splat_points: cs_inv")
        lst = np.ascontiguousarray(list(clouds), np.uint32)
        self._c(self.lib.j3dg_splat_points(self.ctx, int(target), ffi.ptr(P), ffi.ptr(ci), int(flags), ffi.ptr(lst),
                                           len(lst)), "j3dg_splat_points")
