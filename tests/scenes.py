"""Workloads shared by the tests and the benchmark, with the comparison rules
used for parity against the CPU oracle."""
import math

import numpy as np

import j3d_amd as J
from j3d_amd import ffi

MISS = 0xFFFFFFFF


def c0_scene(level=7, w=1024, h=768, seed=1):
    """Seeded icosphere (level 7: 163,842 vertices, 327,680 triangles), camera at
    (0,0,3) looking at the origin, 60 degree vertical fov, edges + shading."""
    v, t = J.icosphere(level, seed=seed, amplitude=0.05)
    diag, pivot = J.scene_framing(v)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    cam = J.camera_with_vertical_fov(60.0)
    return v, t, cam, cs, ci, pivot, diag


def torus_scene(nu, nv, seed=1, azimuth_seed=11):
    """Seeded displaced torus (R=1, r=0.35, fBm 6 octaves amplitude 0.02,
    lacunarity 2, gain 0.5) and an orbit camera at distance 3, 60 degree fov."""
    v, t = J.torus(nu, nv, seed=seed)
    diag, pivot = J.scene_framing(v)
    st = np.array([azimuth_seed], np.uint64)
    lib = ffi.load()
    import ctypes as C
    s = C.c_uint64(azimuth_seed)
    az = 2.0 * math.pi * lib.j3dg_uniform01(C.byref(s))
    el = math.radians(20.0 + 30.0 * lib.j3dg_uniform01(C.byref(s)))
    del st
    cs, ci = J.orbit_camera(3.0, az, el)
    cam = J.camera_with_vertical_fov(60.0)
    light = J.random_hemisphere(azimuth_seed + 1)
    return v, t, cam, cs, ci, pivot, diag, light


def compare_gbuffers(g, og):
    """Bit-exact comparison of every G-buffer field.  Returns a dict of mismatch counts."""
    out = {}
    for k in ("mark", "r", "g", "b", "object_id", "db_id"):
        out[k] = int((g[k] != og[k]).sum())
    for k in ("u", "v", "depth", "barycentric_u", "barycentric_v"):
        out[k] = int((g[k].view(np.uint32) != og[k].view(np.uint32)).sum())
    return out


def compare_rgba(im, oim):
    a = im.view(np.uint8).astype(np.int16)
    b = oim.view(np.uint8).astype(np.int16)
    d = np.abs(a - b)
    return int(d.max()) if d.size else 0, int((d.reshape(im.shape + (4,)).max(axis=-1) > 0).sum())


def random_rays(n, seed, center, radius, t_near=1e-3):
    """Rays from a sphere of radius 2*radius around center toward random points
    near the object (world space, 8 floats each)."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    o = center + 2.0 * radius * o
    tgt = center + radius * rng.uniform(-1, 1, size=(n, 3))
    d = tgt - o
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 6] = t_near
    r[:, 7] = np.finfo(np.float32).max
    return r


def pick_crop(gbuf, side):
    """The side x side window, on a half-overlapping grid of positions, whose
    hit fraction is closest to 3/4: mostly surface, with some silhouette."""
    h, w = gbuf.shape
    side = min(side, w, h)
    hit = (gbuf["object_id"] != 0xFFFFFFFF).astype(np.int64)
    ii = np.zeros((h + 1, w + 1), np.int64)
    ii[1:, 1:] = hit.cumsum(0).cumsum(1)
    best = None
    step = max(1, side // 2)
    for y0 in range(0, h - side + 1, step):
        for x0 in range(0, w - side + 1, step):
            n = ii[y0 + side, x0 + side] - ii[y0, x0 + side] - ii[y0 + side, x0] + ii[y0, x0]
            d = abs(n / float(side * side) - 0.75)
            if best is None or d < best[0]:
                best = (d, x0, y0)
    _, x0, y0 = best
    return (x0, y0, x0 + side - 1, y0 + side - 1)


def shadow_rays(v, t, g, light_dir):
    """The directional-light shadow rays of a frame's hit pixels, rebuilt from
    its G-buffer for a mesh drawn with the identity pose: origin = the hit
    point interpolated from the triangle's vertices with the stored
    barycentrics, in the order and float32 rounding finish_hit uses
    (j3d/canvas.cpp:838-848), direction = the light, t in [1e-3, FLT_MAX].
    Returns (rays n x 8, flat indices of the hit pixels)."""
    flat = g.reshape(-1)
    idx = np.nonzero(flat["object_id"] != MISS)[0]
    tri = t[flat["object_id"][idx]]
    bu = flat["barycentric_u"][idx].astype(np.float32)
    bv = flat["barycentric_v"][idx].astype(np.float32)
    s0 = (np.float32(1) - bu) - bv
    v0, v1, v2 = (v[tri[:, k]].astype(np.float32) for k in range(3))
    pos = (v0 * s0[:, None] + bu[:, None] * v1) + bv[:, None] * v2
    rays = np.zeros((len(idx), 8), np.float32)
    rays[:, 0:3] = pos
    rays[:, 3:6] = np.asarray(light_dir, np.float32)[:3]
    rays[:, 6] = 1e-3
    rays[:, 7] = np.finfo(np.float32).max
    return rays, idx
