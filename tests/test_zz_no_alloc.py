"""j3dg_trace_rays_device, the ray query that allocates nothing, through the C
ABI: its results on device buffers equal j3dg_trace_rays's on the same rays,
bit for bit, on two calls in a row (the second with fewer rays, so stale queue
words or counters in the shared working array would show), and the entry
point refuses what its fixed working memory cannot take.

Device buffers come from render targets: a target of n x 1 pixels holds 32
bytes per pixel, one ray (8 floats) per pixel, written with
j3dg_write_target_gbuffer and addressed with j3dg_target_device_ptrs."""
import ctypes as C

import numpy as np
import pytest

import j3d_amd as J
from j3d_amd import ffi
from scenes import random_rays

pytestmark = pytest.mark.gpu


def _dev_ptr(r, tg):
    g = C.c_void_p()
    assert r.lib.j3dg_target_device_ptrs(r.ctx, tg, C.byref(g), None) == 0
    return g.value


def _trace_device(r, slots, rays):
    n = rays.shape[0]
    t_in = r.create_target(n, 1)
    t_out = r.create_target(n, 1)
    try:
        r.write_gbuffer(t_in, np.ascontiguousarray(rays, np.float32).view(ffi.PIXEL_DTYPE).reshape(1, n))
        # a sentinel in the outputs: a call that leaves a ray unwritten shows
        sentinel = np.full((1, n), 0xAB, np.uint8).repeat(32, axis=0).T.copy().view(ffi.PIXEL_DTYPE).reshape(1, n)
        r.write_gbuffer(t_out, sentinel)
        din, dout = _dev_ptr(r, t_in), _dev_ptr(r, t_out)
        sl = np.ascontiguousarray(slots, np.uint32)
        rc = r.lib.j3dg_trace_rays_device(r.ctx, ffi.ptr(sl), len(sl), din, n, dout, dout + 4 * n, dout + 8 * n,
                                          dout + 12 * n, dout + 16 * n)
        assert rc == 0, r.lib.j3dg_last_error(r.ctx)
        g, _ = r.read_target(t_out)
    finally:
        r.destroy_target(t_in)
        r.destroy_target(t_out)
    raw = g.view(np.uint8).reshape(-1)[:20 * n]
    out = {"obj": raw[0:4 * n].view(np.uint32), "tri": raw[4 * n:8 * n].view(np.uint32),
           "t": raw[8 * n:12 * n].view(np.float32), "u": raw[12 * n:16 * n].view(np.float32),
           "v": raw[16 * n:20 * n].view(np.float32)}
    return out


def _same(a, b):
    for k in ("obj", "tri"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_device_ray_query_equals_host_query(renderer):
    r = renderer
    tv, tt = J.torus(300, 130, seed=5)
    iv, it = J.icosphere(4, seed=2)
    pose = np.eye(4, dtype=np.float32)
    pose[3, :3] = (0.4, -0.2, 0.3)
    a = r.upload_mesh(tv, tt, db_id=J.make_db_id(1, 0))
    b = r.upload_mesh(iv, it, db_id=J.make_db_id(1, 1), cs=pose.reshape(16))
    try:
        for n, seed in ((60000, 3), (7001, 4)):  # the second call has fewer rays
            rays = random_rays(n, seed, np.zeros(3), 1.4)
            print("no-alloc query: %d rays, seed %d" % (n, seed), flush=True)
            got = _trace_device(r, [a, b], rays)
            ref = r.trace_rays([a, b], rays)
            _same(got, ref)
            assert (got["obj"] != 0xFFFFFFFF).sum() > n // 10
        # the allocating entry point still works after the fixed-memory one
        rays = random_rays(5000, 9, np.zeros(3), 1.4)
        _same(r.trace_rays([a], rays), _trace_device(r, [a], rays))
    finally:
        r.remove_mesh(a)
        r.remove_mesh(b)


def test_device_ray_query_refusals(renderer):
    r = renderer
    v, t = J.icosphere(2, seed=1)
    s = r.upload_mesh(v, t)
    try:
        sl = np.full(17, s, np.uint32)
        assert r.lib.j3dg_trace_rays_device(r.ctx, ffi.ptr(sl), 17, 1, 4, 1, 1, 1, 1, 1) == 1  # > 16 slots
        one = np.array([s], np.uint32)
        assert r.lib.j3dg_trace_rays_device(r.ctx, ffi.ptr(one), 1, None, 4, 1, 1, 1, 1, 1) == 1  # null rays
        assert r.lib.j3dg_trace_rays_device(r.ctx, ffi.ptr(one), 1, 1, 4, 1, None, 1, 1, 1) == 1  # null output
        assert r.lib.j3dg_trace_rays_device(r.ctx, ffi.ptr(one), 1, None, 0, None, None, None, None, None) == 0
    finally:
        r.remove_mesh(s)
