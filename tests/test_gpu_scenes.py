"""GPU parity of scenes with more objects than the frame's object table in LDS
holds (J3DG_MAX_OBJECTS = 16): the top-level BVH over the objects' world boxes
(the reference's qbvh_two_level_with_transformations, j3d/canvas.cpp:762).

Bar as in test_gpu_parity.py: G-buffers bit-exact against the oracle, which
loops over every object (no top level), RGBA within 1 LSB; ray queries
bit-exact.
"""
import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import MISS, compare_gbuffers, compare_rgba, random_rays

pytestmark = pytest.mark.gpu

EDGES = ffi.EDGES | ffi.SHADING


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _pose(rng, spread):
    """Column-major rigid pose (rotation + translation), float32."""
    m = np.eye(4)
    m[:3, :3] = _rotation(rng)
    m[:3, 3] = rng.uniform(-spread, spread, 3)
    return np.ascontiguousarray(m.T, np.float32).reshape(16)


def _bases(small):
    if small:
        return [J.torus(24, 12, seed=3), J.icosphere(2, seed=4, amplitude=0.05), J.torus(16, 30, r=0.2, seed=5)]
    return [J.torus(60, 30, seed=3), J.icosphere(3, seed=4, amplitude=0.05), J.torus(40, 70, r=0.2, seed=5)]


def _scene(renderer, n, seed, spread, small=False, colors=False, width=8, builder=ffi.BUILD_HOST):
    rng = np.random.default_rng(seed)
    bases = _bases(small)
    slots, meshes = [], []
    for i in range(n):
        v, t = bases[i % len(bases)]
        cs = _pose(rng, spread)
        db = J.make_db_id(1, i)
        vc = None
        if colors and i % 2 == 0:
            vc = rng.uniform(0, 1, size=(len(v), 3)).astype(np.float32)
        s = renderer.upload_mesh(v, t, db_id=db, cs=cs, width=width, builder=builder)
        if vc is not None:
            renderer.set_mesh_colors(s, vertex_colors=vc)
        slots.append(s)
        meshes.append(O.Mesh(v, t, db_id=db, cs=cs, vertex_colors=vc))
    return slots, meshes


def _frame(slots, spread, flags, w=320, h=240, az=0.7, el=0.4):
    cs, ci = J.orbit_camera(4.0 * spread + 4.0, az, el)
    cam = J.camera_with_vertical_fov(60.0)
    diag = 2.0 * spread + 2.0
    return J.make_frame(w, h, cam, cs, ci, np.zeros(3, np.float32), diag, slots, flags=flags,
                        light_dir=J.random_hemisphere(9))


def _render_both(r, meshes, frame):
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    tg = r.create_target(frame.width, frame.height)
    try:
        g, im = r.render(frame, tg)
    finally:
        r.destroy_target(tg)
    og, oim, st = O.render(meshes, frame, mc, cav)
    return g, im, og, oim, st


def _assert_parity(g, im, og, oim, label):
    mm = compare_gbuffers(g, og)
    assert sum(mm.values()) == 0, "%s: G-buffer mismatches %s" % (label, mm)
    dmax, ndiff = compare_rgba(im, oim)
    hit = g["object_id"] != MISS
    print("%s: %d hit pixels on %d objects, RGBA max diff %d on %d pixels" % (
        label, int(hit.sum()), len(np.unique(g["db_id"][hit])), dmax, ndiff))
    assert dmax <= 1, "%s: RGBA differs by %d" % (label, dmax)


def _remove(r, slots):
    for s in slots:
        r.remove_mesh(s)


@pytest.mark.parametrize("n, small", [(17, False), (40, False), (300, True)])
def test_many_objects_frame(renderer, n, small):
    spread = 2.0 if n <= 40 else 6.0
    slots, meshes = _scene(renderer, n, 100 + n, spread, small=small)
    try:
        fr = _frame(slots, spread, EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT)
        g, im, og, oim, st = _render_both(renderer, meshes, fr)
        _assert_parity(g, im, og, oim, "%d objects" % n)
        hit = g["object_id"] != MISS
        assert len(np.unique(g["db_id"][hit])) > min(n, 16) // 2
        assert st["shadowed"] > 0
    finally:
        _remove(renderer, slots)


def test_many_objects_split_tuning_and_point_light(renderer):
    """The split pipeline has no top-level variant: such scenes run as one
    launch whatever the tuning says, with the same result."""
    spread = 2.0
    slots, meshes = _scene(renderer, 24, 7, spread)
    tun = renderer.get_tuning()
    renderer.set_tuning(**dict(tun, split_passes=1))
    try:
        fr = _frame(slots, spread, EDGES | ffi.SHADOW, az=2.1, el=0.9)
        g, im, og, oim, st = _render_both(renderer, meshes, fr)
        _assert_parity(g, im, og, oim, "24 objects, point light, split tuning")
    finally:
        renderer.set_tuning(**tun)
        _remove(renderer, slots)


def test_many_objects_rays(renderer):
    spread = 2.0
    slots, meshes = _scene(renderer, 40, 11, spread)
    try:
        rays = random_rays(30000, 5, np.zeros(3), spread + 1.0)
        got = renderer.trace_rays(slots, rays)
        ref = O.closest_hits(meshes, rays)
        for k in ("obj", "tri"):
            assert np.array_equal(got[k], ref[k]), k
        for k in ("t", "u", "v"):
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
        assert np.array_equal(renderer.any_hit(slots, rays), O.any_hits(meshes, rays))
        print("40 objects: %d of %d rays hit, %d objects hit" % (
            int((ref["obj"] != MISS).sum()), len(rays), len(np.unique(ref["obj"][ref["obj"] != MISS]))))
    finally:
        _remove(renderer, slots)


def test_many_objects_colors(renderer):
    spread = 2.0
    slots, meshes = _scene(renderer, 20, 13, spread, colors=True)
    try:
        fr = _frame(slots, spread, EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT | ffi.VERTEXCOLORS)
        g, im, og, oim, st = _render_both(renderer, meshes, fr)
        _assert_parity(g, im, og, oim, "20 objects, vertex colours")
        assert (g["mark"] & 2).any()
    finally:
        _remove(renderer, slots)


def test_many_objects_moved(renderer):
    """Objects moved between frames (j3dg_set_mesh_cs): the top level is
    rebuilt from the new poses."""
    spread = 2.0
    slots, meshes = _scene(renderer, 20, 17, spread)
    try:
        fr = _frame(slots, spread, EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT)
        g, im, og, oim, _ = _render_both(renderer, meshes, fr)
        _assert_parity(g, im, og, oim, "20 objects")
        rng = np.random.default_rng(99)
        for k in range(0, 20, 3):
            cs = _pose(rng, spread)
            renderer.set_mesh_cs(slots[k], cs)
            meshes[k].cs = cs
        g2, im2, og2, oim2, _ = _render_both(renderer, meshes, fr)
        _assert_parity(g2, im2, og2, oim2, "20 objects, 7 moved")
        assert not np.array_equal(g2["depth"], g["depth"])
    finally:
        _remove(renderer, slots)


def test_many_objects_learned_rejected(renderer):
    spread = 2.0
    slots, _ = _scene(renderer, 17, 19, spread, small=True)
    try:
        fr = _frame(slots, spread, EDGES | ffi.LEARNED)
        tg = renderer.create_target(fr.width, fr.height)
        try:
            with pytest.raises(J.J3dgError):
                renderer.render(fr, tg)
        finally:
            renderer.destroy_target(tg)
    finally:
        _remove(renderer, slots)


@pytest.mark.parametrize("width, builder", [(4, ffi.BUILD_HOST), (8, ffi.BUILD_DEVICE), (4, ffi.BUILD_DEVICE)])
def test_many_objects_widths_builders_counting(renderer, width, builder):
    """The top level over 4-wide trees and device-built trees, rendered by the
    counting instance of the kernel (per-ray statistics on)."""
    spread = 2.0
    slots, meshes = _scene(renderer, 30, 23, spread, width=width, builder=builder)
    renderer.set_counting(True)
    try:
        fr = _frame(slots, spread, EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, az=1.2, el=0.2)
        g, im, og, oim, st = _render_both(renderer, meshes, fr)
        _assert_parity(g, im, og, oim, "30 objects W%d builder %d" % (width, builder))
        gst = renderer.stats()
        assert gst["stack_overflow"] == 0
        assert gst["primary_rays"] == fr.width * fr.height
    finally:
        renderer.set_counting(False)
        _remove(renderer, slots)
