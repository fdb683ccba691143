"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: every G-buffer field bit-exact (hit ids, db ids, depth, normals,
barycentrics, marks); RGBA within 1 LSB per channel (acos in the edge pass
may differ by an ulp between the device and host math libraries).
"""
import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import MISS, c0_scene, compare_gbuffers, compare_rgba, random_rays, torus_scene

pytestmark = pytest.mark.gpu

EDGES = ffi.EDGES | ffi.SHADING


def _render_both(r, meshes_gpu, meshes_orc, frame, matcap=0):
    mc, cav = J.make_matcap(matcap)
    r.set_matcap(mc, cav)
    tg = r.create_target(frame.width, frame.height)
    try:
        g, im = r.render(frame, tg)
    finally:
        r.destroy_target(tg)
    og, oim, st = O.render(meshes_orc, frame, mc, cav)
    return g, im, og, oim, st


def _assert_parity(g, im, og, oim, label):
    mm = compare_gbuffers(g, og)
    assert sum(mm.values()) == 0, "%s: G-buffer mismatches %s" % (label, mm)
    dmax, ndiff = compare_rgba(im, oim)
    print("%s: %d hit pixels, RGBA max diff %d on %d pixels" % (label, int((g["object_id"] != MISS).sum()), dmax,
                                                                ndiff))
    assert dmax <= 1, "%s: RGBA differs by %d" % (label, dmax)


def test_c0_full_size(renderer):
    v, t, cam, cs, ci, pivot, diag = c0_scene()
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    fr = J.make_frame(1024, 768, cam, cs, ci, pivot, diag, [slot], flags=EDGES)
    g, im, og, oim, st = _render_both(renderer, None, [O.Mesh(v, t, db_id=J.make_db_id(1, 0))], fr)
    _assert_parity(g, im, og, oim, "C0 1024x768")
    assert st["hits"] > 100000
    renderer.remove_mesh(slot)


@pytest.mark.parametrize("flags", [EDGES | ffi.SHADOW, EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, ffi.SHADING,
                                   ffi.WIREFRAME | ffi.SHADING, ffi.ONE_BIT])
def test_modes_icosphere(renderer, flags):
    v, t, cam, cs, ci, pivot, diag = c0_scene(level=4)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 3))
    fr = J.make_frame(320, 240, cam, cs, ci, pivot, diag, [slot], flags=flags,
                      light_dir=J.random_hemisphere(5))
    g, im, og, oim, _ = _render_both(renderer, None, [O.Mesh(v, t, db_id=J.make_db_id(1, 3))], fr)
    _assert_parity(g, im, og, oim, "icosphere L4 flags=%#x" % flags)
    renderer.remove_mesh(slot)


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("light_kind", ["directional", "point"])
def test_torus_shadow(renderer, split, light_kind):
    """Shadow rays in one launch (split 0: a lane traces its pixel's shadow ray
    after the primary hit) and in three (split 1: primary, k_finish, shadow)."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(400, 160)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    flags = EDGES | ffi.SHADOW | (ffi.DIRECTIONAL_LIGHT if light_kind == "directional" else 0)
    fr = J.make_frame(640, 360, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light)
    tun = renderer.get_tuning()
    renderer.set_tuning(**dict(tun, split_passes=split))
    try:
        g, im, og, oim, st = _render_both(renderer, None, [O.Mesh(v, t, db_id=J.make_db_id(1, 0))], fr)
        gst = renderer.stats()
    finally:
        renderer.set_tuning(**tun)
    _assert_parity(g, im, og, oim, "torus 128k tris 640x360 shadow split=%d %s" % (split, light_kind))
    assert st["shadowed"] > 0
    assert gst["shadow_rays"] == st["shadow_rays"] and gst["shadowed"] == st["shadowed"], (gst, st)
    renderer.remove_mesh(slot)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("leaf_size", [1, 2, 3, 4])
def test_leaf_sizes_and_widths(renderer, width, leaf_size):
    """Every BVH shape the builder offers gives the oracle's frame and hits.  An
    8-wide node and a leaf's first two triangles share one 96-byte fetch, a
    leaf's third and fourth triangle are fetched apart, and a one-triangle
    leaf at the end of the triangle array reads the spare record."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(300, 130, seed=5)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(2, 1), width=width, leaf_size=leaf_size)
    m = O.Mesh(v, t, db_id=J.make_db_id(2, 1))
    fr = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [slot], flags=EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT,
                      light_dir=light)
    g, im, og, oim, st = _render_both(renderer, None, [m], fr)
    _assert_parity(g, im, og, oim, "torus 78k tris W%d leaf %d" % (width, leaf_size))
    assert st["shadowed"] > 0
    rays = random_rays(50000, 17, np.zeros(3), 1.4)
    got = renderer.trace_rays([slot], rays)
    ref = O.closest_hits([m], rays)
    for k in ("obj", "tri"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert np.array_equal(renderer.any_hit([slot], rays), O.any_hits([m], rays))
    renderer.remove_mesh(slot)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("leaf_size", [1, 2, 4])
def test_device_builder(renderer, width, leaf_size):
    """The BVH, normals and completion records built on the GPU from the raw
    arrays (BUILD_DEVICE) give the oracle's frame and hits, and the same
    normals, bounds and stack-safe traversal as the host build."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(300, 130, seed=5)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(2, 1), width=width, leaf_size=leaf_size,
                                builder=ffi.BUILD_DEVICE)
    host = renderer.upload_mesh(v, t, width=width, leaf_size=leaf_size)
    try:
        n_dev = renderer.mesh_normals(slot, len(t))
        assert np.array_equal(n_dev.view(np.uint32), renderer.mesh_normals(host, len(t)).view(np.uint32))
        info, hinfo = renderer.mesh_info(slot), renderer.mesh_info(host)
        for k in ("n_vertices", "n_triangles", "width", "bb_min", "bb_max"):
            assert info[k] == hinfo[k], k
        assert 0 < info["n_nodes"] < len(t) and info["max_depth"] >= 2
        print("device build W%d leaf %d: %d nodes, depth %d (host %d nodes, depth %d)" % (
            width, leaf_size, info["n_nodes"], info["max_depth"], hinfo["n_nodes"], hinfo["max_depth"]))
    finally:
        renderer.remove_mesh(host)
    m = O.Mesh(v, t, db_id=J.make_db_id(2, 1))
    fr = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [slot], flags=EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT,
                      light_dir=light)
    g, im, og, oim, st = _render_both(renderer, None, [m], fr)
    _assert_parity(g, im, og, oim, "torus 78k tris device build W%d leaf %d" % (width, leaf_size))
    assert st["shadowed"] > 0
    rays = random_rays(50000, 17, np.zeros(3), 1.4)
    got = renderer.trace_rays([slot], rays)
    ref = O.closest_hits([m], rays)
    for k in ("obj", "tri"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert np.array_equal(renderer.any_hit([slot], rays), O.any_hits([m], rays))
    renderer.remove_mesh(slot)


def _degenerate_meshes():
    """Inputs that stress the device build: every Morton code equal (identical
    triangles: ties go to the lowest triangle id), a flat strip with one empty
    axis and a 1e4:1 aspect, a soup whose NaN vertices give NaN boxes, and
    meshes of one, two and three triangles (one triangle is built on the
    host)."""
    rng = np.random.default_rng(4)
    same_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    same_t = np.tile(np.array([[0, 1, 2]], np.uint32), (3000, 1))
    g = 60
    ys, xs = np.meshgrid(np.arange(g + 1, dtype=np.float32), np.arange(g + 1, dtype=np.float32), indexing="ij")
    strip_v = np.stack([xs.ravel() * 1e-2, ys.ravel() * 1e-6, np.zeros(xs.size, np.float32)], 1).astype(np.float32)
    a = (np.arange(g)[None, :] + (g + 1) * np.arange(g)[:, None]).ravel().astype(np.uint32)
    strip_t = np.concatenate([np.stack([a, a + 1, a + g + 1], 1), np.stack([a + 1, a + g + 2, a + g + 1], 1)])
    small_v = rng.uniform(-1, 1, size=(9, 3)).astype(np.float32)
    nan_v = rng.uniform(-1, 1, size=(3000, 3)).astype(np.float32)
    nan_v[::7, rng.integers(0, 3)] = np.nan  # every 7th vertex has a NaN coordinate: its triangles never hit
    nan_t = np.arange(3000, dtype=np.uint32).reshape(-1, 3)
    return [("identical x3000", same_v, same_t), ("flat strip", strip_v, strip_t.astype(np.uint32)),
            ("soup with NaN vertices", nan_v, nan_t),
            ("1 tri", small_v, np.array([[0, 1, 2]], np.uint32)),
            ("2 tris", small_v, np.array([[0, 1, 2], [3, 4, 5]], np.uint32)),
            ("3 tris", small_v, np.arange(9, dtype=np.uint32).reshape(3, 3))]


@pytest.mark.parametrize("case", range(6))
def test_device_builder_degenerate(renderer, case):
    label, v, t = _degenerate_meshes()[case]
    lo, hi = np.nanmin(v, 0), np.nanmax(v, 0)
    center = (lo + hi) * 0.5
    radius = float(max(np.linalg.norm(hi - lo), 1e-3))
    # half the rays aimed at random points of random triangles, half at random points near the mesh
    rng = np.random.default_rng(31 + case)
    k = rng.integers(0, len(t), 2000)
    b = rng.dirichlet(np.ones(3), 2000).astype(np.float32)
    target = (v[t[k, 0]] * b[:, :1] + v[t[k, 1]] * b[:, 1:2] + v[t[k, 2]] * b[:, 2:3]).astype(np.float32)
    o = rng.normal(size=(2000, 3))
    o = target + 2.0 * radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    aimed = np.zeros((2000, 8), np.float32)
    aimed[:, 0:3] = o
    aimed[:, 3:6] = target - o
    aimed[:, 6] = 1e-3
    aimed[:, 7] = np.finfo(np.float32).max
    rays = np.concatenate([aimed, random_rays(2000, 23 + case, center, radius * 0.6)])
    m = O.Mesh(v, t)
    ref = O.closest_hits([m], rays, brute=True)  # brute force: the oracle's own BVH needs no NaN handling
    for width in (4, 8):
        slot = renderer.upload_mesh(v, t, width=width, builder=ffi.BUILD_DEVICE)
        got = renderer.trace_rays([slot], rays)
        for k in ("obj", "tri"):
            assert np.array_equal(got[k], ref[k]), (label, width, k)
        for k in ("t", "u", "v"):
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (label, width, k)
        assert np.array_equal(renderer.any_hit([slot], rays), O.any_hits([m], rays, brute=True)), (label, width)
        renderer.remove_mesh(slot)
    print("%s: %d of %d rays hit" % (label, int((ref["obj"] != MISS).sum()), len(rays)))


@pytest.mark.parametrize("builder", [ffi.BUILD_HOST, ffi.BUILD_DEVICE])
def test_rays_that_are_not_finite_hit_nothing(renderer, builder):
    """Rays with a NaN or infinite origin or direction component hit nothing
    and are not traversed (their slab tests would be NaN, which min / max
    drop): closest, any and all hits, with the finite rays of the same call
    unchanged; also through the top level of a 17-object scene."""
    v, t = J.torus(120, 60, seed=8)
    rays = random_rays(3000, 29, np.zeros(3), 1.4)
    bad = np.arange(0, 3000, 5)
    comp = np.array([0, 1, 2, 3, 4, 5])[bad % 6]
    vals = np.array([np.nan, np.inf, -np.inf], np.float32)[(bad // 6) % 3]
    rays[bad, comp] = vals
    slot = renderer.upload_mesh(v, t, builder=builder)
    m = O.Mesh(v, t)
    try:
        got = renderer.trace_rays([slot], rays)
        ref = O.closest_hits([m], rays, brute=True)
        for k in ("obj", "tri"):
            assert np.array_equal(got[k], ref[k]), k
        assert (got["obj"][bad] == MISS).all()
        assert np.array_equal(renderer.any_hit([slot], rays), O.any_hits([m], rays, brute=True))
        ah = renderer.all_hits(slot, rays, 8)
        assert (ah["count"][bad] == 0).all()
        assert np.array_equal(ah["count"], O.all_hits(m, rays, 8, brute=True)["count"])
        slots = [slot] + [renderer.upload_mesh(v, t, builder=builder) for _ in range(16)]
        try:
            got17 = renderer.trace_rays(slots, rays)
            assert np.array_equal(got17["tri"], got["tri"]) and (got17["obj"][bad] == MISS).all()
        finally:
            for s2 in slots[1:]:
                renderer.remove_mesh(s2)
    finally:
        renderer.remove_mesh(slot)


def test_frame_with_camera_that_is_not_finite(renderer):
    """A NaN in the camera matrix makes every primary ray non-finite: the
    frame traces no node beyond the root and comes out all background, as the
    oracle's (whose triangle test rejects such rays)."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(100, 50, seed=3)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(2, 2))
    try:
        cs = np.array(cs, np.float32)
        cs[12] = np.nan
        fr = J.make_frame(160, 120, cam, cs, ci, pivot, diag, [slot], flags=EDGES | ffi.SHADOW, light_dir=light)
        g, im, og, oim, st = _render_both(renderer, None, [O.Mesh(v, t, db_id=J.make_db_id(2, 2))], fr)
        _assert_parity(g, im, og, oim, "NaN camera")
        assert (g["object_id"] == MISS).all()
    finally:
        renderer.remove_mesh(slot)


def test_pose_must_be_finite(renderer):
    v, t = J.torus(20, 10, seed=8)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[13] = np.nan
    with pytest.raises(J.J3dgError):
        renderer.upload_mesh(v, t, cs=cs)
    slot = renderer.upload_mesh(v, t)
    try:
        with pytest.raises(J.J3dgError):
            renderer.set_mesh_cs(slot, cs)
    finally:
        renderer.remove_mesh(slot)


def test_mixed_builders_two_objects(renderer):
    """A host-built and a device-built object in one two-level scene."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(200, 90, seed=2)
    v2, t2 = J.icosphere(4, seed=3, amplitude=0.05)
    cs2 = np.eye(4, dtype=np.float32)
    cs2[3, :3] = (0.4, 0.3, 0.2)  # column-major: translation in the last column
    cs2 = cs2.reshape(16)
    s1 = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    s2 = renderer.upload_mesh(v2, t2, db_id=J.make_db_id(1, 1), cs=cs2, builder=ffi.BUILD_DEVICE)
    ms = [O.Mesh(v, t, db_id=J.make_db_id(1, 0)), O.Mesh(v2, t2, db_id=J.make_db_id(1, 1), cs=cs2)]
    fr = J.make_frame(400, 300, cam, cs, ci, pivot, diag, [s1, s2], flags=EDGES | ffi.SHADOW, light_dir=light)
    g, im, og, oim, st = _render_both(renderer, None, ms, fr)
    _assert_parity(g, im, og, oim, "torus host + icosphere device")
    assert len(np.unique(g["db_id"][g["object_id"] != MISS])) == 2
    renderer.remove_mesh(s1)
    renderer.remove_mesh(s2)


def test_random_rays_closest_and_any(renderer):
    v, t = J.torus(300, 120, seed=3)
    slot = renderer.upload_mesh(v, t)
    rays = random_rays(200000, 9, np.zeros(3), 1.4)
    got = renderer.trace_rays([slot], rays)
    m = O.Mesh(v, t)
    ref = O.closest_hits([m], rays)
    for k in ("obj", "tri"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    hit = renderer.any_hit([slot], rays)
    assert np.array_equal(hit, O.any_hits([m], rays))
    renderer.remove_mesh(slot)


def test_banded_tiles_large_frame(renderer):
    """A frame's tiles are split into eight bands with a queue each (one per
    XCD), each band sorted heaviest first on repeat frames: a 1280x720 frame
    (14,400 tiles, 57 sort blocks, bands of 7-8 blocks) rendered twice, the
    second time in the banded cost order, and a 1000x600 frame whose last
    block is partial; every pixel must come out once and right."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(500, 200, seed=9)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    meshes = [O.Mesh(v, t, db_id=J.make_db_id(1, 0))]
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    flags = EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    try:
        for w, h in ((1280, 720), (1000, 600)):
            fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light)
            tg = renderer.create_target(w, h)
            try:
                renderer.render(fr, tg, readback=False)
                # stale contents a skipped tile would leave behind
                junk = np.zeros((h, w), ffi.PIXEL_DTYPE)
                junk["object_id"] = 7
                junk["depth"] = 1.0
                renderer.write_gbuffer(tg, junk)
                g, im = renderer.render(fr, tg)
            finally:
                renderer.destroy_target(tg)
            og, oim, _ = O.render(meshes, fr, mc, cav)
            _assert_parity(g, im, og, oim, "banded %dx%d" % (w, h))
    finally:
        renderer.remove_mesh(slot)


def test_cost_ordered_tiles_repeat_frames(renderer):
    """Frames after the first hand out tiles heaviest first (by the previous
    frame's per-pixel work); results must not change: same frame three times,
    a different rectangle, then the first rectangle again."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(400, 160, seed=7)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    meshes = [O.Mesh(v, t, db_id=J.make_db_id(1, 0))]
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    # 1024x576 = 9216 tiles and a rectangle of 8487: both at least the 8192 tiles given a cost order
    tg = renderer.create_target(1024, 576)
    flags = EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    try:
        for k, rect in enumerate([None, None, None, (20, 10, 1000, 560), (20, 10, 1000, 560), None]):
            fr = J.make_frame(1024, 576, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light, rect=rect)
            g, im = renderer.render(fr, tg)
            og = np.zeros((576, 1024), O.PIXEL_DTYPE)
            og["object_id"] = 0xFFFFFFFF
            # the oracle G-buffer carries over like the target's: render into the previous result
            if k:
                og = prev_og
            og, _ = O.update_canvas(meshes, fr, gbuf=og)
            oim = O.canvas_to_image(og, fr, mc, cav)
            _assert_parity(g, im, og, oim, "repeat %d rect %s" % (k, rect))
            prev_og = og
    finally:
        renderer.destroy_target(tg)
        renderer.remove_mesh(slot)


def test_autotune_keeps_results(renderer, tmp_path):
    """The autotuner tries 9 (depth, rays per item) points and 3 more refill
    thresholds, applies a valid configuration, and the frame it tunes still
    matches the oracle bit for bit; a cached result is reused unchanged."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(300, 120, seed=11)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    flags = EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    fr = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light)
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    tg = renderer.create_target(480, 270)
    base = renderer.get_tuning()
    cache = str(tmp_path / "tune.json")
    try:
        rep = renderer.autotune(fr, tg, repeats=2, cache=cache, key="torus@480x270")
        assert rep["n_tried"] == 14 and not rep["cached"] and rep["width"] == 8
        tun = renderer.get_tuning()
        assert tun == rep["tuning"]
        assert tun["stack_depth"] in (8, 16, 32) and tun["rays_per_item"] in (1, 2, 4)
        assert tun["refill_min"] in (8, 16, 32, 48) and tun["coop_tail"] in (1, 3, 4)
        assert rep["best_ms"] <= rep["baseline_ms"]
        g, im = renderer.render(fr, tg)
        renderer.set_tuning(**base)
        again = renderer.autotune(fr, tg, cache=cache, key="torus@480x270")
        assert again["cached"] and renderer.get_tuning() == rep["tuning"]
    finally:
        renderer.set_tuning(**base)
        renderer.destroy_target(tg)
        renderer.remove_mesh(slot)
    og, oim, _ = O.render([O.Mesh(v, t, db_id=J.make_db_id(1, 0))], fr, mc, cav)
    _assert_parity(g, im, og, oim, "autotuned torus 480x270")


def test_autotune_chooses_width(renderer):
    """j3dg_autotune_widths tunes the same frame over width-4 and width-8
    uploads of its mesh (14 settings each), applies the faster tuning and
    reports its width; the frame at the chosen width matches the oracle."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(300, 120, seed=11)
    s4 = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0), width=4)
    s8 = renderer.upload_mesh(v, t, db_id=J.make_db_id(1, 0), width=8)
    flags = EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    f4 = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [s4], flags=flags, light_dir=light)
    f8 = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [s8], flags=flags, light_dir=light)
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    tg = renderer.create_target(480, 270)
    base = renderer.get_tuning()
    try:
        rep = renderer.autotune_widths(f4, f8, tg, repeats=2)
        assert rep["width"] in (4, 8) and rep["n_tried"] == 28
        assert renderer.get_tuning() == rep["tuning"]
        g, im = renderer.render(f8 if rep["width"] == 8 else f4, tg)
        only8 = renderer.autotune_widths(None, f8, tg, repeats=2)
        assert only8["width"] == 8 and only8["n_tried"] == 14
    finally:
        renderer.set_tuning(**base)
        renderer.destroy_target(tg)
        renderer.remove_mesh(s4)
        renderer.remove_mesh(s8)
    og, oim, _ = O.render([O.Mesh(v, t, db_id=J.make_db_id(1, 0))], f8, mc, cav)
    _assert_parity(g, im, og, oim, "width-autotuned torus 480x270")


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("source", ["vertex", "texture", "both"])
def test_surface_colors(renderer, source, split):
    """Textured and vertex-coloured hits (j3d/canvas.cpp:803-834) with shadows:
    random vertex colours in [-0.1, 1.1] (byte wrap-around included), random
    per-corner uv in [-0.2, 1.2] (clamping) on a 37x23 texture; colour and
    shadow marks combine (mark 3).  One launch and the split pipeline."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(200, 90, seed=13)
    rng = np.random.default_rng(17)
    vc = rng.uniform(-0.1, 1.1, (len(v), 3)).astype(np.float32)
    uv = rng.uniform(-0.2, 1.2, (len(t), 6)).astype(np.float32)
    tex = rng.integers(0, 2 ** 32, (23, 37), dtype=np.uint64).astype(np.uint32)
    slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(3, 2))
    kw = {"vertex": dict(vertex_colors=vc), "texture": dict(uv=uv, texture=tex),
          "both": dict(vertex_colors=vc, uv=uv, texture=tex)}[source]
    renderer.set_mesh_colors(slot, **kw)
    flags = EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT | ffi.TEXTURED | ffi.VERTEXCOLORS
    fr = J.make_frame(480, 270, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light)
    tun = renderer.get_tuning()
    renderer.set_tuning(**dict(tun, split_passes=split))
    try:
        g, im, og, oim, st = _render_both(renderer, None, [O.Mesh(v, t, db_id=J.make_db_id(3, 2), **kw)], fr)
    finally:
        renderer.set_tuning(**tun)
    _assert_parity(g, im, og, oim, "coloured torus (%s) split=%d" % (source, split))
    hit = g["object_id"] != MISS
    assert set(np.unique(g["mark"][hit])) == {2, 3}
    # removing the colours gives the uncoloured frame (mark 0 / 1)
    renderer.set_mesh_colors(slot)
    tg = renderer.create_target(480, 270)
    try:
        g2, _ = renderer.render(fr, tg)
    finally:
        renderer.destroy_target(tg)
    assert set(np.unique(g2["mark"][g2["object_id"] != MISS])) == {0, 1}
    renderer.remove_mesh(slot)


def test_surface_color_arguments(renderer):
    v, t = J.torus(20, 10, seed=1)
    slot = renderer.upload_mesh(v, t)
    with pytest.raises(J.J3dgError):
        renderer.set_mesh_colors(slot, uv=np.zeros((len(t), 6), np.float32))  # uv without a texture
    with pytest.raises(J.J3dgError):
        renderer.set_mesh_colors(slot, texture=np.zeros((2, 2), np.uint32))  # texture without uv
    with pytest.raises(J.J3dgError):
        renderer.set_mesh_colors(99, vertex_colors=np.zeros((len(v), 3), np.float32))
    renderer.remove_mesh(slot)


def test_tail_modes_give_the_oracles_results(renderer):
    """Every tail mode -- coop_tail 0 (none), 1 (the cooperative tail), 2-4
    (the split tail from 8, 16 or 32 rays left, lanes of a group traversing
    parts of a ray's tree on stacks of their own and regrouping as rays
    finish) -- at stack depth 8 (spilling) and 16 gives the oracle's frames
    and hits bit for bit: a 960x540 frame of a 312k-triangle torus with
    directional shadows, a frame of three posed objects with point-light
    shadows (a ray moves on from object to object), closest- and any-hit
    queries over the three."""
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(600, 260, seed=13)
    parts = [(v, t, None), J.torus(200, 90, R=0.5, r=0.18, seed=14) + (J.translation_camera(0.9, 0.4, 0.2)[0],),
             J.icosphere(4, seed=15) + (J.translation_camera(-0.8, -0.3, 0.5)[0],)]
    slots = [renderer.upload_mesh(a, b, db_id=J.make_db_id(5, k), cs=p) for k, (a, b, p) in enumerate(parts)]
    meshes = [O.Mesh(a, b, cs=p, db_id=J.make_db_id(5, k)) for k, (a, b, p) in enumerate(parts)]
    frames = [J.make_frame(960, 540, cam, cs, ci, pivot, diag, [slots[0]], flags=EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT,
                           light_dir=light),
              J.make_frame(640, 360, cam, cs, ci, pivot, diag * 1.6, slots, flags=ffi.SHADING | ffi.SHADOW)]
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    ref = [O.render([meshes[0]], frames[0], mc, cav)[:2], O.render(meshes, frames[1], mc, cav)[:2]]
    rays = random_rays(100000, 21, np.zeros(3), 1.6)
    ref_hits = O.closest_hits(meshes, rays)
    ref_any = O.any_hits(meshes, rays)
    base = renderer.get_tuning()
    try:
        for depth in (8, 16):
            for mode in range(5):
                renderer.set_tuning(**dict(base, coop_tail=mode, stack_depth=depth))
                for fr, (og, oim) in zip(frames, ref):
                    tg = renderer.create_target(fr.width, fr.height)
                    try:
                        g, im = renderer.render(fr, tg)
                    finally:
                        renderer.destroy_target(tg)
                    _assert_parity(g, im, og, oim, "tail mode %d, depth %d, %dx%d" % (mode, depth, fr.width, fr.height))
                got = renderer.trace_rays(slots, rays)
                for k in ("obj", "tri"):
                    assert np.array_equal(got[k], ref_hits[k]), (mode, depth, k)
                for k in ("t", "u", "v"):
                    assert np.array_equal(got[k].view(np.uint32), ref_hits[k].view(np.uint32)), (mode, depth, k)
                assert np.array_equal(renderer.any_hit(slots, rays), ref_any), (mode, depth)
    finally:
        renderer.set_tuning(**base)
        for s in slots:
            renderer.remove_mesh(s)
