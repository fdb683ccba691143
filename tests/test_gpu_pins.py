"""GPU parity on the hand-pinned cases of tests/test_oracle_pins.py and on the
scene branches of render_scene / update_canvas: the HIP path through the C ABI
against the oracle on the same inputs.

* the normal transform under a rotated camera and rotated object poses
  (j3d/canvas.cpp:790-792);
* the point light and its shadow rays (canvas.cpp:703-705, 836-857), every
  tiny-occluder case, full frame;
* the shading pass on the hand-made G-buffers (up neighbour, row 0, wireframe,
  one-bit), written into a target and shaded alone;
* an empty scene (canvas.cpp:744-760) over a sub-rectangle of a target that
  holds earlier contents, and render_scene(nullptr) (canvas.cpp:883-890).

Bar: G-buffer bit-exact, RGBA within 1 LSB (0 expected: no acos differences in
these frames' edge colours beyond that).
"""
import numpy as np
import pytest

import j3d_amd as J
import oracle as O
import pin_scenes as S
from j3d_amd import ffi
from scenes import MISS, compare_gbuffers, compare_rgba

pytestmark = pytest.mark.gpu

IDENT = np.eye(4, dtype=np.float32).reshape(16)


def _jframe(fr, slots, flags=None, rect=None):
    return J.Frame(fr.width, fr.height, fr.proj_inv, fr.cs, fr.cs_inv, fr.pivot, fr.diagonal, slots,
                   flags=fr.flags if flags is None else flags, near_plane=fr.near_plane, light_dir=fr.light_dir,
                   rect=fr.rect if rect is None else rect)


def _parity(g, im, og, oim, label):
    mm = compare_gbuffers(g, og)
    assert sum(mm.values()) == 0, "%s: G-buffer mismatches %s" % (label, mm)
    dmax, ndiff = compare_rgba(im, oim)
    assert dmax <= 1, "%s: RGBA differs by %d on %d pixels" % (label, dmax, ndiff)


def _render(r, jf, w, h, gbuf0=None, stages=ffi.STAGE_ALL):
    tg = r.create_target(w, h)
    try:
        if gbuf0 is not None:
            r.write_gbuffer(tg, gbuf0)
        return r.render(jf, tg, stages=stages)
    finally:
        r.destroy_target(tg)


def test_normal_transform_rotated_camera_and_poses(renderer):
    """The pinned plane scene at 96 x 72, plus a displaced torus under a second
    rotated pose, with point-light shadows: every pixel's normal, depth and
    mark equal the oracle's."""
    r = renderer
    v, t, ocs, fr = S.normal_scene(96, 72, flags=O.EDGES | O.SHADING | O.SHADOW)
    tv, tt = J.torus(120, 60, seed=3)
    tcs = S.rigid(S.rot("y", -0.6) @ S.rot("x", 1.1), [0.4, 0.3, 0.8])
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    a = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0), cs=ocs)
    b = r.upload_mesh(tv, tt, db_id=J.make_db_id(1, 1), cs=tcs)
    try:
        g, im = _render(r, _jframe(fr, [a, b]), fr.width, fr.height)
    finally:
        r.remove_mesh(a)
        r.remove_mesh(b)
    meshes = [O.Mesh(v, t, cs=ocs, db_id=J.make_db_id(1, 0)), O.Mesh(tv, tt, cs=tcs, db_id=J.make_db_id(1, 1))]
    og, oim, st = O.render(meshes, fr, mc, cav)
    _parity(g, im, og, oim, "rotated camera and poses")
    assert st["shadowed"] > 0 and (g["db_id"] == J.make_db_id(1, 1)).sum() > 100


@pytest.mark.parametrize("name", sorted(S.SHADOW_CASES))
def test_point_light_occluders(renderer, name):
    r = renderer
    where, rad, expect = S.SHADOW_CASES[name]
    v, t, fr, _ = S.shadow_scene(where, rad, full=True)
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    try:
        g, im = _render(r, _jframe(fr, [slot]), fr.width, fr.height)
    finally:
        r.remove_mesh(slot)
    og, oim, _ = O.render([O.Mesh(v, t, db_id=J.make_db_id(1, 0))], fr, mc, cav)
    _parity(g, im, og, oim, "point light, occluder %s" % name)
    x, y = S.SHADOW_PIXEL
    assert g["mark"][y, x] & 1 == expect


def _shade_cases():
    mc0, cav0 = O.matcap(0)
    return {
        "up_step+": (S.row_step_gbuf(0.4), O.EDGES | O.SHADING, mc0, cav0),
        "up_step-": (S.row_step_gbuf(-0.4), O.EDGES | O.SHADING, mc0, cav0),
        "row_zero": (S.row_zero_gbuf(), O.EDGES | O.SHADING, mc0, cav0),
        "wireframe": (S.wire_gbuf(), O.WIREFRAME | O.SHADING, S.const_matcap(0xff5a6b7c), 0xff102030),
        "one_bit_808080": (S.one_bit_edge_gbuf(), O.ONE_BIT, S.const_matcap(0xff808080), 0xff000000),
        "one_bit_000000": (S.one_bit_edge_gbuf(), O.ONE_BIT, S.const_matcap(0xff000000), 0xff000000),
    }


@pytest.mark.parametrize("name", sorted(_shade_cases()))
def test_shading_pass_on_pinned_gbuffers(renderer, name):
    g0, flags, mc, cav = _shade_cases()[name]
    h, w = g0.shape
    r = renderer
    r.set_matcap(mc, cav)
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, flags)
    g, im = _render(r, _jframe(fr, []), w, h, gbuf0=g0, stages=ffi.STAGE_SHADE)
    assert sum(compare_gbuffers(g, g0).values()) == 0  # shading leaves the G-buffer alone
    oim = O.canvas_to_image(g0, fr, mc, cav)
    dmax, ndiff = compare_rgba(im, oim)
    assert dmax <= 1, (name, dmax, ndiff)
    if name.startswith("one_bit") or name == "wireframe":
        assert np.array_equal(im, oim)


def _garbage_gbuf(w, h, seed):
    """Earlier contents of a target: hits with marks, colours and barycentrics."""
    rng = np.random.default_rng(seed)
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    for k in ("mark", "r", "g", "b"):
        g[k] = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    for k in ("u", "v", "depth", "barycentric_u", "barycentric_v"):
        g[k] = rng.uniform(-1, 1, size=(h, w)).astype(np.float32)
    g["object_id"] = rng.integers(0, 1000, size=(h, w), dtype=np.uint32)
    g["db_id"] = rng.integers(1, 1 << 30, size=(h, w), dtype=np.uint32)
    return g


@pytest.mark.parametrize("flags", [ffi.EDGES | ffi.SHADING, ffi.EDGES | ffi.SHADING | ffi.SHADOW])
def test_empty_scene_over_a_rectangle(renderer, flags):
    """canvas.cpp:744-760: no objects -> every pixel of the rectangle gets
    db_id 0, object_id -1, u = v = 0, depth FLT_MAX; its mark, colour and
    barycentrics and every pixel outside the rectangle keep what the target
    held."""
    r = renderer
    w, h = 40, 30
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    g0 = _garbage_gbuf(w, h, 5)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    jf = J.make_frame(w, h, J.camera_with_vertical_fov(60.0), cs, ci, (0, 0, 0), 2.0, [], flags=flags,
                      rect=(3, 4, 30, 21))
    g, im = _render(r, jf, w, h, gbuf0=g0)
    og, st = O.update_canvas([], jf, gbuf=g0.copy())
    assert st["primary_rays"] == 0
    inside = np.zeros((h, w), bool)
    inside[4:22, 3:31] = True
    assert np.all(og["object_id"][inside] == MISS) and np.all(og["depth"][inside] == np.finfo(np.float32).max)
    assert np.array_equal(og[~inside], g0[~inside]) and np.array_equal(og["mark"], g0["mark"])
    oim = O.canvas_to_image(og, jf, mc, cav)
    _parity(g, im, og, oim, "empty scene")


def test_render_scene_without_a_scene(renderer):
    """render_scene(nullptr) (canvas.cpp:883-890): every pixel of the target
    (the rectangle does not apply) gets db_id 0 and object_id -1 and keeps
    every other field; the image is the background."""
    r = renderer
    w, h = 37, 23
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    g0 = _garbage_gbuf(w, h, 9)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    jf = J.make_frame(w, h, J.camera_with_vertical_fov(60.0), cs, ci, (0, 0, 0), 2.0, [],
                      flags=ffi.EDGES | ffi.SHADING | ffi.NO_SCENE, rect=(5, 5, 10, 10))
    g, im = _render(r, jf, w, h, gbuf0=g0)
    exp = g0.copy()
    exp["db_id"] = 0
    exp["object_id"] = MISS
    assert sum(compare_gbuffers(g, exp).values()) == 0
    assert np.array_equal(im, O.background(w, h))


def test_read_target_into_strided_rows(renderer):
    """j3dg_read_target_strided: rows of the target's G-buffer and image land
    `stride` elements apart (jtk::image rows, canvas.cpp:964-966), equal to
    the packed readback, the padding untouched; a stride below the width is
    refused."""
    r = renderer
    v, t = J.icosphere(3, seed=7, amplitude=0.05)
    diag, pivot = J.scene_framing(v)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    slot = r.upload_mesh(v, t)
    w, h = 37, 29
    tg = r.create_target(w, h)
    try:
        fr = J.make_frame(w, h, J.camera_with_vertical_fov(60.0), cs, ci, pivot, diag, [slot])
        g, im = r.render(fr, tg)
        gs, ims = r.read_target_strided(tg, 40, 48)
        assert sum(compare_gbuffers(gs[:, :w], g).values()) == 0 and np.array_equal(ims[:, :w], im)
        assert np.all(gs["object_id"][:, w:] == 0) and np.all(ims[:, w:] == 0)
        with pytest.raises(J.J3dgError):
            r.read_target_strided(tg, w - 1, w)
    finally:
        r.destroy_target(tg)
        r.remove_mesh(slot)
