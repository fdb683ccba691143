"""j3dg_render_batch: many frames (meshes of uneven size, cameras, resolutions,
modes) in one persistent launch captured in a HIP graph.  Each frame of a
batch must equal the same frame rendered alone (G-buffer and RGBA bit for
bit), on the first launch and on graph replays with other work in between,
and after the frame table changes under an unchanged graph."""
import math

import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import compare_gbuffers, compare_rgba

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (160, 96), (256, 256), (320, 180), (97, 61)]
FLAG_SETS = [ffi.EDGES | ffi.SHADING, ffi.EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, ffi.SHADING | ffi.SHADOW,
             ffi.WIREFRAME | ffi.SHADING, ffi.ONE_BIT | ffi.EDGES]


@pytest.fixture(scope="module")
def scene(renderer):
    meshes = []
    for k, (nu, nv) in enumerate([(40, 20), (120, 60), (300, 150), (700, 300)]):
        v, t = J.torus(nu, nv, seed=10 + k)
        slot = renderer.upload_mesh(v, t, db_id=J.make_db_id(2, k))
        meshes.append((slot, v, t, J.make_db_id(2, k)))
    v, t = J.icosphere(4, seed=21)
    meshes.append((renderer.upload_mesh(v, t, db_id=J.make_db_id(3, 0)), v, t, J.make_db_id(3, 0)))
    yield meshes
    for slot, _, _, _ in meshes:
        renderer.remove_mesh(slot)


def make_frames(meshes, n, seed):
    rng = np.random.default_rng(seed)
    cam = J.camera_with_vertical_fov(60.0)
    frames = []
    for i in range(n):
        slot, v, _, _ = meshes[i % len(meshes)]
        w, h = SIZES[i % len(SIZES)]
        cs, ci = J.orbit_camera(float(rng.uniform(2.5, 3.5)), float(rng.uniform(0, 2 * math.pi)),
                                float(rng.uniform(0.1, 1.0)))
        diag, pivot = J.scene_framing(v)
        light = rng.normal(size=3).astype(np.float32)
        light[2] = abs(light[2])
        light /= np.linalg.norm(light)
        rect = None if i % 3 else (3, 2, w - 5, h - 3)
        frames.append(J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=FLAG_SETS[i % len(FLAG_SETS)],
                                   light_dir=light, rect=rect))
    return frames


def render_single(r, frames, before=None):
    """Each frame rendered alone into a new target, which first gets the
    G-buffer `before[i]` when given: a miss leaves mark, colour and
    barycentrics as the target held them (j3d/canvas.cpp:859-866), and a
    sub-rectangle every pixel outside it."""
    out = []
    for i, fr in enumerate(frames):
        tg = r.create_target(fr.width, fr.height)
        if before is not None:
            r.write_gbuffer(tg, before[i])
        out.append(r.render(fr, tg))
        r.destroy_target(tg)
    return out


def check(r, frames, targets, expected, label):
    for i, (fr, tg, (g0, im0)) in enumerate(zip(frames, targets, expected)):
        g, im = r.read_target(tg)
        bad = {k: n for k, n in compare_gbuffers(g, g0).items() if n}
        assert not bad, (label, i, bad)
        assert compare_rgba(im, im0)[0] == 0, (label, i)


def test_batch_equals_single_frames(renderer, scene):
    r = renderer
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    frames = make_frames(scene, 13, seed=1)
    expected = render_single(r, frames)
    targets = [r.create_target(fr.width, fr.height) for fr in frames]
    try:
        r.render_batch(frames, targets)
        r.synchronize()
        check(r, frames, targets, expected, "first launch")
        # graph replays, with eager work on the same stream in between
        other = r.create_target(64, 48)
        for k in range(3):
            r.render(frames[0], other)
            r.render_batch(frames, targets)
            r.synchronize()
            check(r, frames, targets, expected, "replay %d" % k)
        r.destroy_target(other)
        # new cameras, same shapes: the graph is replayed over the updated tables;
        # the targets hold the first frames' G-buffers
        frames2 = make_frames(scene, 13, seed=2)
        expected2 = render_single(r, frames2, before=[g for g, _ in expected])
        r.render_batch(frames2, targets)
        r.synchronize()
        check(r, frames2, targets, expected2, "updated tables")
    finally:
        for tg in targets:
            r.destroy_target(tg)


def test_batch_against_oracle_and_counters(renderer, scene):
    r = renderer
    mc, cav = J.make_matcap(1)
    r.set_matcap(mc, cav)
    frames = make_frames(scene, 6, seed=3)
    targets = [r.create_target(fr.width, fr.height) for fr in frames]
    try:
        r.set_counting(True)
        r.render_batch(frames, targets)
        r.synchronize()
        st = r.stats()
        r.set_counting(False)
        total_px = 0
        for i, (fr, tg) in enumerate(zip(frames, targets)):
            slot, v, t, db = scene[i % len(scene)]
            g, im = r.read_target(tg)
            og = np.zeros((fr.height, fr.width), O.PIXEL_DTYPE)
            og["object_id"] = 0xFFFFFFFF
            og, ost = O.update_canvas([O.Mesh(v, t, db_id=db)], fr, gbuf=og)
            oim = O.canvas_to_image(og, fr, mc, cav)
            bad = compare_gbuffers(g, og)
            assert all(c == 0 for c in bad.values()), (i, bad)
            assert compare_rgba(im, oim)[0] <= 1
            x0, y0, x1, y1 = fr.rect if fr.rect else (0, 0, fr.width - 1, fr.height - 1)
            total_px += (x1 - x0 + 1) * (y1 - y0 + 1)
        assert st["primary_rays"] == total_px
    finally:
        for tg in targets:
            r.destroy_target(tg)


def test_batch_rejects_mixed_widths_and_no_scene(renderer, scene):
    r = renderer
    v, t = J.torus(30, 20, seed=5)
    s4 = r.upload_mesh(v, t, width=4)
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.orbit_camera(3.0, 0.3, 0.3)
    diag, pivot = J.scene_framing(v)
    fa = J.make_frame(64, 48, cam, cs, ci, pivot, diag, [scene[0][0]])
    fb = J.make_frame(64, 48, cam, cs, ci, pivot, diag, [s4])
    ta, tb = r.create_target(64, 48), r.create_target(64, 48)
    try:
        with pytest.raises(J.J3dgError):
            r.render_batch([fa, fb], [ta, tb])
        fn = J.make_frame(64, 48, cam, cs, ci, pivot, diag, [scene[0][0]], flags=ffi.EDGES | ffi.NO_SCENE)
        with pytest.raises(J.J3dgError):
            r.render_batch([fn], [ta])
    finally:
        r.destroy_target(ta)
        r.destroy_target(tb)
        r.remove_mesh(s4)


def with_flags(fr, extra, margin=0.0):
    return J.Frame(fr.width, fr.height, fr.proj_inv, fr.cs, fr.cs_inv, fr.pivot, fr.diagonal, fr.objects,
                   flags=fr.flags | extra, near_plane=fr.near_plane, light_dir=fr.light_dir, bg_top=fr.bg_top,
                   bg_bottom=fr.bg_bottom, rect=fr.rect, margin=margin)


def adversarial_proposals(g, rng):
    """Per pixel a bound of one kind: the plain hit distance, half of it (the
    bounded pass finds nothing: fallback), far beyond it, negative, inf, NaN,
    random, one ulp short; a fifth of the pixels predicted to miss."""
    h, w = g.shape
    depth = g["depth"]
    kinds = rng.integers(0, 8, size=(h, w))
    with np.errstate(over="ignore", invalid="ignore"):
        tb = np.where(kinds == 0, depth, 0)
        tb = np.where(kinds == 1, depth * 0.5, tb)
        tb = np.where(kinds == 2, depth * 3.0 + 1.0, tb)
        tb = np.where(kinds == 3, -1.0, tb)
        tb = np.where(kinds == 4, np.inf, tb)
        tb = np.where(kinds == 5, np.nan, tb)
        tb = np.where(kinds == 6, rng.uniform(0, 10, size=(h, w)), tb)
        tb = np.where(kinds == 7, np.nextafter(depth, np.float32(0)), tb)
    logit = np.where(rng.random((h, w)) < 0.8, 1.0, -1.0)
    return np.stack([tb, logit], axis=-1).astype(np.float32), logit


def test_batch_learned_frames_with_kept_proposals(renderer, scene):
    """Learned frames in a batch, each reading its own target's proposals
    (J3DG_KEEP_PROPOSAL, every kind of bound): every frame equals its plain
    frame rendered alone, and the bounded pass takes exactly the pixels of
    the rectangles predicted to hit (resolved + fallback)."""
    r = renderer
    rng = np.random.default_rng(17)
    frames = make_frames(scene, 10, seed=31)
    plain = render_single(r, frames)
    learned = [with_flags(fr, ffi.LEARNED | ffi.KEEP_PROPOSAL) for fr in frames]
    targets = [r.create_target(fr.width, fr.height) for fr in frames]
    try:
        n_pred = 0
        for fr, tg, (g, _) in zip(frames, targets, plain):
            pairs, logit = adversarial_proposals(g, rng)
            r.write_proposal(tg, pairs)
            x0, y0, x1, y1 = fr.rect
            n_pred += int((logit[y0:y1 + 1, x0:x1 + 1] > 0).sum())
        r.render_batch(learned, targets)
        r.synchronize()
        check(r, frames, targets, plain, "learned batch")
        st = r.stats()
        assert st["learned_resolved"] + st["learned_fallback"] == n_pred
        assert st["learned_fallback"] > 0 and st["learned_resolved"] > 0
    finally:
        for tg in targets:
            r.destroy_target(tg)


def test_batch_learned_frames_with_the_network(renderer):
    """A mesh with a trained proposal network, four learned frames in one
    batch: the network runs per frame before the graph launch (its time is
    reported) and fills each target's proposals exactly as the same frame
    rendered alone does; every frame equals its plain frame."""
    r = renderer
    v, t = J.torus(200, 100, seed=12)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(4, 0))
    targets = []
    try:
        r.train_proposal(slot, n_rays=1 << 21, epochs=2, batch=8192, seed=3)
        mesh = [(slot, v, t, J.make_db_id(4, 0))]
        frames = make_frames(mesh, 4, seed=41)
        plain = render_single(r, frames)
        learned = [with_flags(fr, ffi.LEARNED, margin=0.1) for fr in frames]
        alone = []
        for fr in learned:
            tg = r.create_target(fr.width, fr.height)
            r.render(fr, tg)
            alone.append(r.read_proposal(tg))
            r.destroy_target(tg)
        assert sum(int((p[..., 1] > 0).sum()) for p in alone) > 0  # the network predicts hits
        targets = [r.create_target(fr.width, fr.height) for fr in frames]
        r.reset_timings()
        r.render_batch(learned, targets)
        r.synchronize()
        check(r, frames, targets, plain, "learned batch (network)")
        for i, (tg, p0) in enumerate(zip(targets, alone)):
            assert np.array_equal(r.read_proposal(tg).view(np.uint32), p0.view(np.uint32)), i
        st = r.stats()
        assert st["learned_resolved"] + st["learned_fallback"] > 0
        assert float(r.timings()["mlp_ms"][-1]) > 0.0
    finally:
        for tg in targets:
            r.destroy_target(tg)
        r.remove_mesh(slot)


def test_batch_autotune_over_widths(renderer, scene):
    """j3dg_autotune_batch times one graph launch of the batch for BVH width
    {4, 8} x stack depth {8, 16, 32} x rays per work item {1, 2, 4} (the frames
    given once per width), applies the fastest tuning and reports its width;
    the batch at that setting still equals its frames rendered alone."""
    r = renderer
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    s4 = [(r.upload_mesh(v, t, db_id=db, width=4), v, t, db) for _, v, t, db in scene]
    frames8 = make_frames(scene, 7, seed=5)
    frames4 = make_frames(s4, 7, seed=5)
    targets = [r.create_target(fr.width, fr.height) for fr in frames8]
    base = r.get_tuning()
    try:
        rep = r.autotune_batch(frames4, frames8, targets, repeats=2)
        assert rep["width"] in (4, 8) and rep["n_tried"] == 20 and rep["tuning"]["coop_tail"] in (1, 3, 4)
        assert r.get_tuning() == rep["tuning"]
        chosen = frames8 if rep["width"] == 8 else frames4
        expected = render_single(r, chosen, before=[r.read_target(tg)[0] for tg in targets])
        r.render_batch(chosen, targets)
        r.synchronize()
        check(r, chosen, targets, expected, "autotuned batch")
    finally:
        r.set_tuning(**base)
        for tg in targets:
            r.destroy_target(tg)
        for slot, _, _, _ in s4:
            r.remove_mesh(slot)
