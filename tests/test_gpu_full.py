"""Full-size frames against the oracle.

* The 28M-triangle torus at 1920 x 1080 with directional shadow rays (the
  benchmark frame): every pixel of the G-buffer bit-exact, RGBA within 1 LSB,
  with the near-tie pixels (two nearest hits within 1e-6 relative depth)
  counted and printed.
* The 300M-triangle torus (15000 x 10000 grid) built on the GPU, one orbit
  camera at 3840 x 2160 with shadows: a 256 x 256 crop (sub-rectangle
  semantics of j3d/canvas.cpp:677-699) bit-exact against the oracle.

Both build the BVH on the device (the results do not depend on the BVH;
smaller meshes test both builders).
"""
import math
import os
import time

import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import MISS, compare_gbuffers, compare_rgba, pick_crop, torus_scene

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FLAGS = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT


def _threads():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def test_torus28m_1080p_full_frame(renderer):
    r = renderer
    t0 = time.time()
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(5000, 2800, seed=1)
    assert len(t) == 28_000_000
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0), builder=ffi.BUILD_DEVICE)
    try:
        fr = J.make_frame(1920, 1080, cam, cs, ci, pivot, diag, [slot], flags=FLAGS, light_dir=light)
        tg = r.create_target(1920, 1080)
        r.set_counting(True)
        g, im = r.render(fr, tg)
        st = r.stats()
        r.set_counting(False)
        r.destroy_target(tg)
    finally:
        r.remove_mesh(slot)
    print("GPU frame done %.1fs" % (time.time() - t0), flush=True)
    m = O.Mesh(v, t, db_id=J.make_db_id(1, 0), threads=_threads())
    og, oim, ost = O.render([m], fr, mc, cav, threads=_threads(), count_ties=True)
    del m
    mm = compare_gbuffers(g, og)
    dmax, ndiff = compare_rgba(im, oim)
    print("28M 1080p: %d hit pixels, %d shadowed, %d near-tie pixels, G-buffer mismatches %s, RGBA max diff %d "
          "(%.1fs)" % (ost["hits"], ost["shadowed"], ost["near_ties"], mm, dmax, time.time() - t0), flush=True)
    assert sum(mm.values()) == 0, mm
    assert dmax <= 1
    assert st["primary_rays"] == 1920 * 1080 and st["hits"] == ost["hits"] and st["shadowed"] == ost["shadowed"]
    assert st["stack_overflow"] == 0


def test_torus300m_4k_crop(renderer):
    r = renderer
    t0 = time.time()
    v, t = J.torus(15000, 10000, seed=1)
    assert len(t) == 300_000_000
    diag, pivot = J.scene_framing(v)
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.orbit_camera(3.1, 2.2, math.radians(35.0))
    light = J.random_hemisphere(12)
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0), builder=ffi.BUILD_DEVICE)
    try:
        w, h = 3840, 2160
        fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=FLAGS, light_dir=light)
        tg = r.create_target(w, h)
        g, im = r.render(fr, tg)
        r.destroy_target(tg)
    finally:
        r.remove_mesh(slot)
    print("GPU 300M frame done %.1fs" % (time.time() - t0), flush=True)
    # a 256 x 256 window, mostly surface with some silhouette, away from the
    # frame's border, with a one-pixel halo above and to the right for the
    # edge pass (canvas.cpp:596-628)
    x0, y0, x1, y1 = pick_crop(g[1:, :-1], 256)
    y0, y1 = y0 + 1, y1 + 1
    hit = g["object_id"][y0:y1 + 1, x0:x1 + 1] != MISS
    assert 0.2 < hit.mean() < 0.98, hit.mean()
    m = O.Mesh(v, t, db_id=J.make_db_id(1, 0), threads=_threads())
    print("oracle BVH %.1fs" % (time.time() - t0), flush=True)
    cf = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=FLAGS, light_dir=light,
                      rect=(x0, y0 - 1, x1 + 1, y1))
    og = g.copy()
    og[y0 - 1:y1 + 1, x0:x1 + 2] = np.zeros(1, O.PIXEL_DTYPE)[0]
    og["object_id"][y0 - 1:y1 + 1, x0:x1 + 2] = MISS
    og, ost = O.update_canvas([m], cf, gbuf=og, threads=_threads(), count_ties=True)
    del m
    mm = compare_gbuffers(g[y0 - 1:y1 + 1, x0:x1 + 2], og[y0 - 1:y1 + 1, x0:x1 + 2])
    oim = O.canvas_to_image(og, fr, mc, cav)
    dmax, _ = compare_rgba(im[y0:y1 + 1, x0:x1 + 1], oim[y0:y1 + 1, x0:x1 + 1])
    print("300M 4K crop (%d,%d)-(%d,%d): %d hits, %d shadowed, %d near ties, mismatches %s, RGBA max diff %d "
          "(%.1fs)" % (x0, y0, x1, y1, ost["hits"], ost["shadowed"], ost["near_ties"], mm, dmax, time.time() - t0),
          flush=True)
    assert sum(mm.values()) == 0, mm
    assert dmax <= 1
