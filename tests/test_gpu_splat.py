"""GPU parity of the point-cloud splat (j3dg_splat_points,
canvas::render_pointclouds_on_image, j3d/canvas.cpp:952-1030) against the
oracle, which draws point after point in order: G-buffer and image bit-exact
after the clouds are drawn over a rendered frame.

The clouds put many points in each pixel and hold exact duplicates (within a
cloud and across clouds: the earlier one stays), NaN points, points behind
the camera and off screen, so the depth merge, its tie rule and the rejects
are all exercised."""
import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi

pytestmark = pytest.mark.gpu


def _clouds(v, n, seed):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(v), n)
    xyz = v[idx] + rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
    xyz[rng.integers(0, n, n // 20)] = xyz[rng.integers(0, n, n // 20)]  # exact duplicates at other indices
    xyz[rng.integers(0, n, 50)] = np.nan
    xyz[rng.integers(0, n, 200)] *= np.float32(-40.0)  # behind the camera or off screen
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    col = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    a = {"xyz": xyz.astype(np.float32), "normals": nrm, "colors": col, "db_id": J.make_db_id(2, 1)}
    # cloud b repeats some of a's points exactly (a keeps those ties) and adds its own
    m = n // 4
    take = rng.integers(0, n, m)
    b_xyz = np.concatenate([a["xyz"][take], v[rng.integers(0, len(v), m)] * np.float32(1.02)])
    b = {"xyz": b_xyz, "colors": rng.integers(0, 2 ** 32, 2 * m, dtype=np.uint64).astype(np.uint32),
         "db_id": J.make_db_id(2, 2)}
    # cloud c: no colours, normals, a pose (rotation about z and a shift)
    ang = 0.3
    cs = np.eye(4, dtype=np.float32)
    cs[0, 0], cs[0, 1], cs[1, 0], cs[1, 1] = np.cos(ang), np.sin(ang), -np.sin(ang), np.cos(ang)
    cs[3, :3] = (0.1, -0.05, 0.2)
    c = {"xyz": v[rng.integers(0, len(v), m)], "normals": nrm[:m], "db_id": J.make_db_id(2, 3),
         "cs": cs.reshape(16)}
    return [a, b, c]


def _upload(r, cl):
    return r.upload_points(cl["xyz"], normals=cl.get("normals"), colors=cl.get("colors"), db_id=cl["db_id"],
                           cs=cl.get("cs"))


@pytest.mark.parametrize("w, h, n_points, flags", [
    (160, 120, 300000, ffi.EDGES | ffi.SHADING),
    (160, 120, 300000, ffi.EDGES),
    (320, 200, 200000, ffi.EDGES | ffi.SHADING | ffi.ONE_BIT),
    (1920, 1080, 2000000, ffi.EDGES | ffi.SHADING),
])
def test_splat_matches_oracle(renderer, w, h, n_points, flags):
    v, t = J.torus(120, 60, seed=8)
    slot = renderer.upload_mesh(v, t)
    clouds = _clouds(v, n_points, seed=w + n_points)
    ids = [_upload(renderer, c) for c in clouds]
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.orbit_camera(3.0, 0.7, 0.4)
    P, _ = J.projection(cam, w, h)
    fr = J.make_frame(w, h, cam, cs, ci, np.zeros(3, np.float32), 2.0, [slot], flags=flags)
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    tg = renderer.create_target(w, h)
    try:
        g0, im0 = renderer.render(fr, tg)
        renderer.splat_points(tg, P, ci, flags, ids)
        g1, im1 = renderer.read_target(tg)
        og, oim = O.splat_points(g0, im0, P, ci, flags, clouds)
        for k in og.dtype.names:
            a, b = g1[k], og[k]
            if a.dtype.kind == "f":
                a, b = a.view(np.uint32), b.view(np.uint32)
            assert np.array_equal(a, b), (k, int((a != b).sum()))
        assert np.array_equal(im1, oim), int((im1 != oim).sum())
        drawn = (og["db_id"] >> 29) == 2
        n_b = int((og["db_id"] == clouds[1]["db_id"]).sum())
        n_c = int((og["db_id"] == clouds[2]["db_id"]).sum())
        print("%dx%d, %d points: %d pixels drawn (%d from the second cloud, %d from the third), "
              "%d surface pixels kept" % (w, h, n_points, int(drawn.sum()), n_b, n_c,
                                          int(((g0["db_id"] != 0) & ~drawn).sum())))
        assert drawn.sum() > 0.05 * w * h and n_b > 0 and n_c > 0
        assert ((g0["db_id"] != 0) & ~drawn).any()  # the surface hides some pixels' points
        # a second call starts again from the G-buffer as it now is: points drawn
        # there hold depth 1/(1/w), against which they may be redrawn, so only
        # the oracle on the same input decides
        renderer.splat_points(tg, P, ci, flags, ids[:1])
        g2, im2 = renderer.read_target(tg)
        og2, oim2 = O.splat_points(g1, im1, P, ci, flags, clouds[:1])
        assert np.array_equal(g2.view(np.uint8), og2.view(np.uint8)) and np.array_equal(im2, oim2)
    finally:
        renderer.destroy_target(tg)
        for i in ids:
            renderer.remove_points(i)
        renderer.remove_mesh(slot)


def test_splat_no_clouds_empty_cloud_and_errors(renderer):
    w, h = 64, 48
    v, t = J.torus(40, 20, seed=2)
    slot = renderer.upload_mesh(v, t)
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.orbit_camera(3.0, 0.2, 0.3)
    P, _ = J.projection(cam, w, h)
    fr = J.make_frame(w, h, cam, cs, ci, np.zeros(3, np.float32), 2.0, [slot])
    tg = renderer.create_target(w, h)
    empty = renderer.upload_points(np.zeros((0, 3), np.float32))
    try:
        g0, im0 = renderer.render(fr, tg)
        renderer.splat_points(tg, P, ci, 0, [])
        renderer.splat_points(tg, P, ci, 0, [empty])
        g1, im1 = renderer.read_target(tg)
        assert np.array_equal(g0.view(np.uint8), g1.view(np.uint8)) and np.array_equal(im0, im1)
        with pytest.raises(J.J3dgError):
            renderer.splat_points(tg, P, ci, 0, [empty + 1000])
        with pytest.raises(J.J3dgError):
            renderer.splat_points(tg + 1000, P, ci, 0, [empty])
        # a moved cloud (j3dg_set_points_cs) is drawn where its new pose puts it
        pts = v[::7].copy()
        c = renderer.upload_points(pts, db_id=J.make_db_id(2, 9))
        shift = np.eye(4, dtype=np.float32)
        shift[3, :3] = (0.05, 0.0, 0.0)
        renderer.set_points_cs(c, shift.reshape(16))
        renderer.splat_points(tg, P, ci, 0, [c])
        g2, im2 = renderer.read_target(tg)
        og, oim = O.splat_points(g1, im1, P, ci, 0, [{"xyz": pts, "db_id": J.make_db_id(2, 9),
                                                      "cs": shift.reshape(16)}])
        assert np.array_equal(g2.view(np.uint8), og.view(np.uint8)) and np.array_equal(im2, oim)
        renderer.remove_points(c)
    finally:
        renderer.remove_points(empty)
        renderer.destroy_target(tg)
        renderer.remove_mesh(slot)
