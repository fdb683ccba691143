"""The oracle's point-cloud splat (canvas::render_pointclouds_on_image,
j3d/canvas.cpp:952-1030) pinned by closed forms: a point placed on the
ray-cast line of a pixel centre lands in that pixel with depth = its view
depth; depth order and the strict test (earlier point, earlier cloud and the
surface keep ties); the frustum test; the colour rules of canvas.cpp:994-997.
jtk's point renderer itself is absent, so parity with it is unpinned."""
import numpy as np

import oracle as O


def _pose(seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    m = np.eye(4)
    m[:3, :3] = q
    m[:3, 3] = rng.uniform(-2, 2, 3)
    cs = np.ascontiguousarray(m.T, np.float32).reshape(16)
    ci = np.ascontiguousarray(np.linalg.inv(m).T, np.float32).reshape(16)
    return cs, ci, m


def _view(w, h, seed=1):
    P, Pi = O.projection(O.default_camera(), w, h)
    cs, ci, m = _pose(seed)
    return P, Pi, cs, ci, m


def _empty(w, h, rgba=0xff102030):
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    g["object_id"] = 0xFFFFFFFF
    return g, np.full((h, w), rgba, np.uint32)


def _points_on_pixels(Pi, m, w, h, depth, order):
    """World points on the primary ray of each pixel centre (canvas.cpp:773-776)
    at view depth `depth`, float64 then rounded once."""
    pi = Pi.reshape(4, 4).T.astype(np.float64)
    ys, xs = np.divmod(order, w)
    ndc = np.stack([2 * (xs + 0.5) / w - 1, 2 * (ys + 0.5) / h - 1, np.full(len(order), 0.1), np.ones(len(order))])
    d = pi @ ndc
    d[3] = 0.0
    pc = d * depth  # the ray direction has view z = -1: camera-space point at view depth `depth`
    pc[3] = 1.0
    return (m @ pc)[:3].T.astype(np.float32)


def test_point_on_pixel_ray_lands_in_that_pixel():
    w, h = 48, 36
    P, Pi, cs, ci, m = _view(w, h)
    rng = np.random.default_rng(2)
    order = rng.permutation(w * h)
    depth = rng.uniform(0.5, 50.0, w * h)
    xyz = _points_on_pixels(Pi, m, w, h, depth, order)
    g, im = _empty(w, h)
    og, oim = O.splat_points(g, im, P, ci, 0, [{"xyz": xyz, "db_id": 7}])
    flat = og.reshape(-1)
    assert np.array_equal(flat["object_id"][order], np.arange(w * h))
    assert np.all(flat["db_id"] == 7)
    rel = np.abs(flat["depth"][order].astype(np.float64) - depth) / depth
    assert rel.max() < 1e-5
    assert np.all(oim == 0xffffffff)  # no colours: the colour passed to present


def test_nearest_point_wins_and_ties_keep_the_first():
    w, h = 16, 12
    P, Pi, cs, ci, m = _view(w, h, seed=3)
    pix = np.array([5 * w + 7] * 4)
    xyz = _points_on_pixels(Pi, m, w, h, np.array([4.0, 2.0, 3.0, 2.0]), pix)
    xyz[3] = xyz[1]  # an exact duplicate of point 1, later
    cols = np.array([0xff0000ff, 0xff00ff00, 0xffff0000, 0xff00ffff], np.uint32)
    g, im = _empty(w, h)
    og, oim = O.splat_points(g, im, P, ci, 0, [{"xyz": xyz, "colors": cols, "db_id": 1},
                                               {"xyz": xyz[1:2], "colors": cols[3:], "db_id": 2}])
    p = og.reshape(-1)[pix[0]]
    assert p["object_id"] == 1 and p["db_id"] == 1
    assert oim.reshape(-1)[pix[0]] == cols[1]
    assert np.count_nonzero(og["db_id"]) == 1


def test_surface_depth_test():
    """A point in front of the surface (view depth 2 < 5) is drawn, one behind
    it (6 > 5) is not; pixels without a surface (db_id 0) take any point."""
    w, h = 16, 12
    P, Pi, cs, ci, m = _view(w, h, seed=4)
    pix = np.array([3 * w + 2, 4 * w + 9, 8 * w + 1])
    xyz = _points_on_pixels(Pi, m, w, h, np.array([2.0, 6.0, 50.0]), pix)
    g, im = _empty(w, h)
    for i in pix[:2]:
        g.reshape(-1)[i]["db_id"] = 99
        g.reshape(-1)[i]["object_id"] = 5
        g.reshape(-1)[i]["depth"] = 5.0
    og, _ = O.splat_points(g, im, P, ci, 0, [{"xyz": xyz, "db_id": 3}])
    f = og.reshape(-1)
    assert f[pix[0]]["db_id"] == 3 and f[pix[0]]["object_id"] == 0
    assert f[pix[1]]["db_id"] == 99 and f[pix[1]]["object_id"] == 5 and f[pix[1]]["depth"] == 5.0
    assert f[pix[2]]["db_id"] == 3 and f[pix[2]]["object_id"] == 2


def test_side_planes_and_non_finite_points_are_dropped():
    """Behind the camera, outside the four side planes, or NaN: not drawn.  The
    default far plane is FLT_MAX, so the frustum's z row is infinite and clip z
    is not tested: a point at view depth 1e30 is drawn."""
    w, h = 16, 12
    P, Pi, cs, ci, m = _view(w, h, seed=5)
    assert np.isinf(P[14])
    pix = np.array([6 * w + 6, 2 * w + 3])
    front = _points_on_pixels(Pi, m, w, h, np.array([3.0]), pix[:1])[0]
    far = _points_on_pixels(Pi, m, w, h, np.array([1e30]), pix[1:])[0]
    behind = _points_on_pixels(Pi, m, w, h, np.array([-3.0]), pix[:1])[0]  # behind the camera
    pi = Pi.reshape(4, 4).T.astype(np.float64)
    side = []
    for nx, ny in ((1.5, 0.0), (0.0, -1.2), (-1.01, 0.3)):
        d = pi @ np.array([nx, ny, 0.1, 1.0])
        d[3] = 0.0
        pc = d * 3.0
        pc[3] = 1.0
        side.append((m @ pc)[:3])
    nan = np.array([np.nan, 0, 0], np.float32)
    xyz = np.vstack([behind, np.array(side, np.float32), nan, front, far]).astype(np.float32)
    g, im = _empty(w, h)
    og, _ = O.splat_points(g, im, P, ci, 0, [{"xyz": xyz, "db_id": 4}])
    assert np.count_nonzero(og["db_id"]) == 2
    assert og.reshape(-1)[pix[0]]["object_id"] == 5
    assert og.reshape(-1)[pix[1]]["object_id"] == 6


def test_colour_rules():
    """canvas.cpp:994-995: normals only with shading, colours not in one-bit
    mode; a normal toward the camera keeps the colour, one across the view
    blackens it, alpha stays."""
    w, h = 8, 8
    P, Pi, cs, ci, m = _view(w, h, seed=6)
    pix = np.array([1 * w + 1, 2 * w + 5, 6 * w + 3])
    xyz = _points_on_pixels(Pi, m, w, h, np.array([3.0, 3.0, 3.0]), pix)
    rot = m[:3, :3]
    n_cam = np.array([[0, 0, 1], [1, 0, 0], [0, 0.6, 0.8]], np.float64)
    nrm = (rot @ n_cam.T).T.astype(np.float32)  # world normals whose camera-space images are n_cam
    cols = np.array([0x80c86432, 0x80c86432, 0x80c86432], np.uint32)
    g, im = _empty(w, h)
    cloud = {"xyz": xyz, "normals": nrm, "colors": cols, "db_id": 1}
    _, plain = O.splat_points(g, im, P, ci, 0, [cloud])
    _, shaded = O.splat_points(g, im, P, ci, O.SHADING, [cloud])
    _, onebit = O.splat_points(g, im, P, ci, O.ONE_BIT, [cloud])
    _, both = O.splat_points(g, im, P, ci, O.ONE_BIT | O.SHADING, [cloud])
    f = lambda a: a.reshape(-1)[pix]
    assert np.all(f(plain) == cols)
    s = f(shaded)
    assert s[0] in (0x80c86432, 0x80c76331, 0x80c76332, 0x80c76432, 0x80c86431, 0x80c86331, 0x80c76431, 0x80c86332)
    assert s[1] == 0x80000000
    ch = [(int(s[2]) >> (8 * j)) & 0xFF for j in range(3)]
    assert all(abs(c - int(0.8 * v)) <= 1 for c, v in zip(ch, (0x32, 0x64, 0xc8)))
    assert np.all(f(onebit) == 0xffffffff)
    assert f(both)[1] == 0xff000000 and f(both)[0] in (0xffffffff, 0xfffefefe, 0xfffffefe, 0xfffefeff, 0xfffeffff,
                                                       0xfffffeff, 0xfffefffe, 0xfffffffe)
