"""Pins of the oracle's hot-path branches that identity poses cannot see -- CPU only.

Each case is built so that its expected result follows from j3d/canvas.cpp by
hand (geometry in float64, constant matcaps, hand-made G-buffers), and so that
a plausible slip in the restatement gives a different answer:

* the G-buffer normal n = object_cs * (CS^-1 * n_tri) (canvas.cpp:790-792)
  under a rotated camera and a rotated object pose, where CS * n and the
  swapped order CS^-1 * (object_cs * n) differ;
* the point light at CS * (pivot + 3 * diagonal, 1) (canvas.cpp:703-705), the
  shadow ray from the hit point along the unnormalised light - pos with
  t in [1e-3, FLT_MAX] (canvas.cpp:848-853): tiny occluders placed on one
  pixel's shadow ray halfway to the light, beyond it, and within 1e-3 of the
  start in units of |light - pos|;
* the edge pass's up neighbour (canvas.cpp:596, 621-628, 647): row y - 1, and
  row 0 itself for y = 0;
* the wireframe grey (canvas.cpp:449-496) and the one-bit dot pattern
  (canvas.cpp:498-579) on hand-made G-buffers with constant matcaps.
"""
import math

import numpy as np
import pytest

import oracle as O
import pin_scenes as S

IDENT = np.eye(4, dtype=np.float32).reshape(16)


# ------------------------------------------- normal transform (a8) -------

def test_normal_transform_rotated_camera_and_pose():
    """u, v = xy of object_cs * (CS^-1 * n) for a plane seen by a rotated camera
    under a rotated object pose; CS * n and the swapped order are far off."""
    v, t, ocs, fr = S.normal_scene()
    m = O.Mesh(v, t, cs=ocs)
    assert np.array_equal(m.n, np.array([[0, 0, 1], [0, 0, 1]], np.float32))  # cross((2,0,0),(2,2,0)) / 4
    g, st = O.update_canvas([m], fr)
    assert st["hits"] == fr.width * fr.height
    n = np.array([0.0, 0.0, 1.0, 0.0])
    expect = S.mat(ocs) @ (S.mat(fr.cs_inv) @ n)
    wrong = {"CS n": S.mat(ocs) @ (S.mat(fr.cs) @ n), "CS^-1 (object_cs n)": S.mat(fr.cs_inv) @ (S.mat(ocs) @ n),
             "object_cs n": S.mat(ocs) @ n}
    for k, wv in wrong.items():
        assert np.hypot(wv[0] - expect[0], wv[1] - expect[1]) > 0.05, k  # the case tells these apart
    assert np.all(np.abs(g["u"].astype(np.float64) - expect[0]) <= 1e-6)
    assert np.all(np.abs(g["v"].astype(np.float64) - expect[1]) <= 1e-6)
    # depth is t along the unnormalised camera direction (z = -1): the view-space z of the hit
    Mo = S.mat(ocs)
    pn = Mo[:3, :3] @ np.array([0, 0, 1.0])
    p0 = Mo[:3, 3]
    for (x, y) in [(0, 0), (7, 5), (15, 11)]:
        o, d = S.pixel_ray(fr, x, y)
        tt = float((p0 - o) @ pn / (d @ pn))
        assert abs(float(g["depth"][y, x]) - tt) <= 2e-6 * tt


# --------------------------------------------- point light (a2, a10) -----

def _shadow_pixel(name):
    where, r, expect = S.SHADOW_CASES[name]
    v, t, fr, geo = S.shadow_scene(where, r)
    g, _ = O.update_canvas([O.Mesh(v, t)], fr)
    x, y = S.SHADOW_PIXEL
    px = g[y, x]
    assert px["object_id"] in (0, 1)  # the camera ray reaches the ground past the occluder
    assert S.line_distance(geo["c"], geo["o"], geo["d"]) > 2 * r
    return px, fr, geo, r, expect


def test_point_light_position():
    """An occluder halfway from the pixel's surface point to CS * (pivot + 3 diag)
    shadows it; the same occluder halfway to the light without CS, or with
    CS^-1, would not lie on the shadow ray."""
    px, fr, geo, r, expect = _shadow_pixel("halfway")
    P = geo["P"]
    p = np.asarray(fr.pivot, np.float64) + 3.0 * float(fr.diagonal)
    for wrong in (p, (S.mat(fr.cs_inv) @ np.append(p, 1.0))[:3]):
        assert S.line_distance(geo["c"], P, wrong - P) > 3 * r
    assert px["mark"] & 1 == expect == 1


def test_point_light_ray_continues_past_the_light():
    """t_far = FLT_MAX (canvas.cpp:853): an occluder beyond the light still shadows."""
    px, _, _, _, expect = _shadow_pixel("beyond")
    assert px["mark"] & 1 == expect == 1


@pytest.mark.parametrize("name", ["near_0.02", "near_0.3"])
def test_point_light_direction_is_unnormalised(name):
    """t_near = 1e-3 counts in units of |light - pos| (the direction is not
    normalised, canvas.cpp:848): an occluder 0.02 from the surface point toward
    the light lies at t < 1e-3 and is skipped; one at 0.3 is not.  With a
    normalised direction both would shadow."""
    px, _, geo, _, expect = _shadow_pixel(name)
    delta = float(np.linalg.norm(geo["c"] - geo["P"]))
    dist = float(np.linalg.norm(geo["L"] - geo["P"]))
    assert delta > 1e-3 and (delta / dist < 1e-3) == (expect == 0)
    assert px["mark"] & 1 == expect


# ------------------------------------------- edge pass up neighbour ------

def _texel(mc, u, v):
    U = int(math.floor(0.5 + (float(np.float32(u)) + 1) * 511 * 0.5))
    V = int(math.floor(0.5 + (1 - float(np.float32(v))) * 511 * 0.5))
    return int(mc[V, U])


def _edge_color(p1, p2, x1, y1, x2, y2, fr, mc, cav):
    """get_angle_color(compute_convex_cos_angle(...)) (canvas.cpp:287-389) in float64
    for pixel p1 at (x1, y1) and its neighbour p2 at (x2, y2)."""
    Pi = np.asarray(fr.proj_inv, np.float64)

    def ray(xx, yy):
        sx = 2.0 * ((xx + 0.5) / fr.width) - 1.0
        sy = 2.0 * ((yy + 0.5) / fr.height) - 1.0
        return np.array([Pi[0] * sx, Pi[5] * sy, -1.0])
    u1, v1, u2, v2 = (float(p1["u"]), float(p1["v"]), float(p2["u"]), float(p2["v"]))
    a = float(p1["depth"]) * ray(x1, y1)
    b = float(p2["depth"]) * ray(x2, y2)
    n1 = np.array([u1, v1, math.sqrt(1 - u1 * u1 - v1 * v1)])
    n2 = np.array([u2, v2, math.sqrt(1 - u2 * u2 - v2 * v2)])
    dd = (b - a) / np.linalg.norm(b - a)
    angle = float(dd @ n1) if abs(n1 @ n2 - 1) > 1e-4 else 0.0
    tex = _texel(mc, u1, v1)
    hp = 1.57079632679489
    s = math.sqrt(1 - (hp - abs(math.acos(angle) - hp)) / hp)
    k = 1.2 if angle > 0 else 0.5
    ch = [min(255, int(((tex >> sh) & 0xff) * (1 - s) + ((cav >> sh) & 0xff) * s * k)) for sh in (0, 8, 16)]
    return angle, ch


def _bytes(c):
    return [(int(c) >> sh) & 0xff for sh in (0, 8, 16)]


@pytest.mark.parametrize("step", [+0.4, -0.4])
def test_edge_pass_up_neighbour_is_the_row_above(step):
    """Rows 0-1 and rows 2-4 differ in normal and depth.  Row 2 finds its edge
    in row 1 (the right neighbours are equal), computed at screen (x, y - 1);
    rows 0, 1, 3, 4 see equal up neighbours and stay plain, row 0 reading
    itself."""
    g = S.row_step_gbuf(step)
    h, w = g.shape
    mc, cav = O.matcap(0)
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, O.EDGES | O.SHADING)
    im = O.canvas_to_image(g, fr, mc, cav)
    plain_a, plain_b = _texel(mc, 0.1, -0.2), _texel(mc, 0.25, -0.2)
    for x in range(w - 1):
        angle, exp = _edge_color(g[2, x], g[1, x], x, 2, x, 1, fr, mc, cav)
        _, wrong = _edge_color(g[2, x], g[1, x], x, 2, x, 3, fr, mc, cav)  # screen row y + 1 instead
        assert angle != 0 and max(abs(a - b) for a, b in zip(wrong, exp)) >= 2
        assert max(abs(a - b) for a, b in zip(_bytes(im[2, x]), exp)) <= 1, (x, _bytes(im[2, x]), exp)
        # a farther pixel under a nearer row is a concave step (angle > 0), a nearer one convex
        assert (angle > 0) == (step > 0)
    assert np.all(im[0:2, :] == plain_a)
    assert np.all(im[3:, :] == plain_b)
    assert im[2, w - 1] == plain_b  # the last column is plain


def test_edge_pass_row_zero_reads_itself():
    """For y = 0 the 'up' pixel is row 0 itself (canvas.cpp:596, 647): a row 0
    that differs from row 1 gets no edge from it, while row 1 does."""
    g = S.row_zero_gbuf()
    h, w = g.shape
    mc, cav = O.matcap(0)
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, O.EDGES | O.SHADING)
    im = O.canvas_to_image(g, fr, mc, cav)
    assert np.all(im[0, :] == _texel(mc, 0.3, 0.1))
    for x in range(w - 1):
        _, exp = _edge_color(g[1, x], g[0, x], x, 1, x, 0, fr, mc, cav)
        assert max(abs(a - b) for a, b in zip(_bytes(im[1, x]), exp)) <= 1
        assert im[1, x] != _texel(mc, -0.2, 0.15)
    assert np.all(im[2, :] == _texel(mc, -0.2, 0.15))


# --------------------------------------------------- wireframe / one-bit -

def test_wireframe_grey_from_id_changes():
    """canvas.cpp:449-496: a hit pixel whose right or up neighbour is a hit with
    another triangle id is grey (255 * (u^2 + v^2) / 2, truncated); misses and
    equal ids leave the plain matcap colour; the last column is plain; row 0's
    up neighbour is itself."""
    g = S.wire_gbuf()
    h, w = g.shape
    tex, cav = 0xff5a6b7c, 0xff102030
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, O.WIREFRAME | O.SHADING)
    im = O.canvas_to_image(g, fr, S.const_matcap(tex), cav)
    grey = 0xff000000 | (45 << 16) | (45 << 8) | 45  # (unsigned char)(255 * 0.18f) = 45
    dark = 0xff000000 | ((0x5a >> 2) << 16) | ((0x6b >> 2) << 8) | (0x7c >> 2)
    bg = O.background(w, h)
    # Row by row, from the rule (right != own and right hit) or (up != own and up hit):
    # row 0: x=1 right 8 != 7 -> grey; x=0 right 7, up itself -> plain; x=2,3 right 8 -> plain; x=4 last column.
    # row 1: x=0 right 7, up 7 -> plain; x=1 right 7, up 7 -> plain; x=2 right 8 -> grey;
    #        x=3 right 8, up 8 -> plain; x=4 last column -> plain.
    # row 2: x=0 miss -> background; x=1 right 7, up 7 -> plain; x=2 right 7, up 7 -> plain;
    #        x=3 right 7, up 8 -> grey; x=4 last column -> plain.
    # row 3: x=0 right 9 == 9, up miss -> plain, shadowed; x=1 right miss, up 7 != 9 -> grey;
    #        x=2 miss -> background; x=3 right 3 != 7 -> grey; x=4 last column -> plain.
    exp = np.array([
        [tex, grey, tex, tex, tex],
        [tex, tex, grey, tex, tex],
        [bg[2, 0], tex, tex, grey, tex],
        [dark, grey, bg[3, 2], grey, tex],
    ], np.uint32)
    assert np.array_equal(im, exp), (im, exp)


@pytest.mark.parametrize("tex,res", [(0xff808080, 4), (0xff7f7f7f, 3), (0xff000000, 1)])
def test_one_bit_dot_pattern(tex, res):
    """canvas.cpp:498-579 on a flat G-buffer with a constant matcap:
    res = ((R + G + B) >> 7) + 1, a pixel is black iff x % res == 0 and
    y % res == 0 (the last column uses w - 1), and a shadowed texel (channels
    >> 2) changes res."""
    w, h = 9, 7
    mc = S.const_matcap(tex)
    g = S.flat_gbuf(w, h, u=0.0, v=0.0)
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, O.ONE_BIT)
    im = O.canvas_to_image(g, fr, mc, 0xff000000)
    ys, xs = np.mgrid[0:h, 0:w]
    exp = np.where((xs % res == 0) & (ys % res == 0), 0xff000000, 0xffffffff).astype(np.uint32)
    assert np.array_equal(im, exp)
    g["mark"] = 1
    im = O.canvas_to_image(g, fr, mc, 0xff000000)
    res_s = ((((tex & 0xff) >> 2) * 3) >> 7) + 1
    exp = np.where((xs % res_s == 0) & (ys % res_s == 0), 0xff000000, 0xffffffff).astype(np.uint32)
    assert np.array_equal(im, exp)


def one_bit_edge_expected(tex, res, g):
    """The one-bit image of S.one_bit_edge_gbuf() with a constant matcap, from
    canvas.cpp:498-579 by hand: columns 2 | 3 differ by u 0 -> 0.5
    (n.n' = sqrt(0.75) = 0.866 < 0.95); row 3 has u = 0.2 (then 0.5 from
    column 3), so its up neighbour differs by 0.2 (n.n' = sqrt(0.96) = 0.98,
    no forced black) and its column-2 edge is 0.2 vs 0.5:
    0.1 + sqrt(0.96 * 0.75) = 0.9485 < 0.95."""
    h, w = g.shape
    ys, xs = np.mgrid[0:h, 0:w]
    dot = (xs % res == 0) & (ys % res == 0)
    black = dot.copy()
    edge = np.zeros((h, w), bool)
    edge[:, 2] = True
    black[edge] = (~dot[edge]) if res == 1 else True
    black[:, w - 1] = ((w - 1) % res == 0) & (ys[:, w - 1] % res == 0)
    return np.where(black, 0xff000000, 0xffffffff).astype(np.uint32)


def test_one_bit_edges():
    """Where the right (else the up) neighbour's normal differs by more than
    0.001, angle = n1 . n2; |angle| < 0.95 makes the pixel black (res > 1) or
    inverts the dot (res = 1); 0.95 <= |angle| keeps the dot pattern."""
    g = S.one_bit_edge_gbuf()
    h, w = g.shape
    fr = S.Frame(w, h, IDENT, (0, 0, 0), 1.0, O.ONE_BIT)
    for tex, res in [(0xff808080, 4), (0xff000000, 1)]:
        im = O.canvas_to_image(g, fr, S.const_matcap(tex), 0xff000000)
        exp = one_bit_edge_expected(tex, res, g)
        assert np.array_equal(im, exp), (tex, im, exp)


# ------------------------------------------------- scene framing ----------

def test_prepare_and_unzoom_closed_form():
    """prepare_scene + unzoom (j3d/scene.cpp:50-110): the union of the world
    boxes (each local box's two corners through its cs, then min / max per
    axis), diagonal = largest extent, pivot = box centre, camera at
    pivot + (0, 0, 2 diagonal) looking down -z."""
    t = S.rigid(np.eye(3), [1.0, 0.0, 0.0])
    rz = S.rigid(S.rot("z", math.pi / 2), [0.0, 0.0, 0.0])  # (x, y) -> (-y, x)
    d, pivot, cs, ci = O.prepare_and_unzoom([(t, (0, 0, 0), (1, 2, 3)), (rz, (-1, -1, -1), (0, 0, 0.5))])
    # world boxes: [1,2]x[0,2]x[0,3] and corners (1,-1,-1), (~0,0,0.5) -> [~0,1]x[-1,0]x[-1,0.5]
    lo, hi = np.array([0.0, -1.0, -1.0]), np.array([2.0, 2.0, 3.0])
    assert abs(d - 4.0) < 1e-6
    assert np.allclose(pivot, (lo + hi) / 2, atol=1e-6)
    assert np.allclose(S.mat(cs), [[1, 0, 0, 1.0], [0, 1, 0, 0.5], [0, 0, 1, 1.0 + 8.0], [0, 0, 0, 1]], atol=1e-6)
    assert np.allclose(S.mat(ci) @ S.mat(cs), np.eye(4), atol=1e-6)
    # one box alone (a scene of one point cloud): its own extent
    d1, p1, _, _ = O.prepare_and_unzoom([(t, (0, 0, 0), (1, 2, 3))])
    assert d1 == 3.0 and np.array_equal(p1, np.array([1.5, 1.0, 1.5], np.float32))
    # the order of the boxes does not change the union
    d2, p2, _, _ = O.prepare_and_unzoom([(rz, (-1, -1, -1), (0, 0, 0.5)), (t, (0, 0, 0), (1, 2, 3))])
    assert d2 == d and np.array_equal(p2, pivot)
