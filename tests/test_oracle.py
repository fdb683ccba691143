"""Pins the CPU oracle (oracle/oracle.cpp) -- CPU only.

The reference ships no tests or golden vectors (its jtk submodule, which holds
the traversal code, is absent), so each oracle function is pinned by what the
mathematics fixes: closed forms of the camera model (j3d/camera.cpp), known
answers of the ray/triangle test, a float64 brute force for closest hits,
invariants of ray generation, and special cases of the edge pass whose result
follows from j3d/canvas.cpp by hand.  A plausible slip (a dropped term, a
wrong sign or index, a swapped operand) fails at least one of these.
"""
import math

import numpy as np
import pytest

import oracle as O

F32MAX = np.finfo(np.float32).max


# ------------------------------------------------------------------ camera --

def test_default_camera_fields():
    c = O.default_camera()
    assert c.focal_length == np.float32(35) and c.near_plane == np.float32(0.1)
    assert c.far_plane == F32MAX and c.fit_film == 1 and c.zoom == 1
    assert c.film_w_inch == np.float32(1.024) and c.film_h_inch == np.float32(0.768)


@pytest.mark.parametrize("w,h", [(1024, 768), (1920, 1080), (256, 256), (800, 600), (640, 480), (300, 900)])
def test_projection_closed_form(w, h):
    c = O.default_camera()
    P, Pi = O.projection(c, w, h)
    f32 = np.float32
    n = f32(0.1)
    top = (f32(0.768) * f32(25.4) / f32(2)) / f32(35) * n
    right = (f32(1.024) * f32(25.4) / f32(2)) / f32(35) * n
    film = f32(1.024) / f32(0.768)
    dev = f32(w) / f32(h)
    # Overscan: the axis that is too short for the device is widened
    if film > dev:
        top = top * (film / dev)
    else:
        right = right * (dev / film)
    assert P[0] == f32(2) * n / (right + right)
    assert P[5] == -f32(2) * n / (top + top)
    assert P[8] == 0 and P[9] == 0 and P[11] == -1
    assert P[10] == -1  # -(FLT_MAX + n) / (FLT_MAX - n) rounds to -1
    assert P[14] == -np.inf  # 2 * FLT_MAX overflows
    assert Pi[0] == f32(1) / P[0] and Pi[5] == f32(1) / P[5]
    assert Pi[11] == 0 and math.copysign(1.0, Pi[11]) < 0  # 1 / -inf = -0
    assert Pi[14] == -1 and Pi[15] == 0 and Pi[12] == 0 and Pi[13] == 0
    assert np.all(np.isfinite(Pi))


def test_ray_generation_invariants():
    """dir_cam.z == -1 exactly for every pixel; row 0 is the image top."""
    c = O.default_camera()
    w, h = 64, 48
    _, Pi = O.projection(c, w, h)
    ys = []
    for y in range(h):
        for x in (0, w // 2, w - 1):
            sx = np.float32(2) * ((np.float32(x) + np.float32(0.5)) / np.float32(w)) - np.float32(1)
            sy = np.float32(2) * ((np.float32(y) + np.float32(0.5)) / np.float32(h)) - np.float32(1)
            d = O.mat_vec(Pi, np.array([sx, sy, c.near_plane, 1], np.float32))
            assert d[2] == -1.0
            assert d[0] == Pi[0] * sx  # only the x term survives
            ys.append(d[1])
    assert ys[0] > 0 > ys[-1]  # P[5] < 0 flips y: row 0 looks up


def test_invert_orthonormal():
    a = 0.7
    cs = np.array([math.cos(a), math.sin(a), 0, 0, -math.sin(a), math.cos(a), 0, 0, 0, 0, 1, 0, 1.5, -2, 3, 1],
                  np.float32)
    inv = O.invert_orthonormal(cs)
    prod = cs.reshape(4, 4).T.astype(np.float64) @ inv.reshape(4, 4).T.astype(np.float64)
    assert np.allclose(prod, np.eye(4), atol=1e-6)


# ---------------------------------------------------------- ray/triangle --

V0, V1, V2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]


def test_triangle_known_answer():
    hit, t, u, v = O.intersect_triangle([0.25, 0.25, 3], [0, 0, -1], V0, V1, V2)
    assert hit and (t, u, v) == (3.0, 0.25, 0.25)
    # (u, v) are the weights of v1 and v2
    hit, t, u, v = O.intersect_triangle([0.75, 0.125, 2], [0, 0, -1], V0, V1, V2)
    assert hit and (t, u, v) == (2.0, 0.75, 0.125)


def test_triangle_two_sided_and_inclusive():
    assert O.intersect_triangle([0.25, 0.25, -3], [0, 0, 1], V0, V1, V2)[0]      # from below
    assert O.intersect_triangle([0.25, 0.25, 3], [0, 0, -1], V0, V2, V1)[0]      # reversed winding
    assert O.intersect_triangle([0, 0, 3], [0, 0, -1], V0, V1, V2)[0]            # vertex
    assert O.intersect_triangle([0.5, 0.5, 3], [0, 0, -1], V0, V1, V2)[0]        # u + v == 1
    assert O.intersect_triangle([0.5, 0, 3], [0, 0, -1], V0, V1, V2)[0]          # v == 0 edge
    assert not O.intersect_triangle([0.5, 0.5001, 3], [0, 0, -1], V0, V1, V2)[0]
    assert not O.intersect_triangle([-1e-6, 0.5, 3], [0, 0, -1], V0, V1, V2)[0]


def test_triangle_range_inclusive():
    o, d = [0.25, 0.25, 3], [0, 0, -1]
    assert O.intersect_triangle(o, d, V0, V1, V2, t_near=3.0)[0]
    assert not O.intersect_triangle(o, d, V0, V1, V2, t_near=float(np.nextafter(np.float32(3), np.float32(4))))[0]
    assert O.intersect_triangle(o, d, V0, V1, V2, t_far=3.0)[0]
    assert not O.intersect_triangle(o, d, V0, V1, V2, t_far=float(np.nextafter(np.float32(3), np.float32(0))))[0]
    assert not O.intersect_triangle([0.25, 0.25, -3], [0, 0, -1], V0, V1, V2)[0]  # behind the origin


def test_triangle_degenerate_and_nan():
    assert not O.intersect_triangle([0.25, 0.25, 3], [1, 0, 0], V0, V1, V2)[0]  # parallel: det == 0
    assert not O.intersect_triangle([0.25, 0.25, 3], [0, 0, -1], V0, V1, [2, 0, 0])[0]  # collinear triangle
    assert not O.intersect_triangle([0.25, 0.25, 3], [np.nan, 0, -1], V0, V1, V2)[0]


def test_triangle_unnormalised_direction():
    """t is in units of the (unnormalised) direction, as the reference's depth is."""
    hit, t, u, v = O.intersect_triangle([0.25, 0.25, 3], [0, 0, -2], V0, V1, V2)
    assert hit and t == 1.5


def _f64_hits(v, t, rays):
    """All intersections in float64 (Moller-Trumbore, no tie-break): (n_rays, n_tris) t, inf for misses."""
    v = v.astype(np.float64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    e1, e2 = b - a, c - a
    o = rays[:, None, 0:3].astype(np.float64)
    d = rays[:, None, 3:6].astype(np.float64)
    p = np.cross(d, e2[None])
    det = np.einsum("rkj,kj->rk", p, e1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        s = o - a[None]
        u = np.einsum("rkj,rkj->rk", s, p) * inv
        q = np.cross(s, e1[None])
        vv = np.einsum("rkj,rkj->rk", np.broadcast_to(d, q.shape), q) * inv
        tt = np.einsum("kj,rkj->rk", e2, q) * inv
    ok = (u >= 0) & (vv >= 0) & (u + vv <= 1) & (tt >= rays[:, None, 6]) & (tt <= rays[:, None, 7])
    return np.where(ok, tt, np.inf)


def _rays(n, seed, radius=1.2):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = 3.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = rng.uniform(-radius, radius, size=(n, 3)) - o
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, 1e-3, F32MAX
    return r


def _icosahedron_subdivided(level):
    import j3d_amd as J
    return J.icosphere(level, seed=5, amplitude=0.05)


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_closest_hit_matches_float64_brute_force(level):
    v, t = _icosahedron_subdivided(level)
    rays = _rays(400, level)
    got = O.closest_hits([O.Mesh(v, t)], rays)
    allt = _f64_hits(v, t, rays)
    best = allt.min(axis=1)
    hit = np.isfinite(best)
    assert np.array_equal(got["tri"] != 0xFFFFFFFF, hit)
    assert np.allclose(got["t"][hit], best[hit], rtol=1e-5)
    # the chosen triangle is a float64-valid hit within 1e-5 of the minimum
    chosen = allt[np.arange(len(rays))[hit], got["tri"][hit].astype(np.int64)]
    assert np.allclose(chosen, best[hit], rtol=1e-5)


@pytest.mark.parametrize("level", [1, 3])
def test_bvh_equals_brute_force_bitwise(level):
    v, t = _icosahedron_subdivided(level)
    m = O.Mesh(v, t)
    rays = _rays(3000, 10 + level)
    a = O.closest_hits([m], rays)
    b = O.closest_hits([m], rays, brute=True)
    for k in ("obj", "tri"):
        assert np.array_equal(a[k], b[k])
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32))
    assert np.array_equal(O.any_hits([m], rays), O.any_hits([m], rays, brute=True))


def test_tie_break_lowest_triangle_id():
    """Two identical triangles: the lexicographic (t, object, id) minimum picks id 0."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 1, 2]], np.uint32)
    r = np.array([[0.25, 0.25, 3, 0, 0, -1, 0, F32MAX]], np.float32)
    got = O.closest_hits([O.Mesh(v, t)], r, second=True)
    assert got["tri"][0] == 0 and got["t2"][0] == got["t"][0]
    got2 = O.closest_hits([O.Mesh(v, t), O.Mesh(v, t)], r)
    assert got2["obj"][0] == 0


def test_second_nearest_distance():
    v, t = _icosahedron_subdivided(2)
    rays = _rays(300, 77)
    got = O.closest_hits([O.Mesh(v, t)], rays, second=True)
    allt = _f64_hits(v, t, rays)
    srt = np.sort(allt, axis=1)
    has2 = np.isfinite(srt[:, 1])
    assert np.allclose(got["t2"][has2], srt[has2, 1], rtol=1e-5)
    assert np.all(got["t2"][~has2] == F32MAX)


def test_two_level_object_transform():
    """An object translated by its cs is hit where the translated geometry is."""
    v = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = -2.0
    r = np.array([[0.1, 0.2, 5, 0, 0, -1, 0, F32MAX]], np.float32)
    got = O.closest_hits([O.Mesh(v, t, cs=cs)], r)
    assert got["t"][0] == 7.0


# ------------------------------------------------------------- frames -----

class _Frame:
    def __init__(self, w, h, cs, pivot, diag, flags, proj_inv, near=0.1, light=(0, 0, 1)):
        self.width, self.height = w, h
        self.rect = (0, 0, w - 1, h - 1)
        self.proj_inv = proj_inv
        self.near_plane = near
        self.cs = cs
        self.cs_inv = O.invert_orthonormal(cs)
        self.pivot = np.asarray(pivot, np.float32)
        self.diagonal = diag
        self.flags = flags
        self.light_dir = np.asarray(light, np.float32)
        self.bg_top, self.bg_bottom = 0xff000000, 0xff404040


def _plane(z=0.0, size=10.0):
    v = np.array([[-size, -size, z], [size, -size, z], [size, size, z], [-size, size, z]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, t


def test_depth_is_view_z_on_facing_plane():
    v, t = _plane()
    w, h = 32, 24
    _, Pi = O.projection(O.default_camera(), w, h)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = 3.0
    fr = _Frame(w, h, cs, (0, 0, 0), 20.0, O.EDGES | O.SHADING, Pi)
    g, st = O.update_canvas([O.Mesh(v, t)], fr)
    assert st["hits"] == w * h
    d = g["depth"].astype(np.float64)
    assert np.all(np.abs(d - 3.0) <= 3.0 * 2.4e-7)  # within an ulp of the view-space z
    assert np.all(g["u"] == 0) and np.all(g["v"] == 0) and np.all(g["mark"] == 0)


def test_miss_pixels_and_untouched_fields():
    v, t = _plane(z=0.0, size=0.1)
    w, h = 32, 24
    _, Pi = O.projection(O.default_camera(), w, h)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = 3.0
    fr = _Frame(w, h, cs, (0, 0, 0), 0.2, O.EDGES, Pi)
    g0 = np.zeros((h, w), O.PIXEL_DTYPE)
    g0["mark"], g0["r"], g0["barycentric_u"] = 7, 9, 0.5
    g, _ = O.update_canvas([O.Mesh(v, t)], fr, gbuf=g0)
    miss = g["object_id"] == 0xFFFFFFFF
    assert miss.sum() > 0
    assert np.all(g["depth"][miss] == F32MAX) and np.all(g["db_id"][miss] == 0)
    # j3d/canvas.cpp:859-866 leaves mark, r, g, b and the barycentrics alone on a miss
    assert np.all(g["mark"][miss] == 7) and np.all(g["r"][miss] == 9) and np.all(g["barycentric_u"][miss] == 0.5)
    assert np.all(g["mark"][~miss] == 0) and np.all(g["r"][~miss] == 9)


def test_shadow_ray_occluder():
    """A small plate above a large plate shadows it toward an oblique directional
    light; the shadow is the plate's outline translated along the light."""
    v1, t1 = _plane(z=0.0, size=2.0)
    v2, t2 = _plane(z=1.0, size=0.5)
    v = np.vstack([v1, v2])
    t = np.vstack([t1, t2 + 4]).astype(np.uint32)
    w, h = 80, 60
    _, Pi = O.projection(O.default_camera(), w, h)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = 8.0
    L = np.array([0.6, 0.0, 0.8], np.float32)
    fr = _Frame(w, h, cs, (0, 0, 0), 4.0, O.EDGES | O.SHADOW | O.DIRECTIONAL_LIGHT, Pi, light=L)
    g, st = O.update_canvas([O.Mesh(v, t)], fr)
    assert st["shadowed"] > 0
    lower = g["object_id"] < 2
    assert np.all(g["mark"][~lower] == 0)
    ys, xs = np.nonzero(lower)
    for y, x in zip(ys, xs):
        sx = 2.0 * ((x + 0.5) / w) - 1.0
        sy = 2.0 * ((y + 0.5) / h) - 1.0
        d = np.array([Pi[0] * sx, Pi[5] * sy, -1.0])
        p = np.array([0, 0, 8.0]) + g["depth"][y, x] * d
        q = p + L / L[2]  # where the shadow ray crosses z = 1
        m = max(abs(q[0]), abs(q[1]))
        if m < 0.5 - 1e-3:
            assert g["mark"][y, x] == 1
        elif m > 0.5 + 1e-3:
            assert g["mark"][y, x] == 0


def test_color_byte_conversion():
    # (uint8_t)(c * 255.f): truncation toward zero, low 8 bits, NaN / huge -> 0
    cases = {0.0: 0, 0.5: 127, 1.0: 255, 0.999: 254, 1.2: 50, -0.01: 254, -1.0: 1, float("nan"): 0, 1e10: 0}
    for c, b in cases.items():
        assert O.color_byte(c) == b, (c, O.color_byte(c))


def _color_frame(flags, w=40, h=30):
    _, Pi = O.projection(O.default_camera(), w, h)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = 3.0
    return _Frame(w, h, cs, (0, 0, 0), 20.0, flags, Pi)


def test_vertex_colors_constant_and_blend():
    """canvas.cpp:822-834.  A constant colour 0.5 at every vertex gives byte
    127 everywhere (0.5 scales the barycentric sum exactly, and s0 + u + v is
    1 within a few ulps); pure red, green and blue vertices give channels that
    add up to 255 within rounding, each largest at its own vertex."""
    v, t = _plane()
    fr = _color_frame(O.VERTEXCOLORS | O.EDGES)
    g, _ = O.update_canvas([O.Mesh(v, t, vertex_colors=np.full((4, 3), 0.5, np.float32))], fr)
    assert np.all(g["mark"] == 2) and np.all(g["r"] == 127) and np.all(g["g"] == 127) and np.all(g["b"] == 127)
    vc = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 1, 0]], np.float32)
    g0 = np.zeros((30, 40), O.PIXEL_DTYPE)
    g0["object_id"] = 0xFFFFFFFF
    g, _ = O.update_canvas([O.Mesh(v, t, vertex_colors=vc)], fr, gbuf=g0)
    tot = g["r"].astype(int) + g["g"].astype(int) + g["b"].astype(int)
    assert np.all(tot >= 252) and np.all(tot <= 255)
    # the camera looks down -z at the centre of the plane: vertex 0 (-x, -y) shows in the bottom-left
    # quadrant (rows grow downward), vertex 2 (+x, +y) top-right
    assert g["r"][-1, 0] > g["b"][-1, 0] and g["b"][0, -1] > g["r"][0, -1]
    # flags off: no colour, r/g/b untouched
    fr2 = _color_frame(O.EDGES)
    g1 = np.zeros((30, 40), O.PIXEL_DTYPE)
    g1["r"] = 9
    g1, _ = O.update_canvas([O.Mesh(v, t, vertex_colors=vc)], fr2, gbuf=g1)
    assert np.all(g1["mark"] == 0) and np.all(g1["r"] == 9)


def test_texture_lookup_convention_and_priority():
    """canvas.cpp:803-821.  Every corner of both triangles at one uv point
    selects one texel (x = int(u w), y = int(v h), texel = texture[y][x]);
    uv outside [0, 1] clamps to the border texel; with both sources and both
    flags the texture wins."""
    v, t = _plane()
    tex = np.array([[0xff0000aa, 0xff00bb00], [0xffcc0000, 0xff112233]], np.uint32)  # rows y = 0, 1
    exp = {(0.25, 0.25): (0xaa, 0, 0), (0.75, 0.25): (0, 0xbb, 0), (0.25, 0.75): (0, 0, 0xcc),
           (0.75, 0.75): (0x33, 0x22, 0x11), (5.0, -3.0): (0, 0xbb, 0), (-1.0, 9.0): (0, 0, 0xcc)}
    fr = _color_frame(O.TEXTURED | O.VERTEXCOLORS | O.EDGES)
    for (cu, cv), rgb in exp.items():
        uv = np.tile(np.array([cu, cv], np.float32), (2, 3))
        m = O.Mesh(v, t, uv=uv, texture=tex, vertex_colors=np.full((4, 3), 0.5, np.float32))
        g, _ = O.update_canvas([m], fr)
        assert np.all(g["mark"] == 2)
        assert np.all(g["r"] == rgb[0]) and np.all(g["g"] == rgb[1]) and np.all(g["b"] == rgb[2]), (cu, cv)


def test_colour_and_shadow_marks_combine():
    """mark = 2 | 1 on a coloured, shadowed pixel (canvas.cpp:832, 855)."""
    v1, t1 = _plane(z=0.0, size=2.0)
    v2, t2 = _plane(z=1.0, size=0.5)
    v = np.vstack([v1, v2])
    t = np.vstack([t1, t2 + 4]).astype(np.uint32)
    w, h = 80, 60
    _, Pi = O.projection(O.default_camera(), w, h)
    cs = np.eye(4, dtype=np.float32).reshape(16)
    cs[14] = 8.0
    fr = _Frame(w, h, cs, (0, 0, 0), 4.0, O.EDGES | O.SHADOW | O.DIRECTIONAL_LIGHT | O.VERTEXCOLORS, Pi,
                light=np.array([0.6, 0.0, 0.8], np.float32))
    g, st = O.update_canvas([O.Mesh(v, t, vertex_colors=np.full((8, 3), 0.25, np.float32))], fr)
    hit = g["object_id"] != 0xFFFFFFFF
    assert st["shadowed"] > 0 and set(np.unique(g["mark"][hit])) == {2, 3}
    assert np.sum(g["mark"] == 3) == st["shadowed"]


def test_empty_scene_all_miss():
    w, h = 16, 8
    _, Pi = O.projection(O.default_camera(), w, h)
    fr = _Frame(w, h, np.eye(4, dtype=np.float32).reshape(16), (0, 0, 0), 1.0, O.EDGES, Pi)
    g, st = O.update_canvas([], fr)
    assert st["hits"] == 0 and np.all(g["object_id"] == 0xFFFFFFFF) and np.all(g["depth"] == F32MAX)


# ----------------------------------------------------------- shading ------

def test_background_closed_form():
    h, w = 600, 8
    im = O.background(w, h, 0xff000000, 0xff404040)
    assert np.all(im[0] == 0xff000000)
    for y in range(h):
        exact = int(math.floor(y / h * 0x40 + 1e-9))
        ch = int(im[y, 0] & 0xff)
        assert abs(ch - exact) <= 1 and im[y, 0] >> 24 == 0xff
        assert (im[y, 0] >> 8) & 0xff == ch and (im[y, 0] >> 16) & 0xff == ch
    assert np.all(np.diff((im[:, 0] & 0xff).astype(int)) >= 0)


MATCAP_SHA256 = {
    0: "red wax",
}


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_matcap_closed_form(kind):
    im, cav = O.matcap(kind)
    assert cav == {0: 0xFF7D7DFF, 1: 0xff505050, 2: 0xff405060, 3: 0xFF505050}[kind]
    assert np.all(im >> 24 == 0xff)
    if kind == 3:
        return
    assert im[0, 0] == 0xff000000 and im[511, 511] == 0xff000000  # outside the 1.01 disk
    base = {0: (32, 0, 0), 1: (32, 32, 32), 2: (32, 20, 10)}[kind]
    c1 = {0: (200 / 1.5, 50, 37.5), 1: (50, 50, 50), 2: (50, 45, 40)}[kind]
    c2 = {0: (30, 25, 20), 1: (50, 50, 50), 2: (50, 40, 30)}[kind]
    n1 = np.array([0, 0.8, 1]) / math.sqrt(1.64)
    n2 = np.array([0, 0.4, 1]) / math.sqrt(1.16)
    rng = np.random.default_rng(kind)
    for _ in range(200):
        x, y = rng.integers(0, 512, 2)
        u, v = x / 511 * 2 - 1, y / 511 * 2 - 1
        if u * u + v * v > 1.0:
            continue
        sp = np.array([u, v, math.sqrt(1.01 - u * u - v * v)])
        a = (sp @ n1, (sp @ n2) ** 3, sp[2] ** 5, sp[2] ** 50)
        px = int(im[511 - y, x])  # rows are written bottom-up
        for k in range(3):
            exact = base[k] + c1[k] * a[0] + c2[k] * a[1] + 50 * a[2] + 30 * a[3]
            exact = min(max(exact, 0), 255)
            assert abs(((px >> (8 * k)) & 0xff) - int(exact)) <= 1


def _flat_gbuf(w, h, u=0.0, v=0.0, depth=3.0):
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    g["object_id"], g["u"], g["v"], g["depth"], g["db_id"] = 1, u, v, depth, 1 << 29
    return g


def _frame_for(w, h, flags):
    _, Pi = O.projection(O.default_camera(), w, h)
    return _Frame(w, h, np.eye(4, dtype=np.float32).reshape(16), (0, 0, 0), 1.0, flags, Pi)


def test_edge_pass_flat_region_is_plain_texel():
    w, h = 12, 9
    mc, cav = O.matcap(0)
    fr = _frame_for(w, h, O.EDGES | O.SHADING)
    im = O.canvas_to_image(_flat_gbuf(w, h), fr, mc, cav)
    # centre normal (0, 0) -> texel (U, V) = (256, 256)
    assert np.all(im == mc[256, 256])


def test_edge_pass_misses_keep_background_and_shadow_darkens():
    w, h = 12, 9
    mc, cav = O.matcap(0)
    g = _flat_gbuf(w, h)
    g["object_id"][:, :4] = 0xFFFFFFFF
    g["mark"][:, 8:] = 1
    fr = _frame_for(w, h, O.EDGES | O.SHADING)
    im = O.canvas_to_image(g, fr, mc, cav)
    bg = O.background(w, h)
    assert np.array_equal(im[:, :4], bg[:, :4])
    t = int(mc[256, 256])
    dark = 0xff000000 | ((((t >> 16) & 0xff) >> 2) << 16) | ((((t >> 8) & 0xff) >> 2) << 8) | ((t & 0xff) >> 2)
    assert np.all(im[:, 8:] == dark)


def _expected_edge(px, pr, Pi, w, h, mc, cav, x, y):
    """get_angle_color(compute_convex_cos_angle(...)) of j3d/canvas.cpp:287-389 in float64."""
    def ray(xx, yy):
        sx = 2.0 * ((xx + 0.5) / w) - 1.0
        sy = 2.0 * ((yy + 0.5) / h) - 1.0
        return np.array([Pi[0] * sx, Pi[5] * sy, -1.0])
    p1 = px["depth"] * ray(x, y)
    p2 = pr["depth"] * ray(x + 1, y)
    n1 = np.array([px["u"], px["v"], math.sqrt(1 - px["u"] ** 2 - px["v"] ** 2)])
    n2 = np.array([pr["u"], pr["v"], math.sqrt(1 - pr["u"] ** 2 - pr["v"] ** 2)])
    d = (p2 - p1) / np.linalg.norm(p2 - p1)
    angle = float(d @ n1) if abs(n1 @ n2 - 1) > 1e-4 else 0.0
    U = int(math.floor(0.5 + (px["u"] + 1) * 511 * 0.5))
    V = int(math.floor(0.5 + (1 - px["v"]) * 511 * 0.5))
    tex = int(mc[V, U])
    hp = 1.57079632679489
    scale = math.sqrt(1 - (hp - abs(math.acos(angle) - hp)) / hp)
    k = 1.2 if angle > 0 else 0.5
    out = []
    for sh in (0, 8, 16):
        c = (tex >> sh) & 0xff
        c2 = (cav >> sh) & 0xff
        out.append(min(255, int(c * (1 - scale) + c2 * scale * k)))
    return angle, out


@pytest.mark.parametrize("step", [+0.4, -0.4])
def test_edge_pass_concave_convex_and_last_column(step):
    """A depth step between columns 2 and 3: the edge pixel blends toward the
    cavity colour, x1.2 when the neighbour is nearer (concave, angle > 0) and
    x0.5 when it is farther (convex); the last column is always plain."""
    w, h = 6, 4
    mc, cav = O.matcap(0)
    g = _flat_gbuf(w, h, u=0.1, v=-0.2)
    g["depth"][:, 3:] = 3.0 + step
    fr = _frame_for(w, h, O.EDGES | O.SHADING)
    g["u"][:, 3:] = 0.25  # a normal change is what marks an edge
    im = O.canvas_to_image(g, fr, mc, cav)
    angle, exp = _expected_edge(g[0, 2], g[0, 3], fr.proj_inv, w, h, mc, cav, 2, 0)
    assert (angle > 0) == (step < 0) and angle != 0
    got = [(int(im[0, 2]) >> sh) & 0xff for sh in (0, 8, 16)]
    assert max(abs(a - b) for a, b in zip(got, exp)) <= 1
    U = int(math.floor(0.5 + (np.float32(0.1) + 1) * 511 * 0.5))
    V = int(math.floor(0.5 + (1 - np.float32(-0.2)) * 511 * 0.5))
    assert np.all(im[:, 0:2] == mc[V, U])                 # no edge inside the flat region
    assert np.all(im[:, 5] == im[0, 5]) and im[0, 5] != im[0, 2]


def test_edge_pass_near_parallel_normals_give_zero_angle():
    """|n1.n2 - 1| <= 1e-4 (j3d/canvas.cpp:306) sets angle = 0: no visible blend."""
    w, h = 6, 4
    mc, cav = O.matcap(0)
    g = _flat_gbuf(w, h, u=0.1, v=-0.2)
    g["depth"][:, 3:] = 3.4
    g["u"][:, 3:] = 0.11
    im = O.canvas_to_image(g, _frame_for(w, h, O.EDGES | O.SHADING), mc, cav)
    d = np.abs(np.array([im[0, 2]]).view(np.uint8).astype(int) - np.array([im[0, 0]]).view(np.uint8).astype(int))
    assert d.max() <= 1


def test_edge_pass_nan_normal_falls_back_to_unblended():
    w, h = 4, 2
    mc, cav = O.matcap(0)
    g = _flat_gbuf(w, h, u=0.9, v=0.9)  # 1 - u^2 - v^2 < 0: NaN z
    g["u"][:, 1:] = 0.0
    g["v"][:, 1:] = 0.0
    fr = _frame_for(w, h, O.EDGES | O.SHADING)
    im = O.canvas_to_image(g, fr, mc, cav)
    U = int(math.floor(0.5 + (np.float32(0.9) + 1) * 511 * 0.5))
    V = int(math.floor(0.5 + (1 - np.float32(0.9)) * 511 * 0.5))
    # NaN makes |dot - 1| > 1e-4 false, angle = 0: the texel survives within truncation
    d = np.abs(np.array([im[0, 0]]).view(np.uint8).astype(int) - np.array([mc[V, U]]).view(np.uint8).astype(int))
    assert d.max() <= 1


# ---- all hits (find_all_triangles) and the voxel grid (write_vox) ----------

def _cube_mesh():
    """The unit cube [0,1]^3 as 12 triangles."""
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    q = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    t = np.array([[a, b, c] for a, b, c, d in q] + [[a, c, d] for a, b, c, d in q], np.uint32)
    return v, t


def test_all_hits_bvh_equals_brute_force_and_is_sorted():
    import j3d_amd as J
    from scenes import random_rays
    v, t = J.icosphere(3, seed=2, amplitude=0.2)
    rays = random_rays(3000, 8, np.zeros(3), 1.2)
    m = O.Mesh(v, t)
    a = O.all_hits(m, rays, 16)
    b = O.all_hits(m, rays, 16, brute=True)
    for k in ("count", "tri", "t", "u", "v"):
        assert np.array_equal(a[k], b[k]), k
    assert a["count"].max() >= 2
    for i in np.nonzero(a["count"] >= 2)[0][:200]:
        n = min(int(a["count"][i]), 16)
        pairs = list(zip(a["t"][i, :n], a["tri"][i, :n]))
        assert pairs == sorted(pairs)
    # the first hit is the closest hit; a ray hits something iff the any-hit query says so
    ref = O.closest_hits([m], rays)
    hit = a["count"] > 0
    assert np.array_equal(hit, O.any_hits([m], rays).astype(bool))
    assert np.array_equal(a["tri"][hit, 0], ref["tri"][hit])
    assert np.array_equal(a["t"][hit, 0].view(np.uint32), ref["t"][hit].view(np.uint32))
    # truncation keeps the first hits of the order
    c = O.all_hits(m, rays, 1)
    assert np.array_equal(c["count"], a["count"]) and np.array_equal(c["tri"][:, 0], a["tri"][:, 0])


def test_vox_color_index_known_values():
    # vox.cpp:154-171 by hand: r, g minus 16 and b minus 32 (floored at 0), then r>>5 <<5 | g>>5 <<2 | b>>6
    assert O.vox_color_index(0, 0, 0) == 1  # index 0 is "empty": black maps to 1
    assert O.vox_color_index(255, 255, 255) == (7 << 5) | (7 << 2) | 3
    assert O.vox_color_index(255, 0, 0) == 7 << 5
    assert O.vox_color_index(48, 48, 96) == (1 << 5) | (1 << 2) | 1
    assert O.vox_color_index(47, 47, 95) == 1


def test_vox_dims():
    assert O.vox_dims([0, 0, 0], [2, 1, 0.5], 100) == (100, 50, 25)
    assert O.vox_dims([0, 0, 0], [1, 3, 0], 10) == (3, 10, 1)  # a flat axis gets one cell
    assert O.vox_dims([1, 1, 1], [1, 1, 1], 10) == (1, 1, 1)   # a point: 0/0 ratios count as 0


def test_voxelize_cube_shell():
    """Axis rays through the cell centres of the unit cube hit its faces only:
    the voxels set are exactly the 8^3 grid's boundary shell, white (255)."""
    v, t = _cube_mesh()
    dims, data = O.voxelize(O.Mesh(v, t), 8)
    assert dims == (8, 8, 8)
    shell = np.ones((8, 8, 8), bool)
    shell[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(data > 0, shell)
    assert set(np.unique(data[shell]).tolist()) == {255}
    red = np.tile(np.array([[1.0, 0.0, 0.0]], np.float32), (len(v), 1))
    _, data_r = O.voxelize(O.Mesh(v, t, vertex_colors=red), 8)
    assert set(np.unique(data_r[shell]).tolist()) == {7 << 5}
    # a texture wins over vertex colours: every texel blue (0xAABBGGRR)
    tex = np.full((4, 4), 0xFFFF0000, np.uint32)
    uv = np.tile(np.array([[0, 0, 1, 0, 0, 1]], np.float32), (len(t), 1))
    _, data_t = O.voxelize(O.Mesh(v, t, vertex_colors=red, uv=uv, texture=tex), 8)
    assert set(np.unique(data_t[shell]).tolist()) == {3}
