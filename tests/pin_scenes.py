"""Scenes and G-buffers whose results follow from j3d/canvas.cpp by hand.

Shared by tests/test_oracle_pins.py (the oracle against float64 geometry and
hand-derived colours) and tests/test_gpu_pins.py (the HIP path against the
oracle on the same inputs).  Matrices are column-major float32 with the
translation in m[12..14], as the frame descriptor and the oracle take them.
"""
import math

import numpy as np

import oracle as O

MISS = 0xFFFFFFFF


def rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def rigid(R, T):
    """Column-major 4x4 of the map x -> R x + T, as float32."""
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = T
    return np.ascontiguousarray(M.T.reshape(16), np.float32)


def mat(m):
    """A float32 column-major matrix as a float64 4x4 (row, column)."""
    return np.asarray(m, np.float64).reshape(4, 4).T


class Frame:
    """The frame fields the oracle reads (a j3d_amd.Frame carries the same ones)."""

    def __init__(self, w, h, cs, pivot, diag, flags, rect=None, light=(0.0, 0.0, 1.0)):
        _, Pi = O.projection(O.default_camera(), w, h)
        self.width, self.height = w, h
        self.rect = rect if rect is not None else (0, 0, w - 1, h - 1)
        self.proj_inv = Pi
        self.near_plane = np.float32(0.1)
        self.cs = cs
        self.cs_inv = O.invert_orthonormal(cs)
        self.pivot = np.asarray(pivot, np.float32)
        self.diagonal = np.float32(diag)
        self.flags = flags
        self.light_dir = np.asarray(light, np.float32)
        self.bg_top, self.bg_bottom = 0xff000000, 0xff404040


def quad(size, z=0.0):
    v = np.array([[-size, -size, z], [size, -size, z], [size, size, z], [-size, size, z]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, t


def pixel_ray(fr, x, y):
    """World origin and direction of pixel (x, y) in float64 (canvas.cpp:773-782)."""
    Pi = np.asarray(fr.proj_inv, np.float64)
    sx = 2.0 * ((x + 0.5) / fr.width) - 1.0
    sy = 2.0 * ((y + 0.5) / fr.height) - 1.0
    d_cam = np.array([Pi[0] * sx, Pi[5] * sy, -1.0, 0.0])
    C = mat(fr.cs)
    return C[:3, 3].copy(), (C @ d_cam)[:3]


# ---- normal transform: a plane under a rotated pose, seen by a rotated camera

def normal_scene(w=16, h=12, flags=None):
    """(verts, tris, object cs, frame): the plane z = 0 (object space) under
    the pose rot_x(0.4) rot_z(0.25) + (0.2, -0.1, 0.3), camera rot_z(0.7)
    rot_y(0.3) at distance 5 along its own z axis."""
    v, t = quad(10.0)
    ocs = rigid(rot("x", 0.4) @ rot("z", 0.25), [0.2, -0.1, 0.3])
    Rc = rot("z", 0.7) @ rot("y", 0.3)
    cs = rigid(Rc, 5.0 * Rc[:, 2])
    fr = Frame(w, h, cs, (0, 0, 0), 2.0, O.EDGES | O.SHADING if flags is None else flags)
    return v, t, ocs, fr


# ---- point light: one pixel's shadow ray and a tiny occluder on it --------

def light_world(fr):
    """CS * (pivot + 3 * diagonal, 1) (canvas.cpp:703-705) in float64."""
    p = np.asarray(fr.pivot, np.float64) + 3.0 * float(fr.diagonal)
    return (mat(fr.cs) @ np.array([p[0], p[1], p[2], 1.0]))[:3]


def occluder(centre, normal, r):
    """A triangle of circumradius r centred on `centre`, in the plane normal to `normal`."""
    nz = normal / np.linalg.norm(normal)
    a = np.cross(nz, [1.0, 0.0, 0.0] if abs(nz[0]) < 0.9 else [0.0, 1.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(nz, a)
    pts = [centre + r * (math.cos(k * 2 * math.pi / 3) * a + math.sin(k * 2 * math.pi / 3) * b) for k in range(3)]
    return np.array(pts, np.float32)


def line_distance(p, o, d):
    d = d / np.linalg.norm(d)
    q = p - o
    return float(np.linalg.norm(q - (q @ d) * d))


SHADOW_PIXEL = (20, 15)


def shadow_scene(where, r, w=48, h=36, full=False):
    """A ground plane (z = 0) seen by a tilted, turned camera, point light
    (SHADOW without DIRECTIONAL_LIGHT), and one tiny occluder of circumradius
    r on the shadow ray of SHADOW_PIXEL at where(P, L), P the pixel's surface
    point and L the light.  The frame covers that pixel only unless `full`.
    Returns (verts, tris, frame, geometry)."""
    Rc = rot("z", 0.5) @ rot("x", 0.45)
    cs = rigid(Rc, Rc @ np.array([0.0, 0.0, 7.0]) + np.array([0.3, -0.2, 0.0]))
    x, y = SHADOW_PIXEL
    fr = Frame(w, h, cs, (0.1, -0.2, 0.05), 5.0, O.SHADOW, rect=None if full else (x, y, x, y))
    o, d = pixel_ray(fr, x, y)
    P = o + (-o[2] / d[2]) * d
    L = light_world(fr)
    c = where(P, L)
    gv, gt = quad(20.0)
    v = np.vstack([gv, occluder(c, L - P, r)]).astype(np.float32)
    t = np.vstack([gt, np.array([[4, 5, 6]], np.uint32)]).astype(np.uint32)
    return v, t, fr, dict(P=P, L=L, c=c, o=o, d=d)


SHADOW_CASES = {
    # name: (where, occluder radius, expected mark bit 0 of SHADOW_PIXEL)
    "halfway": (lambda P, L: P + 0.5 * (L - P), 0.05, 1),
    "beyond": (lambda P, L: P + 2.0 * (L - P), 0.05, 1),
    "near_0.02": (lambda P, L: P + 0.02 * (L - P) / np.linalg.norm(L - P), 0.004, 0),
    "near_0.3": (lambda P, L: P + 0.3 * (L - P) / np.linalg.norm(L - P), 0.004, 1),
}


# ---- hand-made G-buffers for the shading pass ----------------------------

def flat_gbuf(w, h, u=0.1, v=-0.2, depth=3.0):
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    g["object_id"], g["u"], g["v"], g["depth"], g["db_id"] = 1, u, v, depth, 1 << 29
    return g


def row_step_gbuf(step):
    """Rows 0-1 and rows 2-4 of a 6 x 5 G-buffer differ in normal and depth."""
    g = flat_gbuf(6, 5)
    g["u"][2:, :] = 0.25
    g["depth"][2:, :] = 3.0 + step
    return g


def row_zero_gbuf():
    """Row 0 of a 5 x 3 G-buffer differs from rows 1-2."""
    g = flat_gbuf(5, 3, u=0.3, v=0.1, depth=2.5)
    g["u"][1:, :] = -0.2
    g["v"][1:, :] = 0.15
    g["depth"][1:, :] = 2.9
    return g


WIRE_IDS = np.array([[7, 7, 8, 8, 8],
                     [7, 7, 7, 8, 8],
                     [MISS, 7, 7, 7, 7],
                     [9, 9, MISS, 7, 3]], np.uint32)


def wire_gbuf():
    g = flat_gbuf(5, 4, u=0.6, v=0.0)
    g["object_id"] = WIRE_IDS
    g["mark"][3, 0] = 1
    return g


def one_bit_edge_gbuf():
    g = flat_gbuf(6, 4, u=0.0, v=0.0)
    g["u"][:, 3:] = 0.5
    g["u"][3, :] = 0.2
    g["u"][3, 3:] = 0.5
    return g


def const_matcap(c):
    return np.full((512, 512), c, np.uint32)
