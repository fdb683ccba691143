"""The C++ canvas/scene/camera layer (j3d_amd/host) over the C ABI: CPU checks
of the build and the harness's argument handling, and GPU parity of a frame
driven through it the way j3d's view drives its canvas (view.cpp:421-430)
against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import compare_gbuffers, compare_rgba

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "j3d_amd", "lib")
EXE = os.path.join(LIB, "j3d_render_frame")


def test_canvas_layer_built_and_linked():
    assert os.path.exists(os.path.join(LIB, "libj3d_canvas.so")) and os.path.exists(EXE)
    out = subprocess.run(["ldd", EXE], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert "libj3d_canvas.so" in out and "libj3dg.so" in out and "not found" not in out


@pytest.mark.parametrize("args,msg", [([], "need --scene"), (["--scene", "cube"], "icosphere, torus or ply"), (["--scene", "ply", "/nonexistent.ply", "--size", "8", "8"], "read_ply: cannot open"),
                                      (["--bogus"], "unknown argument")])
def test_harness_rejects_bad_arguments(args, msg):
    r = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 2 and msg in r.stderr


def _oracle_frame(v, t, w, h, cam, cs, ci, flags, light, rect=None, mc_kind=0):
    diag, pivot = J.scene_framing(v)
    fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [0], flags=flags, light_dir=light, rect=rect)
    mc, cav = O.matcap(mc_kind)
    m = O.Mesh(v, t, db_id=1 << 29)
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    g["object_id"] = 0xFFFFFFFF
    g, _ = O.update_canvas([m], fr, gbuf=g)
    im = O.canvas_to_image(g, fr, mc, cav)
    return g, im


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["full", "rect", "torus_shadow", "points"])
def test_canvas_frame_matches_oracle(tmp_path, case):
    w, h = 320, 240
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    light = np.array([0.3, 0.5, 0.81], np.float32)
    light /= np.linalg.norm(light)
    rect = None
    if case in ("torus_shadow", "points"):
        v, t = J.torus(300, 170, seed=2)
        cs, ci = J.orbit_camera(3.0, 0.7, 0.5)
        flags = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
        scene_args = ["--scene", "torus", "300", "170", "2"]
        flag_arg = "edges,shading,shadow,directional"
    else:
        v, t = J.icosphere(5, seed=3, amplitude=0.05)
        flags = ffi.EDGES | ffi.SHADING
        scene_args = ["--scene", "icosphere", "5", "3", "0.05"]
        flag_arg = "edges,shading"
        if case == "rect":
            rect = (40, 30, 250, 200)
    prefix = str(tmp_path / case)
    cmd = [EXE] + scene_args + ["--size", str(w), str(h), "--focal", repr(float(cam.focal_length)),
                                "--cs"] + [repr(float(x)) for x in cs] + ["--flags", flag_arg, "--out", prefix]
    if flags & ffi.DIRECTIONAL_LIGHT:
        cmd += ["--light"] + [repr(float(x)) for x in light]
    if rect:
        cmd += ["--rect"] + [str(x) for x in rect]
    if case == "points":  # render_pointclouds_on_image after canvas_to_image (view.cpp:428)
        cmd += ["--points", "60000", "5"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = np.fromfile(prefix + ".gbuf", ffi.PIXEL_DTYPE).reshape(h, w)
    im = np.fromfile(prefix + ".rgba", np.uint32).reshape(h, w)
    og, oim = _oracle_frame(v, t, w, h, cam, cs, ci, flags, light, rect=rect)
    if case == "points":
        rec = np.fromfile(prefix + ".points", np.float32).reshape(2, -1, 3)
        col = np.fromfile(prefix + ".colors", np.uint32)
        P, _ = J.projection(cam, w, h)
        # the harness's scene CS^-1 is invert_orthonormal(--cs), which can differ
        # from orbit_camera's inverse in the translation's last bit; points see it
        ci_scene = J.invert_orthonormal(cs)
        og, oim = O.splat_points(og, oim, P, ci_scene, flags, [{"xyz": rec[0], "normals": rec[1], "colors": col,
                                                           "db_id": (2 << 29) + 1}])
        assert (g["db_id"] == (2 << 29) + 1).sum() > 1000
    assert all(c == 0 for c in compare_gbuffers(g, og).values()), compare_gbuffers(g, og)
    assert compare_rgba(im, oim)[0] <= 1


@pytest.mark.gpu
@pytest.mark.parametrize("device_build", [False, True])
def test_canvas_many_objects_matches_oracle(tmp_path, device_build):
    """A scene of 20 objects through the canvas layer: more than the frame
    description's objects[] holds, so the canvas passes object_list and the
    frame is traversed through the top-level BVH; with --device-build the
    objects' trees are built on the GPU (add_object's builder)."""
    w, h, n, spacing = 320, 160, 20, 0.9
    v, t = J.icosphere(3, seed=6, amplitude=0.05)
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.orbit_camera(11.0, 1.5, 0.3, target=(n * spacing * 0.5, 0.0, 0.0))
    flags = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    light = np.array([0.3, 0.5, 0.81], np.float32)
    light /= np.linalg.norm(light)
    prefix = str(tmp_path / "many")
    cmd = [EXE, "--scene", "icosphere", "3", "6", "0.05", "--copies", str(n), repr(spacing),
           "--size", str(w), str(h), "--focal", repr(float(cam.focal_length)),
           "--cs"] + [repr(float(x)) for x in cs] + ["--flags", "edges,shading,shadow,directional", "--light"] + \
        [repr(float(x)) for x in light] + ["--out", prefix] + (["--device-build"] if device_build else [])
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = np.fromfile(prefix + ".gbuf", ffi.PIXEL_DTYPE).reshape(h, w)
    im = np.fromfile(prefix + ".rgba", np.uint32).reshape(h, w)
    # the oracle scene: object k translated by k * spacing (float32, as the harness computes it)
    lo, hi = J.bbox(v)
    meshes, blo, bhi = [], None, None
    for k in range(n):
        ocs = np.eye(4, dtype=np.float32)
        ocs[3, 0] = np.float32(k) * np.float32(spacing)
        ocs = ocs.reshape(16)
        meshes.append(O.Mesh(v, t, db_id=(1 << 29) + k, cs=ocs))
        a = (lo + np.array([ocs[12], 0, 0], np.float32)).astype(np.float32)
        b = (hi + np.array([ocs[12], 0, 0], np.float32)).astype(np.float32)
        blo = a if blo is None else np.minimum(blo, a)
        bhi = b if bhi is None else np.maximum(bhi, b)
    ext = bhi - blo  # prepare_scene (j3d/scene.cpp:50-89) over every object
    diag = float(np.float32(max(ext[0], ext[1], ext[2])))
    pivot = ((blo + bhi) * np.float32(0.5)).astype(np.float32)
    fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, list(range(n)), flags=flags, light_dir=light)
    mc, cav = O.matcap(0)
    og = np.zeros((h, w), O.PIXEL_DTYPE)
    og["object_id"] = 0xFFFFFFFF
    og, _ = O.update_canvas(meshes, fr, gbuf=og)
    oim = O.canvas_to_image(og, fr, mc, cav)
    mm = compare_gbuffers(g, og)
    assert all(c == 0 for c in mm.values()), mm
    assert compare_rgba(im, oim)[0] <= 1
    hit = g["object_id"] != 0xFFFFFFFF
    assert len(np.unique(g["db_id"][hit])) > 16


@pytest.mark.gpu
def test_canvas_points_in_scene_framing(tmp_path):
    """prepare_scene + unzoom through the canvas layer (no --cs) for a scene of
    one mesh and one point cloud (j3d/scene.cpp:50-110): the cloud's box, which
    the jitter pushes past the mesh's, enters the diagonal and the pivot, and so
    the camera, t_near and the frame.  The frame and the points drawn over it
    equal the oracle's with the framing it computes over both boxes."""
    w, h = 256, 192
    cam = J.camera_with_vertical_fov(60.0)
    v, t = J.torus(300, 170, seed=2)
    prefix = str(tmp_path / "pts")
    cmd = [EXE, "--scene", "torus", "300", "170", "2", "--size", str(w), str(h), "--focal",
           repr(float(cam.focal_length)), "--flags", "edges,shading", "--points", "60000", "5", "--out", prefix]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import json
    rep = json.loads(r.stdout.strip().splitlines()[-1])
    rec = np.fromfile(prefix + ".points", np.float32).reshape(2, -1, 3)
    col = np.fromfile(prefix + ".colors", np.uint32)
    ident = np.eye(4, dtype=np.float32).reshape(16)
    mlo, mhi = J.bbox(v)
    plo, phi = J.bbox(rec[0])
    diag, pivot, cs, ci = O.prepare_and_unzoom([(ident, mlo, mhi), (ident, plo, phi)])
    diag_mesh, _, _, _ = O.prepare_and_unzoom([(ident, mlo, mhi)])
    assert diag != diag_mesh  # the cloud changes the framing
    assert np.float32(rep["diagonal"]) == np.float32(diag)
    assert np.array_equal(np.array(rep["pivot"], np.float32), pivot)
    flags = ffi.EDGES | ffi.SHADING
    fr = J.Frame(w, h, J.projection(cam, w, h)[1], cs, ci, pivot, diag, [0], flags=flags,
                 near_plane=cam.near_plane)
    g = np.fromfile(prefix + ".gbuf", ffi.PIXEL_DTYPE).reshape(h, w)
    im = np.fromfile(prefix + ".rgba", np.uint32).reshape(h, w)
    mc, cav = O.matcap(0)
    og = np.zeros((h, w), O.PIXEL_DTYPE)
    og["object_id"] = 0xFFFFFFFF
    og, _ = O.update_canvas([O.Mesh(v, t, db_id=1 << 29)], fr, gbuf=og)
    oim = O.canvas_to_image(og, fr, mc, cav)
    P, _ = J.projection(cam, w, h)
    og, oim = O.splat_points(og, oim, P, ci, flags, [{"xyz": rec[0], "normals": rec[1], "colors": col,
                                                      "db_id": (2 << 29) + 1}])
    assert (g["db_id"] == (2 << 29) + 1).sum() > 500
    assert all(c == 0 for c in compare_gbuffers(g, og).values()), compare_gbuffers(g, og)
    assert compare_rgba(im, oim)[0] <= 1


@pytest.mark.gpu
@pytest.mark.parametrize("ascii", [False, True])
def test_canvas_ply_scene_matches_oracle(tmp_path, ascii):
    """A mesh loaded from a PLY file (read_ply, j3d/io.cpp:733-736) through the
    canvas layer, framed by prepare_scene + unzoom: the frame equals the
    oracle's over the arrays written to the file."""
    from j3d_amd import meshio
    w, h = 256, 192
    cam = J.camera_with_vertical_fov(60.0)
    v, t = J.torus(240, 130, seed=8)
    path = str(tmp_path / ("mesh_a.ply" if ascii else "mesh_b.ply"))
    rng = np.random.default_rng(3)
    meshio.write_ply(path, v, t, normals=rng.normal(size=v.shape), colors=np.arange(len(v)), ascii=ascii)
    prefix = str(tmp_path / "ply")
    cmd = [EXE, "--scene", "ply", path, "--size", str(w), str(h), "--focal", repr(float(cam.focal_length)),
           "--flags", "edges,shading", "--out", prefix]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = np.fromfile(prefix + ".gbuf", ffi.PIXEL_DTYPE).reshape(h, w)
    im = np.fromfile(prefix + ".rgba", np.uint32).reshape(h, w)
    ident = np.eye(4, dtype=np.float32).reshape(16)
    lo, hi = J.bbox(v)
    diag, pivot, cs, ci = O.prepare_and_unzoom([(ident, lo, hi)])
    fr = J.Frame(w, h, J.projection(cam, w, h)[1], cs, ci, pivot, diag, [0], flags=ffi.EDGES | ffi.SHADING,
                 near_plane=cam.near_plane)
    mc, cav = O.matcap(0)
    og = np.zeros((h, w), O.PIXEL_DTYPE)
    og["object_id"] = 0xFFFFFFFF
    og, _ = O.update_canvas([O.Mesh(v, t, db_id=1 << 29)], fr, gbuf=og)
    oim = O.canvas_to_image(og, fr, mc, cav)
    assert (g["object_id"] != 0xFFFFFFFF).sum() > 2000
    assert all(c == 0 for c in compare_gbuffers(g, og).values()), compare_gbuffers(g, og)
    assert compare_rgba(im, oim)[0] <= 1
