"""Runs frames and ray queries through whichever build of libj3dg the
J3DG_LIBRARY environment variable names (the CPU emulation in
tests/test_emu_batch.py) with the tail mode EMU_TAIL (the coop_tail tuning:
1 = a group's lanes split each node's children, 2 = they traverse parts of
the ray's tree with stacks of their own) and the LDS stack depth EMU_STACK,
and compares everything with the oracle bit for bit: frames of one object
and of three posed objects (a ray's traversal moves on from object to
object), with shadows, and closest-hit and any-hit queries over the three.
Run with EMU_CUS=16 so that the persistent grid has many waves, each of which
ends in its tail.  Prints one line per check; exit status 0 when everything
matches, 1 on a mismatch (a memory fault in the emulation exits 86)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import j3d_amd as J  # noqa: E402
import oracle as O  # noqa: E402
from j3d_amd import ffi  # noqa: E402
from scenes import compare_gbuffers, compare_rgba, random_rays  # noqa: E402


def posed(angle, axis, offset):
    """Object -> world: a rotation about a coordinate axis, then a translation (column-major 4x4)."""
    c, s = np.cos(angle), np.sin(angle)
    m = np.eye(4, dtype=np.float32)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    m[:3, 3] = offset
    return m.T.reshape(16).copy()


def world(v, cs):
    m = np.asarray(cs, np.float32).reshape(4, 4).T
    return v @ m[:3, :3].T + m[:3, 3]


def main():
    tail = int(os.environ.get("EMU_TAIL", "2"))
    depth = int(os.environ.get("EMU_STACK", "16"))
    r = J.Renderer(0)
    r.set_tuning(**dict(r.get_tuning(), coop_tail=tail, stack_depth=depth))
    print("tuning", r.get_tuning(), flush=True)
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    parts = [(J.torus(300, 150, seed=4), posed(0.0, 0, (0.0, 0.0, 0.0))),
             (J.torus(150, 80, R=0.6, r=0.2, seed=5), posed(1.1, 0, (0.9, 0.3, 0.2))),
             (J.icosphere(3, seed=6), posed(0.4, 1, (-0.8, -0.5, 0.4)))]
    slots, meshes, wv = [], [], []
    for k, ((v, t), cs) in enumerate(parts):
        db = J.make_db_id(4, k)
        slots.append(r.upload_mesh(v, t, db_id=db, cs=cs))
        meshes.append(O.Mesh(v, t, cs=cs, db_id=db))
        wv.append(world(v, cs))
    bad = 0
    cam = J.camera_with_vertical_fov(50.0)
    cases = [("one object", [0], (128, 96), ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, 0.35),
             ("three objects", [0, 1, 2], (112, 80), ffi.SHADING | ffi.SHADOW, 1.9),
             ("three objects, grazing", [0, 1, 2], (96, 64), ffi.EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, 3.0)]
    light = np.array([0.3, 0.5, 0.81], np.float32)
    light /= np.linalg.norm(light)
    for name, objs, (w, h), flags, az in cases:
        pts = np.concatenate([wv[i] for i in objs])
        cs, ci = J.orbit_camera(2.6, az, 0.15 if "grazing" in name else 0.6, target=pts.mean(axis=0))
        diag, pivot = J.scene_framing(pts)
        fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slots[i] for i in objs], flags=flags, light_dir=light)
        tg = r.create_target(w, h)
        g, im = r.render(fr, tg)
        og, oim, _ = O.render([meshes[i] for i in objs], fr, mc, cav)
        mm = compare_gbuffers(g, og)
        d = compare_rgba(im, oim)[0]
        hits = int((g["object_id"] != 0xFFFFFFFF).sum()) if "object_id" in g.dtype.names else -1
        print("frame %s: %d hit pixels, G-buffer mismatches %s, rgba %d" % (name, hits, mm, d), flush=True)
        if sum(mm.values()) or d > 1:
            bad += 1
        r.destroy_target(tg)
    rays = random_rays(20000, 23, np.zeros(3), 1.6)
    got = r.trace_rays(slots, rays)
    ref = O.closest_hits(meshes, rays)
    mism = sum(int((got[k] != ref[k]).sum()) for k in ("obj", "tri"))
    mism += sum(int((got[k].view(np.uint32) != ref[k].view(np.uint32)).sum()) for k in ("t", "u", "v"))
    n_hit = int((ref["tri"] != 0xFFFFFFFF).sum())
    print("closest-hit queries: %d rays, %d hits, %d field mismatches" % (len(rays), n_hit, mism), flush=True)
    bad += mism != 0
    ah = r.any_hit(slots, rays)
    oah = O.any_hits(meshes, rays)
    print("any-hit queries: %d hits, %d mismatches" % (int(oah.sum()), int((ah != oah).sum())), flush=True)
    bad += int((ah != oah).sum()) != 0
    r.close()
    lib = ffi.load()
    if hasattr(lib, "emu_inactive_lane_reads"):
        import ctypes as C
        lib.emu_inactive_lane_reads.restype = C.c_ulonglong
        n = int(lib.emu_inactive_lane_reads())
        print("shuffles that read a lane outside the exec mask: %d" % n)
        bad += n
    print("TAIL OK" if bad == 0 else "TAIL MISMATCH %d" % bad, flush=True)
    sys.exit(0 if bad == 0 else 1)


if __name__ == "__main__":
    main()
