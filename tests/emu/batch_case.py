"""Runs the batch scenario of tests/test_gpu_batch.py through whichever build
of libj3dg the J3DG_LIBRARY environment variable names (the CPU emulation in
tests/test_emu_batch.py): frames of meshes of uneven size, several sizes,
sub-rectangles and modes, rendered one at a time and then as one batch (graph
capture and replays, then updated frame tables), compared bit for bit, and
against the oracle; EMU_TAIL picks the tail mode (the coop_tail tuning).  Prints one line per check; exit status 0 when every
frame matches, 1 on a mismatch (a memory fault in the emulation exits 86)."""
import math
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
from scenes import compare_gbuffers, compare_rgba  # noqa: E402

SIZES = [(64, 48), (40, 30), (56, 56), (72, 40), (37, 23)]
FLAG_SETS = [ffi.EDGES | ffi.SHADING, ffi.EDGES | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, ffi.SHADING | ffi.SHADOW,
             ffi.WIREFRAME | ffi.SHADING, ffi.ONE_BIT | ffi.EDGES]


def meshes(r):
    out = []
    for k, (nu, nv) in enumerate([(20, 10), (60, 30), (150, 75), (300, 120)]):
        v, t = J.torus(nu, nv, seed=10 + k)
        out.append((r.upload_mesh(v, t, db_id=J.make_db_id(2, k)), v, t, J.make_db_id(2, k)))
    v, t = J.icosphere(3, seed=21)
    out.append((r.upload_mesh(v, t, db_id=J.make_db_id(3, 0)), v, t, J.make_db_id(3, 0)))
    return out


def make_frames(ms, n, seed):
    rng = np.random.default_rng(seed)
    cam = J.camera_with_vertical_fov(60.0)
    frames = []
    for i in range(n):
        slot, v, _, _ = ms[i % len(ms)]
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


def main():
    n_frames = int(sys.argv[1]) if len(sys.argv) > 1 else 13
    r = J.Renderer(0)
    if os.environ.get("EMU_TAIL"):  # the tail mode (coop_tail tuning) to run the batches with
        r.set_tuning(**dict(r.get_tuning(), coop_tail=int(os.environ["EMU_TAIL"])))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    ms = meshes(r)
    bad = 0
    for seed in (1, 2):
        frames = make_frames(ms, n_frames, seed)
        expected = []
        for fr in frames:
            tg = r.create_target(fr.width, fr.height)
            expected.append(r.render(fr, tg))
            r.destroy_target(tg)
        targets = [r.create_target(fr.width, fr.height) for fr in frames]
        for rep in range(2):  # the capture, then a replay of the graph
            r.render_batch(frames, targets)
            r.synchronize()
            for i, (tg, (g0, im0)) in enumerate(zip(targets, expected)):
                g, im = r.read_target(tg)
                mm = compare_gbuffers(g, g0)
                d = compare_rgba(im, im0)[0]
                if sum(mm.values()) or d:
                    bad += 1
                    print("seed %d pass %d frame %d differs from its single render: %s rgba %d" % (seed, rep, i, mm, d))
        # every frame against the oracle
        for i, (fr, tg) in enumerate(zip(frames, targets)):
            slot, v, t, db = ms[i % len(ms)]
            g, im = r.read_target(tg)
            og = np.zeros((fr.height, fr.width), O.PIXEL_DTYPE)
            og["object_id"] = 0xFFFFFFFF
            og, _ = O.update_canvas([O.Mesh(v, t, db_id=db)], fr, gbuf=og)
            oim = O.canvas_to_image(og, fr, mc, cav)
            mm = compare_gbuffers(g, og)
            if sum(mm.values()) or compare_rgba(im, oim)[0] > 1:
                bad += 1
                print("seed %d frame %d differs from the oracle: %s" % (seed, i, mm))
        for tg in targets:
            r.destroy_target(tg)
        print("seed %d: %d frames checked" % (seed, n_frames), flush=True)
    r.close()
    lib = ffi.load()
    if hasattr(lib, "emu_inactive_lane_reads"):
        import ctypes as C
        lib.emu_inactive_lane_reads.restype = C.c_ulonglong
        n = int(lib.emu_inactive_lane_reads())
        print("shuffles that read a lane outside the exec mask: %d" % n)
        bad += n
    print("BATCH OK" if bad == 0 else "BATCH MISMATCH %d" % bad, flush=True)
    sys.exit(0 if bad == 0 else 1)


if __name__ == "__main__":
    main()
