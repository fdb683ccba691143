#!/usr/bin/env python3
"""Ray-casting benchmark of j3d's per-frame path on one MI355X.

Workload (default): seeded displaced torus, 5000 x 2800 grid = 14,000,000
vertices / 28,000,000 triangles; orbit camera at distance 3, 60 degree vertical
fov, 1920 x 1080 primary rays plus one directional shadow ray per hit pixel;
edges + matcap shading.  A step is one frame with the mesh and BVH already
resident in HBM: one persistent k_trace launch (ray generation, primary
traversal, hit completion into the G-buffer and the shadow ray, per lane) and
the shading/edge pass (k_shade).  The learned-bounded frame adds the proposal
network (k_infer, bf16 MFMA) before the trace, which then traverses with
tmax = prediction * (1 + margin) and falls back to the unbounded traversal.

value = (primary + shadow rays) / wall time per frame of the learned-bounded
frame (network included), in Mrays/s; plain_mrays is the same frame without
the network.  Both frames' G-buffers must be identical (reported).

The JSON line also carries:
  roofline      the trace kernel's algorithmic bytes per launch (node visits x
                node bytes + triangle tests x 48 B, primary and shadow, + 32 B
                G-buffer write per pixel + the 48 B completion record per hit,
                from the kernel's own per-ray counters) over its mean
                HIP-event duration, against 8 TB/s; `traffic` is the HBM bytes
                per launch measured by rocprofv3 (profiles/), when present.
  cpu_baseline  the CPU oracle renderer (the project's restatement of
                canvas::update_canvas + canvas_to_image with its own BVH and a
                row-parallel std::thread pool) timed on this host's cores on
                full frames of the same workload.

--workload torus300m is the 300M-triangle case: 16 orbit cameras at
3840 x 2160 with shadow rays, the network retrained on that mesh, oracle
parity on a 256 x 256 crop of every camera (see multi_camera).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s @28M-tri 1080p: learned-bounded vs plain GPU traversal; % HBM peak"
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="torus28m", choices=["torus28m", "icosphere327k", "torus300m", "set16x64", "vox28m", "scene256", "splat56m"])
    ap.add_argument("--grid", type=int, nargs=2, default=None, metavar=("NU", "NV"),
                    help="torus300m: torus grid (default 15000 10000 = 300M triangles)")
    ap.add_argument("--resolution", type=int, nargs=2, default=None, metavar=("W", "H"),
                    help="torus300m: frame size (default 3840 2160)")
    ap.add_argument("--cameras", type=int, default=16, help="torus300m: orbit cameras")
    ap.add_argument("--crop", type=int, default=256, help="torus300m: side of the oracle parity crop")
    ap.add_argument("--width", type=int, default=8, choices=[4, 8])
    ap.add_argument("--leaf-size", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="max triangles per BVH leaf (0: the library default, 2 for 8-wide trees)")
    ap.add_argument("--builder", default="host", choices=["host", "device"],
                    help="BVH builder: binned SAH on the host (default) or the linear BVH built on the GPU")
    ap.add_argument("--vox-dim", type=int, default=512, help="vox28m: max_dim of the voxel grid")
    ap.add_argument("--stack-depth", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--coop-tail", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline time")
    ap.add_argument("--profile-only", action="store_true", help="no CPU baseline, short output (for rocprofv3)")
    ap.add_argument("--rows", type=int, default=0, help="render only a centred band of this many rows")
    ap.add_argument("--no-learned", action="store_true", help="plain traversal only")
    ap.add_argument("--train-rays", type=int, default=None, help="default 2^24 (2^25 for torus300m)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--autotune", type=int, default=1, help="autotune stack depth / rays per item / margin")
    ap.add_argument("--orbit-frames", type=int, default=24,
                    help="torus28m: also time an orbiting camera of this many frames per turn (0: off)")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with a 1 GiB device copy (read + write), the reference for profiler byte counters")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(name):
    import j3d_amd as J
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenes import c0_scene, torus_scene
    from j3d_amd import ffi
    if name == "torus28m":
        t0 = time.time()
        v, t, cam, cs, ci, pivot, diag, light = torus_scene(5000, 2800, seed=1)
        log("torus 5000x2800: %d verts %d tris in %.1fs" % (len(v), len(t), time.time() - t0))
        w, h = 1920, 1080
        flags = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
        desc = {"workload": "displaced torus R=1 r=0.35, 5000x2800 grid, 28,000,000 triangles, fBm 6 octaves; "
                            "1920x1080 primary + directional shadow rays, edges + matcap",
                "triangles": int(len(t)), "width": w, "height": h, "shadow": "directional"}
    else:
        v, t, cam, cs, ci, pivot, diag = c0_scene()
        light = np.array([0, 0, 1], np.float32)
        w, h = 1024, 768
        flags = ffi.EDGES | ffi.SHADING
        desc = {"workload": "displaced icosphere level 7, 327,680 triangles; 1024x768 primary rays, edges + matcap",
                "triangles": int(len(t)), "width": w, "height": h, "shadow": "off"}
    return v, t, cam, cs, ci, pivot, diag, light, w, h, flags, desc


def read_traffic(workload_name, fused):
    """HBM bytes per trace launch measured by rocprofv3 --pmc (committed summary)."""
    p = os.path.join(ROOT, "profiles", "k_trace_%s_%s_1080p_traffic.json" % ("fused" if fused else "primary",
                                                                         workload_name))
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def cpu_baseline(v, t, frame, mc, cav, gpu_gbuf, gpu_im, min_seconds):
    """Times the oracle's update_canvas + canvas_to_image (the reference's
    render_scene + canvas_to_image, row-parallel over std::thread) on full
    frames of the same workload, repeated for at least `min_seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    t0 = time.time()
    m = O.Mesh(v, t, db_id=(1 << 29), threads=threads)
    bvh_s = time.time() - t0
    h, w = frame.height, frame.width
    # parity frame (not timed): every G-buffer field against the GPU's, near
    # ties (the two nearest hits within 1e-6 relative depth) counted
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenes import compare_gbuffers
    g = np.zeros((h, w), O.PIXEL_DTYPE)
    g["object_id"] = 0xFFFFFFFF
    g, tst = O.update_canvas([m], frame, gbuf=g, threads=threads, count_ties=True)
    im = O.canvas_to_image(g, frame, mc, cav)
    field_mism = {k: n for k, n in compare_gbuffers(gpu_gbuf, g).items() if n}
    near_ties = int(tst["near_ties"])
    frames, rays, dt = 0, 0, 0.0
    while dt < min_seconds or frames == 0:
        g = np.zeros((h, w), O.PIXEL_DTYPE)
        g["object_id"] = 0xFFFFFFFF
        t0 = time.time()
        g, st = O.update_canvas([m], frame, gbuf=g, threads=threads)
        im = O.canvas_to_image(g, frame, mc, cav)
        dt += time.time() - t0
        rays += st["primary_rays"] + st["shadow_rays"]
        frames += 1
        if frames >= 500:
            break
    mism = int(sum(field_mism.values()))
    d = np.abs(im.view(np.uint8).astype(int) - gpu_im.view(np.uint8).astype(int))
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d full %dx%d frames (%d primary + %d shadow rays each), update_canvas + canvas_to_image, "
                      "%.1fs of CPU time (oracle BVH build %.1fs excluded)" % (frames, w, h, st["primary_rays"],
                                                                             st["shadow_rays"], dt, bvh_s),
            "cpu_model": model, "gbuffer_mismatches_vs_gpu": mism, "gbuffer_field_mismatches": field_mism,
            "near_tie_pixels": near_ties, "rgba_max_diff_vs_gpu": int(d.max()),
            "ms_per_frame": dt / frames * 1e3}


def orbit_sequence(r, tg, cam, pivot, diag, slot, flags, light, w, h, n, passes):
    """A viewer's orbiting camera: n frames whose camera moves 360/n degrees in
    azimuth from one to the next (distance 3, elevation 35 degrees), each
    frame's tile cost order predicted from the frame before it (another
    view).  Counts the rays of every frame, then times `passes` passes over
    the sequence (plain traversal)."""
    import j3d_amd as J
    frames = []
    for k in range(n):
        cs, ci = J.orbit_camera(3.0, 2.0 * math.pi * k / n, math.radians(35.0))
        frames.append(J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light))
    r.set_counting(True)
    rays = 0
    for f in frames:
        r.render(f, tg, readback=False)
        st = r.stats()
        rays += st["primary_rays"] + st["shadow_rays"]
    r.set_counting(False)
    for f in frames:  # warm-up pass
        r.render(f, tg, readback=False)
    r.synchronize()
    r.reset_timings()
    t0 = time.perf_counter()
    for _ in range(passes):
        for f in frames:
            r.render(f, tg, readback=False)
    r.synchronize()
    wall = time.perf_counter() - t0
    tm = r.timings()
    return {"frames": n, "degrees_per_frame": 360.0 / n, "passes": passes, "rays_per_pass": int(rays),
            "mrays": rays * passes / wall / 1e6, "ms_per_frame": wall / (passes * n) * 1e3,
            "trace_ms_mean": float(np.mean(tm["trace_ms"]))}


def bound_ceiling(r, tg, fr, slot, v, t, g, st, steps, warmup):
    """What a perfect learned bound could save, for both ray kinds.

    Primary rays: the plain frame's own hit distances written as the
    proposal (margin 0; misses proposed as misses), rendered with the network
    skipped -- the bounded pass at its best, with no inference cost; counted
    and timed.  Shadow rays (any hit): rebuilt from the G-buffer, traced in
    full (their node visits must equal the frame's), then the occluded ones
    again inside a bound around their nearest occluder (closest-hit query):
    [1e-3, t(1 + 1e-4)] and [t(1 - 1e-4), t(1 + 1e-4)].  An unoccluded ray
    cannot be bounded (a miss must be proven over the whole ray)."""
    import j3d_amd as J
    from j3d_amd import ffi
    from scenes import MISS, shadow_rays
    hit = g["object_id"] != MISS
    pairs = np.stack([np.where(hit, g["depth"], 0), np.where(hit, 1.0, -1.0)], axis=-1).astype(np.float32)
    r.write_proposal(tg, pairs)
    ref_g, _ = r.render(fr, tg)  # plain, over the same target contents as the bounded frame
    pfr = J.Frame(fr.width, fr.height, fr.proj_inv, fr.cs, fr.cs_inv, fr.pivot, fr.diagonal, fr.objects,
                  flags=fr.flags | ffi.LEARNED | ffi.KEEP_PROPOSAL, near_plane=fr.near_plane,
                  light_dir=fr.light_dir, bg_top=fr.bg_top, bg_bottom=fr.bg_bottom, rect=fr.rect, margin=0.0)
    r.set_counting(True)
    pg, _ = r.render(pfr, tg)
    pst = r.stats()
    r.set_counting(False)
    same = gbuffer_mismatches(pg, ref_g) == 0
    for _ in range(warmup):
        r.render(pfr, tg, readback=False)
    r.synchronize()
    r.reset_timings()
    for _ in range(steps):
        r.render(pfr, tg, readback=False)
    r.synchronize()
    perfect_trace_ms = float(np.mean(r.timings()["trace_ms"]))
    out = {"primary": {"node_visits_plain": int(st["node_visits"]), "node_visits_perfect": int(pst["node_visits"]),
                       "fallback": int(pst["learned_fallback"]), "gbuffer_equal_plain": bool(same),
                       "trace_ms_perfect_no_network": perfect_trace_ms}}
    rays, _ = shadow_rays(v, t, g, fr.light_dir)
    r.set_counting(True)
    occ = r.any_hit([slot], rays).astype(bool)
    full = r.stats()
    q = r.trace_rays([slot], rays[occ])
    tocc = q["t"]
    sh = {"rays": int(len(rays)), "occluded": int(occ.sum()), "frame_shadowed": int(st["shadowed"]),
          "node_visits_frame": int(st["shadow_node_visits"]), "node_visits_full": int(full["node_visits"]),
          "node_visits_occluded_full": None}
    occ_rays = rays[occ]
    r.any_hit([slot], occ_rays)
    sh["node_visits_occluded_full"] = int(r.stats()["node_visits"])
    for name, lo, hi in (("upper", None, 1 + 1e-4), ("interval", 1 - 1e-4, 1 + 1e-4)):
        b = occ_rays.copy()
        if lo is not None:
            b[:, 6] = np.maximum(np.float32(1e-3), tocc * np.float32(lo))
        b[:, 7] = tocc * np.float32(hi)
        found = r.any_hit([slot], b).astype(bool)
        sh["node_visits_occluded_" + name] = int(r.stats()["node_visits"])
        sh["found_" + name] = int(found.sum())
    r.set_counting(False)
    unocc = sh["node_visits_full"] - sh["node_visits_occluded_full"]
    sh["node_visits_perfect"] = unocc + min(sh["node_visits_occluded_upper"], sh["node_visits_occluded_interval"])
    out["shadow"] = sh
    plain = st["node_visits"] + st["shadow_node_visits"]
    best = pst["node_visits"] + sh["node_visits_perfect"]
    out["frame_node_visits_plain"] = int(plain)
    out["frame_node_visits_perfect"] = int(best)
    out["node_visit_ratio"] = best / max(1, plain)
    return out


def orbit_cameras(n, seed):
    """n orbit cameras looking at the origin: azimuth stratified over n equal
    sectors, elevation 15..55 degrees, distance 2.6..3.6, drawn from the
    library's seeded uniform generator."""
    import ctypes as C
    import j3d_amd as J
    from j3d_amd import ffi
    lib = ffi.load()
    s = C.c_uint64(seed)
    out = []
    for i in range(n):
        az = 2.0 * math.pi * (i + lib.j3dg_uniform01(C.byref(s))) / n
        el = math.radians(15.0 + 40.0 * lib.j3dg_uniform01(C.byref(s)))
        dist = 2.6 + lib.j3dg_uniform01(C.byref(s))
        out.append(J.orbit_camera(dist, az, el))
    return out


# j3dg_build_opts.builder values (J3DG_BUILD_HOST, J3DG_BUILD_DEVICE)
BUILDERS = {"host": 0, "device": 1}

FUSED_PER_UNIT = ("%d B per node visit, 48 B per triangle test (primary and shadow), 32 B G-buffer write per pixel, "
                  "48 B hit completion record per hit")


def gbuffer_mismatches(a, b):
    """Values (bit patterns) that differ between two G-buffers, over every field."""
    return int(sum(np.count_nonzero(a[k].view(np.uint32) != b[k].view(np.uint32)) if a[k].dtype.itemsize == 4
                   else np.count_nonzero(a[k] != b[k]) for k in a.dtype.names))


def fused_bytes(st, npx, node_bytes):
    """Algorithmic bytes of one fused k_trace launch: node and triangle reads of
    the primary and shadow rays, the 32 B G-buffer write of every pixel and the
    48 B completion record (vertices, normal) of every hit."""
    return ((st["node_visits"] + st["shadow_node_visits"]) * node_bytes +
            (st["tri_tests"] + st["shadow_tri_tests"]) * 48 + 32 * npx + 48 * st["hits"])


def multi_camera(a):
    """300M-triangle workload: seeded displaced torus of 15000 x 10000 grid
    vertices (150M vertices, 300M triangles), `cameras` orbit cameras at
    3840 x 2160 with directional shadow rays, edges + matcap.  A step renders
    every camera once (one persistent k_trace launch + k_shade per camera).  The
    proposal network is retrained on this mesh.  Parity: every camera's
    G-buffer against the oracle on a crop x crop window, and the learned
    frames against the plain ones on every pixel."""
    import j3d_amd as J
    from j3d_amd import ffi
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from scenes import compare_gbuffers, pick_crop
    nu, nv = a.grid or (15000, 10000)
    w, h = a.resolution or (3840, 2160)
    t0 = time.time()
    v, t = J.torus(nu, nv, seed=1)
    log("torus %dx%d: %d verts %d tris in %.1fs" % (nu, nv, len(v), len(t), time.time() - t0))
    diag, pivot = J.scene_framing(v)
    cam = J.camera_with_vertical_fov(60.0)
    light = J.random_hemisphere(12)
    flags = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    mc, cav = J.make_matcap(0)
    r = J.Renderer(0)
    t0 = time.time()
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0), width=8, builder=BUILDERS[a.builder])
    info = r.mesh_info(slot)
    log("upload: build %.2fs copy %.2fs, %d nodes, depth %d, %.2f GB on device (%.1fs)" % (
        info["build_seconds"], info["upload_seconds"], info["n_nodes"], info["max_depth"],
        info["device_bytes"] / 1e9, time.time() - t0))
    r.set_matcap(mc, cav)
    cams = orbit_cameras(a.cameras, 21)
    frames = [J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light) for cs, ci in cams]
    targets = [r.create_target(w, h) for _ in frames]
    npx = w * h

    def counting(frs):
        r.set_counting(True)
        outs, sts = [], []
        for f, tg in zip(frs, targets):
            g, im = r.render(f, tg)
            sts.append(r.stats())
            outs.append((g, im))
        r.set_counting(False)
        return outs, sts

    def timed(frs):
        for _ in range(a.warmup):
            for f, tg in zip(frs, targets):
                r.render(f, tg, readback=False)
        r.synchronize()
        r.reset_timings()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            for f, tg in zip(frs, targets):
                r.render(f, tg, readback=False)
        r.synchronize()
        wall = time.perf_counter() - t0
        tm = r.timings()
        return wall, {k: float(np.mean(x)) for k, x in tm.items()}

    plain, pst = counting(frames)
    assert all(s["stack_overflow"] == 0 for s in pst)
    rays_per_step = sum(s["primary_rays"] + s["shadow_rays"] for s in pst)
    log("counting frames: %d rays per step, %d hit pixels" % (rays_per_step, sum(s["hits"] for s in pst)))
    tune = None
    if a.autotune:
        tune = r.autotune(frames[0], targets[0], repeats=3)
        log("autotune (plain): %s" % tune)
    wall, tm = timed(frames)
    plain_value = rays_per_step * a.steps / wall / 1e6
    plain_ms = wall / a.steps * 1e3
    log("plain: %.1f Mrays/s, %.2f ms per step" % (plain_value, plain_ms))
    value, ms_per_step = plain_value, plain_ms

    # every camera in one persistent k_trace + one k_shade_batch, captured in a HIP
    # graph (j3dg_render_batch); each target holds its camera's plain frame, which
    # the batch must reproduce exactly
    r.render_batch(frames, targets)
    r.synchronize()
    bmism = 0
    for (g0, im0), tg in zip(plain, targets):
        g, im = r.read_target(tg)
        bmism += sum(compare_gbuffers(g, g0).values()) + int(np.count_nonzero(im != im0))
    del g, im
    for _ in range(a.warmup):
        r.render_batch(frames, targets)
    r.synchronize()
    r.reset_timings()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        r.render_batch(frames, targets)
    r.synchronize()
    bwall = time.perf_counter() - t1
    graph_ms = float(np.mean(r.timings()["trace_ms"]))
    balgo = sum(fused_bytes(s, npx, 96) for s in pst)
    batched = {"mrays": rays_per_step * a.steps / bwall / 1e6, "ms_per_step": bwall / a.steps * 1e3,
               "graph_ms": graph_ms, "mismatches_vs_frames": bmism,
               "roofline": {"bound": "hbm", "achieved": balgo / (graph_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                            "unit": "GB/s", "frac": balgo / (graph_ms * 1e-3) / HBM_PEAK,
                            "kernel": "k_trace + k_shade_batch (one graph launch, all cameras)",
                            "algorithmic_bytes_per_launch": int(balgo)}}
    log("batched: %s" % batched)

    learned = None
    if not a.no_learned:
        n_train = a.train_rays or (1 << 25)
        log("training the proposal network on %d rays x %d epochs ..." % (n_train, a.epochs))
        rep_t = r.train_proposal(slot, n_rays=n_train, epochs=a.epochs, seed=7)
        log("training: %s" % rep_t)
        margin = 0.05
        ltune = None
        if a.autotune:
            lf0 = J.make_frame(w, h, cam, cams[0][0], cams[0][1], pivot, diag, [slot], flags=flags | ffi.LEARNED,
                               light_dir=light, margin=margin)
            ltune = r.autotune(lf0, targets[0], repeats=3)
            margin = ltune["margin"]
            log("autotune (learned): %s" % ltune)
        lframes = [J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags | ffi.LEARNED, light_dir=light,
                                margin=margin) for cs, ci in cams]
        louts, lsts = counting(lframes)
        mism = 0
        for (lg, lim), (g, im) in zip(louts, plain):
            mism += int(sum(np.count_nonzero(lg[k].view(np.uint32) != g[k].view(np.uint32))
                            if lg[k].dtype.itemsize == 4 else np.count_nonzero(lg[k] != g[k]) for k in lg.dtype.names))
            mism += int(np.count_nonzero(lim != im))
        del louts
        lwall, ltm = timed(lframes)
        learned = {"mrays": rays_per_step * a.steps / lwall / 1e6, "ms_per_step": lwall / a.steps * 1e3,
                   "mlp_ms_per_frame": ltm["mlp_ms"], "trace_ms_per_frame": ltm["trace_ms"], "margin": margin,
                   "resolved": int(sum(s["learned_resolved"] for s in lsts)),
                   "fallback": int(sum(s["learned_fallback"] for s in lsts)),
                   "primary_node_visits": int(sum(s["node_visits"] for s in lsts)),
                   "plain_primary_node_visits": int(sum(s["node_visits"] for s in pst)),
                   "mismatches_vs_plain": mism,
                   "training": {k: rep_t[k] for k in ("steps", "loss_first", "loss_last", "train_ms", "label_ms",
                                                        "hit_fraction", "val_hit_accuracy", "val_mean_abs_error",
                                                        "val_resolved")},
                   "autotune": ltune}
        log("learned: %.1f Mrays/s, %d G-buffer/RGBA mismatches vs plain" % (learned["mrays"], mism))
        value, ms_per_step = learned["mrays"], learned["ms_per_step"]
        # the learned frames batched: the network per camera, then one graph launch
        r.render_batch(lframes, targets)
        r.synchronize()
        lbm = 0
        for (g0, im0), tg in zip(plain, targets):
            g, im = r.read_target(tg)
            lbm += sum(compare_gbuffers(g, g0).values()) + int(np.count_nonzero(im != im0))
        del g, im
        for _ in range(a.warmup):
            r.render_batch(lframes, targets)
        r.synchronize()
        r.reset_timings()
        t1 = time.perf_counter()
        for _ in range(a.steps):
            r.render_batch(lframes, targets)
        r.synchronize()
        lbwall = time.perf_counter() - t1
        ltm = r.timings()
        learned["batched"] = {"mrays": rays_per_step * a.steps / lbwall / 1e6, "ms_per_step": lbwall / a.steps * 1e3,
                              "graph_ms": float(np.mean(ltm["trace_ms"])), "mlp_ms": float(np.mean(ltm["mlp_ms"])),
                              "mismatches_vs_plain": lbm}
        log("learned, batched: %s" % learned["batched"])
        # the workload's configuration is learned and batched
        value, ms_per_step = learned["batched"]["mrays"], learned["batched"]["ms_per_step"]

    algo = sum(fused_bytes(s, npx, 96) for s in pst) / len(pst)
    achieved = algo / (tm["trace_ms"] * 1e-3)
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None,
            "kernel": "k_trace (one launch: primary, hit completion, shadow)", "launch_ms": tm["trace_ms"],
            "algorithmic_bytes_per_launch": int(algo),
            "per_unit": FUSED_PER_UNIT % 96}

    if a.no_cpu_baseline or a.profile_only:
        r.close()
        out = {"metric": "Mrays/s @300M-tri %d cameras %dx%d (profile run: no oracle parity)" % (a.cameras, w, h),
               "value": value, "plain_mrays": plain_value, "roofline": roof, "batched": batched, "learned": learned}
        print(json.dumps(out), flush=True)
        return
    # oracle parity on a crop of every camera, timed as the CPU baseline
    import oracle as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    log("oracle BVH over %d triangles on %d threads ..." % (len(t), threads))
    t0 = time.time()
    m = O.Mesh(v, t, db_id=J.make_db_id(1, 0), threads=threads)
    bvh_s = time.time() - t0
    crops, cpu_s, cpu_rays, ties = [], 0.0, 0, 0
    g_mism, rgba_max, rgba_px = {}, 0, 0
    for i, (f, (g, im)) in enumerate(zip(frames, plain)):
        x0, y0, x1, y1 = pick_crop(g, a.crop)
        crops.append([x0, y0, x1, y1])
        cf = J.make_frame(w, h, cam, cams[i][0], cams[i][1], pivot, diag, [slot], flags=flags, light_dir=light,
                          rect=(x0, y0, x1, y1))
        og = g.copy()
        og[y0:y1 + 1, x0:x1 + 1] = np.zeros(1, O.PIXEL_DTYPE)[0]
        og["object_id"][y0:y1 + 1, x0:x1 + 1] = 0xFFFFFFFF
        t1 = time.time()
        og, ost = O.update_canvas([m], cf, gbuf=og, threads=threads, count_ties=True)
        cpu_s += time.time() - t1
        cpu_rays += ost["primary_rays"] + ost["shadow_rays"]
        ties += ost.get("near_ties", 0)
        for k, n in compare_gbuffers(g[y0:y1 + 1, x0:x1 + 1], og[y0:y1 + 1, x0:x1 + 1]).items():
            g_mism[k] = g_mism.get(k, 0) + n
        # the shading pass over the whole frame, the crop coming from the oracle
        oim = O.canvas_to_image(og, f, mc, cav)
        d = np.abs(oim.view(np.uint8).astype(np.int16) - im.view(np.uint8).astype(np.int16))
        rgba_max = max(rgba_max, int(d.max()))
        rgba_px += int((d.reshape(im.shape + (4,)).max(axis=-1) > 0).sum())
        log("camera %d crop %s: %s" % (i, crops[-1], {k: n for k, n in g_mism.items() if n}))
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    cpu = {"value": cpu_rays / cpu_s / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": "%d crops of %dx%d pixels (one per camera, %d primary + shadow rays), oracle update_canvas, "
                     "%.1fs (oracle BVH build %.1fs excluded)" % (len(crops), a.crop, a.crop, cpu_rays, cpu_s, bvh_s),
           "cpu_model": model}
    parity = {"crops": crops, "gbuffer_mismatches": g_mism, "near_ties": ties,
              "rgba_max_diff": rgba_max, "rgba_pixels_differing": rgba_px}

    desc = {"workload": "displaced torus R=1 r=0.35, %dx%d grid, %d triangles, fBm 6 octaves; %d orbit cameras "
                        "%dx%d, primary + directional shadow rays, edges + matcap" % (nu, nv, len(t), a.cameras, w, h),
            "triangles": int(len(t)), "width": w, "height": h, "cameras": a.cameras, "shadow": "directional",
            "launches": "value: learned frames batched -- the network per camera, then all cameras in one "
                        "graph-captured k_trace + k_shade_batch; 'plain_mrays' and learned 'mrays': one launch "
                        "per camera; 'batched': plain frames batched"}
    out = {"metric": "Mrays/s @300M-tri %d cameras %dx%d + shadows, learned-bounded and batched (one graph "
                     "launch) vs plain GPU traversal; %% HBM peak" % (a.cameras, w, h),
           "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic", "config": desc, "roofline": roof,
           "plain_mrays": plain_value, "plain_ms_per_step": plain_ms, "batched": batched, "learned": learned,
           "learned_over_plain": (learned["mrays"] / plain_value) if learned else None,
           "learned_batched_over_plain_batched": (learned["batched"]["mrays"] / batched["mrays"]) if learned else None,
           "autotune_plain": tune, "rays_per_step": int(rays_per_step),
           "primary_rays": int(sum(s["primary_rays"] for s in pst)),
           "shadow_rays": int(sum(s["shadow_rays"] for s in pst)),
           "trace_ms_per_frame": tm["trace_ms"], "shade_ms_per_frame": tm["shade_ms"],
           "bvh": {"width": 8, "leaf_size": 2, "builder": a.builder, "nodes": int(info["n_nodes"]),
                   "depth": int(info["max_depth"]), "build_s": info["build_seconds"],
                   "upload_s": info["upload_seconds"], "device_gb": info["device_bytes"] / 1e9},
           "tuning": r.get_tuning(), "parity": parity, "cpu_baseline": cpu}
    r.close()
    print(json.dumps(out), flush=True)


def mesh_set(n_meshes, seed):
    """n displaced tori (R=1, r=0.35, fBm) whose triangle counts are drawn
    log-uniform in [1e4, 1e7]: grid nu x nv with nu/nv ~ R/r, 2 nu nv triangles."""
    import ctypes as C
    import j3d_amd as J
    from j3d_amd import ffi
    lib = ffi.load()
    s = C.c_uint64(seed)
    out = []
    for i in range(n_meshes):
        target = 10.0 ** (4.0 + 3.0 * lib.j3dg_uniform01(C.byref(s)))
        nv = max(3, int(round(math.sqrt(target / 2.0 / 2.857))))
        nu = max(3, int(round(target / 2.0 / nv)))
        v, t = J.torus(nu, nv, seed=100 + i)
        out.append((nu, nv, v, t))
    return out


def mesh_set_bench(a):
    """16 meshes of uneven size (displaced tori, 1e4..1e7 triangles, log-uniform)
    and 64 orbit cameras at resolutions drawn from {256x256, 640x480, 1280x720,
    1920x1080}; camera i shows mesh i mod 16 (single-object frames), edges +
    matcap.  A step renders all 64 frames in one persistent k_trace launch +
    one k_shade_batch, captured in a HIP graph and replayed
    (j3dg_render_batch).  The autotuner (j3dg_autotune_batch) sweeps BVH width
    {4, 8} x rays per work item {1, 2, 4} x LDS stack depth {8, 16, 32} on that
    one launch.  Parity: every batch frame equal to the same frame rendered
    alone, the width-4 and width-8 frames identical, and a 64 x 64 crop of every
    frame against the oracle.  The frames rendered one after another (64
    launches) are timed beside it."""
    import ctypes as C
    import j3d_amd as J
    from j3d_amd import ffi
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from scenes import compare_gbuffers
    lib = ffi.load()
    t0 = time.time()
    meshes = mesh_set(16, 41)
    n_tris = [len(m[3]) for m in meshes]
    log("16 meshes, %d triangles (%d .. %d) in %.1fs" % (sum(n_tris), min(n_tris), max(n_tris), time.time() - t0))
    r = J.Renderer(0)
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    t0 = time.time()
    slots = {4: [], 8: []}
    framing = []
    for i, (nu, nv, v, t) in enumerate(meshes):
        for w in (4, 8):
            slots[w].append(r.upload_mesh(v, t, db_id=J.make_db_id(1, i), width=w))
        framing.append(J.scene_framing(v))
    log("uploads (both widths): %.1fs" % (time.time() - t0))
    s = C.c_uint64(43)
    sizes = [(256, 256), (640, 480), (1280, 720), (1920, 1080)]
    cam = J.camera_with_vertical_fov(60.0)
    cams = []
    for i in range(64):
        az = 2.0 * math.pi * lib.j3dg_uniform01(C.byref(s))
        el = math.radians(-30.0 + 90.0 * lib.j3dg_uniform01(C.byref(s)))
        dist = 2.6 + lib.j3dg_uniform01(C.byref(s))
        wh = sizes[min(3, int(4 * lib.j3dg_uniform01(C.byref(s))))]
        cs, ci = J.orbit_camera(dist, az, el)
        cams.append((cs, ci, wh))
    flags = ffi.EDGES | ffi.SHADING
    frames = {w: [] for w in (4, 8)}
    for i, (cs, ci, (fw, fh)) in enumerate(cams):
        diag, pivot = framing[i % 16]
        for w in (4, 8):
            frames[w].append(J.make_frame(fw, fh, cam, cs, ci, pivot, diag, [slots[w][i % 16]], flags=flags))
    targets = [r.create_target(fw, fh) for _, _, (fw, fh) in cams]
    npx = sum(fw * fh for _, _, (fw, fh) in cams)

    tune = r.autotune_batch(frames[4], frames[8], targets, repeats=3)
    log("batch autotune: %s" % tune)
    bw = tune["width"]
    bf = frames[bw]

    # the frames one at a time: reference output and rays; then the batch
    r.set_counting(True)
    rays, hits, nodes, tris = 0, 0, 0, 0
    single = {4: [], 8: []}
    for w in (4, 8):
        for f, tg in zip(frames[w], targets):
            g, im = r.render(f, tg)
            single[w].append((g, im))
            if w == bw:
                st = r.stats()
                rays += st["primary_rays"] + st["shadow_rays"]
                hits += st["hits"]
                nodes += st["node_visits"]
                tris += st["tri_tests"]
    r.set_counting(False)
    width_mism = 0
    for (g4, i4), (g8, i8) in zip(single[4], single[8]):
        width_mism += sum(compare_gbuffers(g4, g8).values()) + int(np.count_nonzero(i4 != i8))
    # every target holds its frame already: a batch over them must leave them the same
    r.render_batch(bf, targets)
    r.synchronize()
    batch_mism = 0
    outs = []
    for (g0, im0), tg in zip(single[bw], targets):
        g, im = r.read_target(tg)
        outs.append((g, im))
        batch_mism += sum(compare_gbuffers(g, g0).values()) + int(np.count_nonzero(im != im0))
    log("batch vs single frames: %d mismatches; width 4 vs 8: %d" % (batch_mism, width_mism))

    for _ in range(a.warmup):
        r.render_batch(bf, targets)
    r.synchronize()
    r.reset_timings()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        r.render_batch(bf, targets)
    r.synchronize()
    wall = time.perf_counter() - t1
    tm = r.timings()
    value = rays * a.steps / wall / 1e6
    graph_ms = float(np.mean(tm["trace_ms"]))  # one graph launch: k_reset + k_trace + k_shade_batch
    log("batch: %.1f Mrays/s, %.2f ms per launch of 64 frames (graph %.2f ms)" % (value, wall / a.steps * 1e3,
                                                                               graph_ms))
    # the same frames as 64 launches, for comparison
    for f, tg in zip(bf, targets):
        r.render(f, tg, readback=False)
    r.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        for f, tg in zip(bf, targets):
            r.render(f, tg, readback=False)
    r.synchronize()
    seq_wall = time.perf_counter() - t1
    seq_value = rays * a.steps / seq_wall / 1e6
    log("sequential: %.1f Mrays/s, %.2f ms per 64 frames" % (seq_value, seq_wall / a.steps * 1e3))
    node_bytes = 96 if bw == 8 else 64
    algo = nodes * node_bytes + tris * 48 + 32 * npx
    achieved = algo / (graph_ms * 1e-3)

    # oracle parity on a 64 x 64 crop of every batch frame
    import oracle as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    t0 = time.time()
    om = [O.Mesh(v, t, db_id=J.make_db_id(1, i), threads=threads) for i, (_, _, v, t) in enumerate(meshes)]
    bvh_s = time.time() - t0
    crop_mism, cpu_s, cpu_rays = {}, 0.0, 0
    for i, (f, (g, im)) in enumerate(zip(bf, outs)):
        fw, fh = f.width, f.height
        x0, y0 = fw // 2 - 32, fh // 2 - 32
        rect = (x0, y0, x0 + 63, y0 + 63)
        cf = J.make_frame(fw, fh, cam, cams[i][0], cams[i][1], framing[i % 16][1], framing[i % 16][0],
                          [slots[bw][i % 16]], flags=flags, rect=rect)
        og = np.zeros((fh, fw), O.PIXEL_DTYPE)
        og["object_id"] = 0xFFFFFFFF
        t1 = time.time()
        og, ost = O.update_canvas([om[i % 16]], cf, gbuf=og, threads=threads)
        cpu_s += time.time() - t1
        cpu_rays += ost["primary_rays"] + ost["shadow_rays"]
        for k, n in compare_gbuffers(g[y0:y0 + 64, x0:x0 + 64], og[y0:y0 + 64, x0:x0 + 64]).items():
            crop_mism[k] = crop_mism.get(k, 0) + n
    log("oracle crops: %s" % {k: n for k, n in crop_mism.items() if n})
    desc = {"workload": "16 displaced tori, %d triangles in all (%d .. %d, log-uniform from 1e4..1e7); 64 orbit "
                        "cameras at 256x256 / 640x480 / 1280x720 / 1920x1080, camera i -> mesh i mod 16; edges + "
                        "matcap" % (sum(n_tris), min(n_tris), max(n_tris)),
            "meshes": [[int(nu), int(nv), int(len(t))] for nu, nv, _, t in meshes],
            "resolutions": [list(c[2]) for c in cams],
            "launches": "all 64 frames in one persistent k_trace + one k_shade_batch, captured in a HIP graph"}
    out = {"metric": "Mrays/s, 16 meshes x 64 cameras of mixed size in one graph-captured launch", "value": value,
           "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic", "config": desc,
           "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK, "traffic": None,
                        "kernel": "k_trace + k_shade_batch (one graph launch; the shading pass's time included)",
                        "launch_ms": graph_ms, "algorithmic_bytes_per_launch": int(algo),
                        "per_unit": "%d B per node visit, 48 B per triangle test, 32 B G-buffer write per pixel"
                                    % node_bytes},
           "rays_per_step": int(rays), "hit_pixels": int(hits), "autotune": tune, "width": bw,
           "sequential_mrays": seq_value, "sequential_ms_per_step": seq_wall / a.steps * 1e3,
           "batch_vs_single_mismatches": batch_mism, "width4_vs_width8_mismatches": width_mism,
           "parity": {"crop": "64x64 at every frame's centre", "gbuffer_mismatches": crop_mism},
           "cpu_baseline": {"value": cpu_rays / cpu_s / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
                            "sample": "64 crops of 64x64 pixels (%d rays), oracle update_canvas, %.2fs (oracle BVH "
                                      "builds %.1fs excluded)" % (cpu_rays, cpu_s, bvh_s)}}
    r.close()
    print(json.dumps(out), flush=True)


def scene_bench(a):
    """A scene of 256 objects, more than a frame's object table in LDS holds, so
    frames run through the top-level BVH over the objects: four seeded displaced
    tori of 20,000 triangles (100 x 100 grid) repeated in seeded rigid poses in
    a cube of side 16 (5.12M triangles in all), one orbit camera at distance
    30, 1920x1080 primary + directional shadow rays, edges + matcap.  Parity: a
    256 x 256 crop of the G-buffer against the oracle, which loops over all 256
    objects (no top level)."""
    import copy
    import j3d_amd as J
    from j3d_amd import ffi
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from scenes import compare_gbuffers, pick_crop
    rng = np.random.default_rng(5)
    bases = [J.torus(100, 100, seed=s) for s in range(4)]
    poses = []
    for i in range(256):
        q, rr = np.linalg.qr(rng.normal(size=(3, 3)))
        q = q * np.sign(np.diag(rr))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        m = np.eye(4)
        m[:3, :3] = q
        m[:3, 3] = rng.uniform(-8.0, 8.0, 3)
        poses.append(np.ascontiguousarray(m.T, np.float32).reshape(16))
    r = J.Renderer(0)
    t0 = time.time()
    slots = [r.upload_mesh(bases[i % 4][0], bases[i % 4][1], db_id=J.make_db_id(1, i), cs=poses[i],
                           builder=BUILDERS[a.builder]) for i in range(256)]
    log("256 objects, %d triangles, uploaded in %.1fs" % (256 * len(bases[0][1]), time.time() - t0))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    w, h = 1920, 1080
    cs, ci = J.orbit_camera(30.0, 0.7, 0.45)
    cam = J.camera_with_vertical_fov(60.0)
    light = J.random_hemisphere(12)
    flags = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT
    fr = J.make_frame(w, h, cam, cs, ci, np.zeros(3, np.float32), 18.0, slots, flags=flags, light_dir=light)
    tg = r.create_target(w, h)
    r.set_counting(True)
    g, _ = r.render(fr, tg)
    st = r.stats()
    r.set_counting(False)
    rays = int(st["primary_rays"] + st["shadow_rays"])
    for _ in range(a.warmup):
        r.render(fr, tg, readback=False)
    r.synchronize()
    r.reset_timings()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        r.render(fr, tg, readback=False)
    r.synchronize()
    wall = time.perf_counter() - t1
    tm = {k: float(np.mean(v)) for k, v in r.timings().items()}
    value = rays * a.steps / wall / 1e6
    hit = g["object_id"] != 0xFFFFFFFF
    log("scene: %.1f Mrays/s, %.3f ms per frame (trace %.3f ms), %d rays, %d hit pixels on %d objects" % (
        value, wall / a.steps * 1e3, tm["trace_ms"], rays, int(hit.sum()), len(np.unique(g["db_id"][hit]))))
    parity, cpu = None, None
    if not a.no_cpu_baseline and not a.profile_only:
        import oracle as O
        base_m = [O.Mesh(v, t) for v, t in bases]
        inst = []
        for i in range(256):  # instances share their base mesh's oracle BVH
            m = copy.copy(base_m[i % 4])
            m.cs = poses[i]
            m.db_id = J.make_db_id(1, i)
            inst.append(m)
        rect = pick_crop(g, 256)
        frc = J.make_frame(w, h, cam, cs, ci, np.zeros(3, np.float32), 18.0, slots, flags=flags, light_dir=light,
                           rect=rect)
        og = np.zeros((h, w), O.PIXEL_DTYPE)
        og["object_id"] = 0xFFFFFFFF
        t2 = time.time()
        og, ost = O.update_canvas(inst, frc, gbuf=og)
        cpu_s = time.time() - t2
        x0, y0, x1, y1 = rect
        mm = compare_gbuffers(g[y0:y1 + 1, x0:x1 + 1], og[y0:y1 + 1, x0:x1 + 1])
        parity = {"crop": list(rect), "gbuffer_mismatches": mm}
        crays = int(ost["primary_rays"] + ost["shadow_rays"])
        cpu = {"value": crays / cpu_s / 1e6, "unit": "Mrays/s", "cores": os.cpu_count() or 1, "kind": "port",
               "sample": "oracle update_canvas on the 256x256 crop over all 256 objects (%d rays, %.1fs)" % (crays, cpu_s)}
        for m in inst:  # the copies do not own the shared BVHs
            m.bvh = None
        log("parity on crop %s: %s; cpu %.2f Mrays/s" % (rect, mm, cpu["value"]))
    out = {"metric": "Mrays/s (256-object scene through the top-level BVH, 1080p primary + shadow rays)",
           "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "256 objects: 4 displaced tori of 20,000 triangles in seeded rigid poses; "
                                  "1920x1080 primary + directional shadow rays, edges + matcap",
                      "objects": 256, "triangles": 256 * int(len(bases[0][1])), "builder": a.builder},
           "trace_ms": tm["trace_ms"], "primary_node_visits_per_ray": st["node_visits"] / max(1, st["primary_rays"]),
           "shadow_node_visits_per_ray": st["shadow_node_visits"] / max(1, st["shadow_rays"]),
           "parity": parity, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    r.close()


def voxel_bench(a):
    """write_vox's voxel grid (j3d/vox.cpp:278-376) of the 28,000,000-triangle
    torus at max_dim 512 (--vox-dim): rays along x, y and z through every cell
    centre, every hit of each ray (find_all_triangles, vox.cpp:333), each hit to
    its voxel and palette index.  value = voxel rays per second of the whole
    j3dg_voxelize call (hit lists, voxel writes, grid to host), median of
    --steps calls.  Parity: the grid at max_dim 96 against the oracle's."""
    import j3d_amd as J
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    t0 = time.time()
    v, t = J.torus(5000, 2800, seed=1)
    log("torus 5000x2800: %d verts %d tris in %.1fs" % (len(v), len(t), time.time() - t0))
    r = J.Renderer(0)
    slot = r.upload_mesh(v, t, builder=BUILDERS[a.builder])
    info = r.mesh_info(slot)
    log("upload (%s build): build %.2fs copy %.2fs" % (a.builder, info["build_seconds"], info["upload_seconds"]))
    md = a.vox_dim
    for _ in range(max(1, a.warmup)):
        dims, data = r.voxelize(slot, md)
    times = []
    for _ in range(a.steps):
        t1 = time.perf_counter()
        dims, data = r.voxelize(slot, md)
        times.append(time.perf_counter() - t1)
    ms = 1e3 * float(np.median(times))
    rays = dims[1] * dims[2] + dims[2] * dims[0] + dims[0] * dims[1]
    log("voxelize max_dim %d: dims %s, %d rays, %d voxels set, %.2f ms" % (md, dims, rays, int((data > 0).sum()), ms))
    parity, cpu = None, None
    if not a.no_cpu_baseline and not a.profile_only:
        import oracle as O
        t2 = time.time()
        m = O.Mesh(v, t)
        build_s = time.time() - t2
        pd = 96
        gd, gdata = r.voxelize(slot, pd)
        t3 = time.time()
        od, odata = O.voxelize(m, pd)
        cpu_s = time.time() - t3
        parity = {"max_dim": pd, "dims": list(od), "voxels_differing": int((gdata != odata).sum()),
                  "voxels_set": int((odata > 0).sum())}
        crays = od[1] * od[2] + od[2] * od[0] + od[0] * od[1]
        cpu = {"value": crays / cpu_s / 1e6, "unit": "M voxel rays/s", "cores": 1, "kind": "port",
               "sample": "oracle write_vox restatement at max_dim %d (%d rays, %.1fs; oracle BVH build %.1fs "
                         "excluded), one thread" % (pd, crays, cpu_s, build_s)}
        log("parity at max_dim %d: %s; cpu %.3f M rays/s" % (pd, parity, cpu["value"]))
    out = {"metric": "M voxel rays/s (write_vox grid of the 28M-triangle torus, max_dim %d)" % md,
           "value": rays / (ms / 1e3) / 1e6, "unit": "M voxel rays/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "displaced torus R=1 r=0.35, 5000x2800 grid, 28,000,000 triangles; "
                                  "write_vox voxel grid", "max_dim": md, "dims": list(dims), "rays": int(rays),
                      "builder": a.builder},
           "parity": parity, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    r.close()


def splat_bench(a):
    """Point clouds drawn over a rendered frame (canvas::render_pointclouds_on_image,
    j3d/canvas.cpp:952-1030): the 1920x1080 frame of the 28,000,000-triangle
    torus (shadows, edges, matcap), then one cloud of 56,000,000 points -- the
    torus's 14,000,000 vertices in four copies pushed along random directions
    by up to 0.02 -- with normals and colours, shading on.  value = points per
    second of the whole j3dg_splat_points call (z-buffer from the G-buffer,
    depth merge, resolve), median of --steps calls, each over the same rendered
    target.  Parity: the oracle draws the same points in order over the same
    frame; G-buffer and image compared in full."""
    import j3d_amd as J
    from j3d_amd import ffi
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    v, t, cam, cs, ci, pivot, diag, light, w, h, flags, desc = workload("torus28m")
    rng = np.random.default_rng(5)
    copies = 4
    xyz = np.concatenate([v + rng.uniform(-0.02, 0.02, size=v.shape).astype(np.float32) for _ in range(copies)])
    nrm = rng.normal(size=xyz.shape).astype(np.float32)
    col = rng.integers(0, 2 ** 32, len(xyz), dtype=np.uint64).astype(np.uint32)
    n = len(xyz)
    log("cloud: %d points" % n)
    r = J.Renderer(0)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0))
    cloud = r.upload_points(xyz, normals=nrm, colors=col, db_id=J.make_db_id(2, 0))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    tg = r.create_target(w, h)
    fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light)
    P, _ = J.projection(cam, w, h)
    sflags = flags | ffi.SHADING
    times, depth_ms = [], []
    for k in range(a.warmup + a.steps):
        r.render(fr, tg, readback=False)
        r.synchronize()
        t1 = time.perf_counter()
        r.splat_points(tg, P, ci, sflags, [cloud])
        if k >= a.warmup:
            times.append(time.perf_counter() - t1)
            depth_ms.append(r.stats()["splat_ms"])
    ms = 1e3 * float(np.median(times))
    kms = float(np.median(depth_ms))
    g0, im0 = r.render(fr, tg)
    r.splat_points(tg, P, ci, sflags, [cloud])
    g1, im1 = r.read_target(tg)
    drawn = int((g1["db_id"] == J.make_db_id(2, 0)).sum())
    log("splat: %d points in %.3f ms (%.1f G points/s), %d pixels drawn" % (n, ms, n / ms / 1e6, drawn))
    parity, cpu = None, None
    if not a.no_cpu_baseline and not a.profile_only:
        import oracle as O
        t2 = time.time()
        og, oim = O.splat_points(g0, im0, P, ci, sflags, [{"xyz": xyz, "normals": nrm, "colors": col,
                                                            "db_id": J.make_db_id(2, 0)}])
        cpu_s = time.time() - t2
        parity = {"gbuffer_bytes_differing": int((g1.view(np.uint8) != og.view(np.uint8)).sum()),
                  "rgba_differing": int((im1 != oim).sum()), "pixels_drawn": drawn}
        cpu = {"value": n / cpu_s / 1e6, "unit": "M points/s", "cores": 1, "kind": "port",
               "sample": "oracle render_pointclouds_on_image restatement, all %d points over the same frame, "
                         "%.1fs, one thread" % (n, cpu_s)}
        log("parity: %s; cpu %.1f M points/s" % (parity, cpu["value"]))
    # per point: its 12 B of xyz, read once, and the 8 B z-buffer key of its pixel (read, and
    # written by an atomicMax only when it can grow); the keys (16.6 MB) stay in L2/MALL
    point_bytes = 20 * n
    out = {"metric": "G points/s (56M-point cloud splatted over the 28M-triangle 1080p frame)",
           "value": n / (ms / 1e3) / 1e9, "unit": "G points/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "torus28m frame (1920x1080, shadows) + %d points (4 jittered copies of the "
                                  "torus vertices) with normals and colours, shading" % n, "points": n,
                      "width": w, "height": h},
           "roofline": {"bound": "hbm", "achieved": point_bytes / (kms / 1e3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": point_bytes / (kms / 1e3) / 1e9 / 8000.0, "traffic": None,
                        "kernel": "k_splat_depth (HIP events around it, median of the timed calls)",
                        "launch_ms": kms,
                        "per_unit": "12 B xyz + 8 B z-buffer key per point (the keys are a random access "
                                    "served from L2/MALL: the HBM roofline is an upper bound it cannot reach)"},
           "parity": parity, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    r.close()


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame would shard by image "
              "tiles or cameras with a final RGBA gather); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "torus300m":
        return multi_camera(a)
    if a.workload == "set16x64":
        return mesh_set_bench(a)
    if a.workload == "vox28m":
        return voxel_bench(a)
    if a.workload == "scene256":
        return scene_bench(a)
    if a.workload == "splat56m":
        return splat_bench(a)
    import j3d_amd as J
    from j3d_amd import ffi
    v, t, cam, cs, ci, pivot, diag, light, w, h, flags, desc = workload(a.workload)
    mc, cav = J.make_matcap(0)
    r = J.Renderer(0)
    tun = r.get_tuning()
    if a.stack_depth:
        tun["stack_depth"] = a.stack_depth
    if a.block:
        tun["block_size"] = a.block
    if a.coop_tail >= 0:
        tun["coop_tail"] = a.coop_tail
    r.set_tuning(**tun)
    t0 = time.time()
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 0), width=a.width, leaf_size=a.leaf_size,
                         builder=BUILDERS[a.builder])
    info = r.mesh_info(slot)
    log("upload: build %.2fs copy %.2fs, %d nodes, depth %d, %.2f GB on device" % (
        info["build_seconds"], info["upload_seconds"], info["n_nodes"], info["max_depth"],
        info["device_bytes"] / 1e9))
    r.set_matcap(mc, cav)
    tg = r.create_target(w, h)
    rect = None
    if a.rows:
        y0 = max(0, h // 2 - a.rows // 2)
        rect = (0, y0, w - 1, min(h - 1, y0 + a.rows - 1))
    fr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light, rect=rect)

    # counters (node visits, triangle tests) from one counting frame
    r.set_counting(True)
    gpu_g, gpu_im = r.render(fr, tg)
    st = r.stats()
    r.set_counting(False)
    assert st["stack_overflow"] == 0
    rays_per_frame = st["primary_rays"] + st["shadow_rays"]

    tune = None
    if a.autotune:
        tune = r.autotune(fr, tg, repeats=5)
        log("autotune (plain): %s" % tune)

    def timed(frame):
        for _ in range(a.warmup):
            r.render(frame, tg, readback=False)
        r.synchronize()
        r.reset_timings()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            r.render(frame, tg, readback=False)
        r.synchronize()
        wall = time.perf_counter() - t0
        tm = r.timings()
        return wall, {k: float(np.mean(v)) for k, v in tm.items()}

    wall, tm = timed(fr)
    trace_ms, shadow_ms, shade_ms = tm["trace_ms"], tm["shadow_ms"], tm["shade_ms"]
    plain_ms = wall / a.steps * 1e3
    plain_value = rays_per_frame * a.steps / wall / 1e6
    ms_per_step, value = plain_ms, plain_value

    orbit = None
    if a.workload == "torus28m" and a.orbit_frames > 0:
        otg = r.create_target(w, h)  # its own target: misses keep a target's earlier fields
        orbit = orbit_sequence(r, otg, cam, pivot, diag, slot, flags, light, w, h, a.orbit_frames, a.steps)
        r.destroy_target(otg)
        log("orbit: %s" % orbit)

    learned = None
    if not a.no_learned:
        n_train = a.train_rays or (1 << 24)
        log("training the proposal network on %d rays x %d epochs ..." % (n_train, a.epochs))
        rep_t = r.train_proposal(slot, n_rays=n_train, epochs=a.epochs, seed=7)
        log("training: %s" % rep_t)
        lfr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags | ffi.LEARNED, light_dir=light,
                           rect=rect, margin=0.05)
        ltune = None
        if a.autotune:
            ltune = r.autotune(lfr, tg, repeats=5)
            log("autotune (learned): %s" % ltune)
            lfr = J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags | ffi.LEARNED, light_dir=light,
                               rect=rect, margin=ltune["margin"])
        # the reference: a plain frame over the same target contents (a miss keeps the
        # target's earlier mark and barycentrics, j3d/canvas.cpp:859-866)
        pg, pim = r.render(fr, tg)
        r.set_counting(True)
        lg, lim = r.render(lfr, tg)
        lst = r.stats()
        r.set_counting(False)
        mism = gbuffer_mismatches(lg, pg)
        lwall, ltm = timed(lfr)
        learned = {"mrays": rays_per_frame * a.steps / lwall / 1e6, "ms_per_step": lwall / a.steps * 1e3,
                   "mlp_ms": ltm["mlp_ms"], "trace_ms": ltm["trace_ms"], "shadow_stage_ms": ltm["shadow_ms"],
                   "margin": lfr.margin, "resolved": int(lst["learned_resolved"]),
                   "fallback": int(lst["learned_fallback"]), "hit_pixels": int(lst["hits"]),
                   "primary_node_visits": int(lst["node_visits"]), "plain_primary_node_visits": int(st["node_visits"]),
                   "gbuffer_mismatches_vs_plain": mism, "rgba_equal_plain": bool(np.array_equal(lim, pim)),
                   "training": {k: rep_t[k] for k in ("steps", "loss_first", "loss_last", "train_ms", "label_ms",
                                                        "hit_fraction", "val_hit_accuracy", "val_mean_abs_error",
                                                        "val_resolved")},
                   "autotune": ltune}
        if a.workload == "torus28m":
            learned["perfect_bound_ceiling"] = bound_ceiling(r, tg, fr, slot, v, t, gpu_g, st, a.steps, a.warmup)
            log("perfect-bound ceiling: %s" % learned["perfect_bound_ceiling"])
        ms_per_step, value = learned["ms_per_step"], learned["mrays"]

    node_bytes = 96 if a.width == 8 else 64  # quantized wide node
    npx = w * h
    fused = r.get_tuning()["split_passes"] == 0
    if fused:
        algo_bytes = fused_bytes(st, npx, node_bytes)
        per_unit = FUSED_PER_UNIT % node_bytes
        kname = "k_trace (one launch: primary, hit completion, shadow)"
    else:
        # primary k_trace launch: node and triangle reads + 20 B of G-buffer writes per pixel
        algo_bytes = st["node_visits"] * node_bytes + st["tri_tests"] * 48 + 20 * npx
        per_unit = "%d B per node visit, 48 B per triangle test, 20 B per pixel" % node_bytes
        kname = "k_trace (primary pass)"
    achieved = algo_bytes / (trace_ms * 1e-3)
    traffic = read_traffic(a.workload, fused)
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": (traffic or {}).get("hbm_bytes_per_launch") if traffic else None,
            "kernel": kname, "launch_ms": trace_ms,
            "algorithmic_bytes_per_launch": int(algo_bytes), "per_unit": per_unit}

    out = {"metric": METRIC, "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": desc,
           "roofline": roof,
           "plain_mrays": plain_value, "plain_ms_per_step": plain_ms,
           "orbit": orbit,
           "learned": learned,
           "learned_over_plain": (learned["mrays"] / plain_value) if learned else None,
           "autotune_plain": tune,
           "primary_mrays": st["primary_rays"] / (ms_per_step * 1e-3) / 1e6,
           "rays_per_frame": int(rays_per_frame), "primary_rays": int(st["primary_rays"]),
           "shadow_rays": int(st["shadow_rays"]), "hit_pixels": int(st["hits"]),
           "primary_node_visits_per_ray": st["node_visits"] / max(1, st["primary_rays"]),
           "primary_tri_tests_per_ray": st["tri_tests"] / max(1, st["primary_rays"]),
           "trace_ms": trace_ms, "shadow_stage_ms": shadow_ms, "shade_ms": shade_ms,
           "shadow_node_visits_per_ray": st["shadow_node_visits"] / max(1, st["shadow_rays"]),
           "wave_trips": int(st["wave_trips"] + st["shadow_wave_trips"]),
           "lane_steps": int(st["lane_steps"] + st["shadow_lane_steps"]),
           "lane_steps_per_wave_trip": (st["lane_steps"] + st["shadow_lane_steps"]) /
                                       max(1, st["wave_trips"] + st["shadow_wave_trips"]),
           "bvh": {"width": a.width, "leaf_size": a.leaf_size or (2 if a.width == 8 else 4),
                   "builder": a.builder, "upload_s": info["upload_seconds"], "nodes": int(info["n_nodes"]),
                   "depth": int(info["max_depth"]),
                   "build_s": info["build_seconds"], "device_gb": info["device_bytes"] / 1e9},
           "tuning": r.get_tuning()}

    if not a.no_cpu_baseline and not a.profile_only:
        log("cpu baseline for >= %.0fs ..." % a.cpu_seconds)
        out["cpu_baseline"] = cpu_baseline(v, t, fr, mc, cav, gpu_g, gpu_im, a.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    if a.calibrate:
        out["calibration_copy_ms"] = r.calibration_copy(1 << 30, 2)
    r.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
