#!/usr/bin/env python3
"""Writes the golden fixtures in this folder from the CPU oracle.

    python3 tests/golden/make_golden.py

Each fixture is a compressed .npz (no pickled objects) holding the frame's
inputs as plain numbers and the oracle's G-buffer (j3d/pixel.h:11-27 records
as raw 32-byte rows) and RGBA image.  tests/test_golden.py re-renders them
with the oracle (guards against oracle drift) and with the GPU path.  The
matcap fixture is the SHA-256 of each 512 x 512 table; the network fixture is
the oracle's output of a seeded network on seeded rays.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import j3d_amd as J  # noqa: E402
import oracle as O  # noqa: E402
from j3d_amd import ffi  # noqa: E402

W, H = 96, 64

CASES = {
    # name: (mesh, flags, light kind)
    "icosphere3_edges": ("icosphere", ffi.EDGES | ffi.SHADING, None),
    "icosphere3_shadow_point": ("icosphere", ffi.EDGES | ffi.SHADING | ffi.SHADOW, None),
    "torus_shadow_directional": ("torus", ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT, (0.3, 0.5)),
    "icosphere3_wireframe": ("icosphere", ffi.WIREFRAME | ffi.SHADING, None),
    "icosphere3_one_bit": ("icosphere", ffi.ONE_BIT | ffi.EDGES, None),
}


def case_inputs(name):
    mesh, flags, light = CASES[name]
    if mesh == "icosphere":
        v, t = J.icosphere(3, seed=5, amplitude=0.05)
        cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    else:
        v, t = J.torus(60, 34, seed=6)
        cs, ci = J.orbit_camera(3.0, 0.9, 0.45)
    cam = J.camera_with_vertical_fov(60.0)
    diag, pivot = J.scene_framing(v)
    ld = np.array([0.0, 0.0, 1.0], np.float32)
    if light is not None:
        th, ph = light
        ld = np.array([np.cos(ph) * np.cos(th), np.cos(ph) * np.sin(th), np.sin(ph)], np.float32)
        ld /= np.linalg.norm(ld)
    return v, t, cam, cs, ci, pivot, diag, flags, ld


def oracle_render(name):
    v, t, cam, cs, ci, pivot, diag, flags, ld = case_inputs(name)
    fr = J.make_frame(W, H, cam, cs, ci, pivot, diag, [0], flags=flags, light_dir=ld)
    mc, cav = O.matcap(0)
    g = np.zeros((H, W), O.PIXEL_DTYPE)
    g["object_id"] = 0xFFFFFFFF
    g, _ = O.update_canvas([O.Mesh(v, t, db_id=1 << 29)], fr, gbuf=g)
    im = O.canvas_to_image(g, fr, mc, cav)
    return g, im, fr


def mlp_case():
    from test_mlp_oracle import NORM, random_params, random_rays
    p = random_params(91)
    rays = random_rays(92, 256, NORM)
    return p, NORM, rays


def main():
    meta = {"width": W, "height": H, "cases": {}}
    for name in CASES:
        g, im, fr = oracle_render(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), gbuf=g.view(np.uint8).reshape(H, W, 32), rgba=im)
        meta["cases"][name] = {"flags": int(CASES[name][1]), "hits": int((g["object_id"] != 0xFFFFFFFF).sum())}
    hashes = {}
    for k in range(4):
        mc, cav = O.matcap(k)
        hashes[str(k)] = {"sha256": hashlib.sha256(np.ascontiguousarray(mc, "<u4").tobytes()).hexdigest(),
                          "cavity": int(cav)}
    meta["matcap_sha256"] = hashes
    p, nrm, rays = mlp_case()
    np.savez_compressed(os.path.join(HERE, "mlp_forward.npz"), y_fp32=O.mlp_forward(p, nrm, rays, bf16=False),
                        y_bf16=O.mlp_forward(p, nrm, rays, bf16=True))
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
