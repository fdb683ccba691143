"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py from
the CPU oracle): the oracle must still produce them bit for bit (guards
against oracle drift), the product's host matcap tables must hash to the
committed SHA-256, and the GPU path must reproduce the frames (G-buffer
bit-exact, RGBA within 1) and the network outputs (within the stated
tolerance)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import compare_gbuffers, compare_rgba

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden as MG  # noqa: E402

META = json.load(open(os.path.join(HERE, "golden.json")))


def load_case(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    g = np.ascontiguousarray(z["gbuf"]).view(O.PIXEL_DTYPE).reshape(MG.H, MG.W)
    return g, z["rgba"]


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_oracle_reproduces_fixture(name):
    g0, im0 = load_case(name)
    g, im, _ = MG.oracle_render(name)
    assert np.array_equal(g.view(np.uint8), g0.view(np.uint8))
    assert np.array_equal(im, im0)
    assert int((g["object_id"] != 0xFFFFFFFF).sum()) == META["cases"][name]["hits"]


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_matcap_tables_hash(kind):
    exp = META["matcap_sha256"][str(kind)]
    mc, cav = O.matcap(kind)
    assert hashlib.sha256(np.ascontiguousarray(mc, "<u4").tobytes()).hexdigest() == exp["sha256"]
    # the product's own host tables (j3dg_make_matcap) hash the same
    pm, pcav = J.make_matcap(kind)
    assert hashlib.sha256(np.ascontiguousarray(pm, "<u4").tobytes()).hexdigest() == exp["sha256"]
    assert int(cav) == int(pcav) == exp["cavity"]


def test_oracle_network_reproduces_fixture():
    z = np.load(os.path.join(HERE, "mlp_forward.npz"), allow_pickle=False)
    p, nrm, rays = MG.mlp_case()
    assert np.array_equal(O.mlp_forward(p, nrm, rays, bf16=False), z["y_fp32"])
    assert np.array_equal(O.mlp_forward(p, nrm, rays, bf16=True), z["y_bf16"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_gpu_reproduces_fixture(renderer, name):
    g0, im0 = load_case(name)
    v, t, cam, cs, ci, pivot, diag, flags, ld = MG.case_inputs(name)
    slot = renderer.upload_mesh(v, t, db_id=1 << 29)
    mc, cav = J.make_matcap(0)
    renderer.set_matcap(mc, cav)
    tg = renderer.create_target(MG.W, MG.H)
    fr = J.make_frame(MG.W, MG.H, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=ld)
    g, im = renderer.render(fr, tg)
    renderer.destroy_target(tg)
    renderer.remove_mesh(slot)
    assert all(c == 0 for c in compare_gbuffers(g, g0).values()), compare_gbuffers(g, g0)
    assert compare_rgba(im, im0)[0] <= 1


@pytest.mark.gpu
def test_gpu_network_matches_fixture(renderer):
    from test_gpu_mlp import FWD_MAX
    z = np.load(os.path.join(HERE, "mlp_forward.npz"), allow_pickle=False)
    p, nrm, rays = MG.mlp_case()
    y = renderer.mlp_forward(p, nrm, rays)
    rel = np.abs(y.astype(np.float64) - z["y_bf16"]) / (1.0 + np.abs(z["y_bf16"]))
    assert rel.max() < FWD_MAX
