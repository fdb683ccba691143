"""GPU tests of the learned tmax proposal: the bf16-MFMA network kernels
against the CPU oracle (oracle/mlp_oracle.cpp, bf16 rounding emulated), Adam,
exactness of the bounded pass with learned, perfect and adversarial
proposals, checkpoints, and the gain of a trained network over plain
traversal."""
import math

import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import c0_scene, compare_gbuffers, compare_rgba, shadow_rays, torus_scene
from test_mlp_oracle import NORM, OFF, bf16, random_labels, random_params, random_rays

pytestmark = pytest.mark.gpu

# Tolerances of the GPU network against the oracle.  Both round the same
# operands to bf16; they differ only in summation (fp32 matrix-core sums vs
# double) and in sin/cos (hardware v_sin/v_cos vs libm), which moves an
# activation across a bf16 rounding boundary now and then (one bf16 ulp,
# 2^-8 relative, of one of 64 terms).
FWD_MAX = 2e-2      # max |y_gpu - y_cpu| / (1 + |y_cpu|)
FWD_MEAN = 5e-4     # mean |y_gpu - y_cpu|
GRAD_REL = 1e-2     # per parameter group: |g_gpu - g_cpu| / |g_cpu| (L2 norms)
GRAD_COS = 0.9999   # cosine between the full gradients


def test_forward_matches_oracle(renderer):
    p = random_params(41)
    rays = random_rays(42, 5000, NORM)  # not a multiple of 32: the last chunk is partial
    y = renderer.mlp_forward(p, NORM, rays)
    yo = O.mlp_forward(p, NORM, rays, bf16=True)
    assert np.all(np.isfinite(y))
    d = np.abs(y.astype(np.float64) - yo)
    rel = d / (1.0 + np.abs(yo))
    print("forward: max rel %.3g mean abs %.3g exact %.3f" % (rel.max(), d.mean(), float(np.mean(d == 0))))
    assert rel.max() < FWD_MAX
    assert d.mean() < FWD_MEAN


def test_forward_feature_wiring(renderer):
    # the closed-form network of the oracle test: y0 = relu(bf16(feature idx)), y1 = 0.75
    rays = random_rays(43, 1000, NORM)
    f = O.mlp_features(rays, NORM)
    for idx in (0, 3, 5, 6, 7, 17, 40, 77):
        p = np.zeros(ffi.MLP_PARAMS, np.float32)
        p[OFF["W1"] + idx] = 1.0
        p[OFF["W2"]] = 1.0
        p[OFF["W3"]] = 1.0
        p[OFF["W4"]] = 1.0
        p[OFF["b4"] + 1] = 0.75
        y = renderer.mlp_forward(p, NORM, rays)
        exp = np.maximum(bf16(f[:, idx]), 0)
        if idx < 6:
            # raw inputs: the same float32 arithmetic, then bf16
            assert np.array_equal(y[:, 0], exp), idx
        else:
            # sin/cos: hardware v_sin/v_cos vs libm, at most one bf16 step apart
            assert np.all(np.abs(y[:, 0] - exp) <= 2.0 ** -8 * np.maximum(np.abs(exp), 2.0 ** -6)), idx
        assert np.all(y[:, 1] == np.float32(0.75))


def test_gradients_match_oracle(renderer):
    p = random_params(51)
    rays = random_rays(52, 4000, NORM)
    lab = random_labels(53, 4000)
    g, loss = renderer.mlp_gradients(p, NORM, rays, lab)
    go, lo = O.mlp_gradients(p, NORM, rays, lab, bf16=True)
    assert np.all(np.isfinite(g))
    print("loss gpu %.6f cpu %.6f" % (loss, lo))
    assert abs(loss - lo) <= 1e-4 * max(1.0, abs(lo))
    bounds = [OFF[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")] + [ffi.MLP_PARAMS]
    for a, b in zip(bounds[:-1], bounds[1:]):
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-12)
        print("group %d..%d rel %.3g" % (a, b, rel))
        assert rel < GRAD_REL, (a, b, rel)
    cos = float(np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos > GRAD_COS


def test_gradients_of_tiny_batches(renderer):
    # 1 ray and 33 rays: masking of the unused lanes of a chunk.  With so few
    # rays one activation that rounds to the neighbouring bf16 value (a 2^-8
    # relative step) is not averaged away, hence the wider bounds.
    p = random_params(55)
    for n in (1, 33):
        rays = random_rays(56 + n, n, NORM)
        lab = random_labels(57 + n, n)
        g, loss = renderer.mlp_gradients(p, NORM, rays, lab)
        go, lo = O.mlp_gradients(p, NORM, rays, lab, bf16=True)
        print("n=%d loss gpu %.6f cpu %.6f grad rel %.3g" % (n, loss, lo, np.linalg.norm(g - go) / np.linalg.norm(go)))
        assert abs(loss - lo) <= 2e-3 * max(1.0, abs(lo))
        assert np.linalg.norm(g - go) <= 3e-2 * np.linalg.norm(go)


def test_adam_step(renderer):
    rng = np.random.default_rng(61)
    p = rng.normal(size=ffi.MLP_PARAMS).astype(np.float32)
    m = (0.1 * rng.normal(size=ffi.MLP_PARAMS)).astype(np.float32)
    v = (0.01 * rng.random(ffi.MLP_PARAMS)).astype(np.float32)
    g = rng.normal(size=ffi.MLP_PARAMS).astype(np.float32)
    step, lr = 7, 1e-3
    p2, m2, v2 = renderer.mlp_adam_step(p, m, v, g, step, lr)
    # Adam (Kingma & Ba 2015, algorithm 1), beta1 0.9, beta2 0.999, eps 1e-8
    me = 0.9 * m.astype(np.float64) + 0.1 * g
    ve = 0.999 * v.astype(np.float64) + 0.001 * g.astype(np.float64) ** 2
    pe = p - lr * (me / (1 - 0.9 ** step)) / (np.sqrt(ve / (1 - 0.999 ** step)) + 1e-8)
    assert np.allclose(m2, me, rtol=1e-5, atol=1e-7)
    assert np.allclose(v2, ve, rtol=1e-5, atol=1e-9)
    assert np.allclose(p2, pe, rtol=1e-5, atol=1e-6)


def _frame(slot, w, h, cam, cs, ci, pivot, diag, light, flags, margin=0.0, rect=None):
    return J.make_frame(w, h, cam, cs, ci, pivot, diag, [slot], flags=flags, light_dir=light, rect=rect,
                        margin=margin)


FLAGS = ffi.EDGES | ffi.SHADING | ffi.SHADOW | ffi.DIRECTIONAL_LIGHT


def _render_pair(r, fr_plain, fr_learned, tg):
    r.set_counting(True)
    g0, im0 = r.render(fr_plain, tg)
    s0 = r.stats()
    g1, im1 = r.render(fr_learned, tg)
    s1 = r.stats()
    r.set_counting(False)
    return g0, im0, s0, g1, im1, s1


def test_adversarial_proposals_are_exact(renderer):
    """The bounded pass + fallback gives the plain result for any proposal:
    too small, too large, negative, zero, inf, NaN, and wrong hit predictions."""
    r = renderer
    v, t, cam, cs, ci, pivot, diag = c0_scene(level=5, w=320, h=240)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 9))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    w, h = 320, 240
    tg = r.create_target(w, h)
    light = np.array([0.3, 0.4, 0.866], np.float32)
    plain = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS)
    keep = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS | ffi.LEARNED | ffi.KEEP_PROPOSAL)
    g0, im0 = r.render(plain, tg)
    depth = g0["depth"].astype(np.float32)
    hit = g0["object_id"] != 0xFFFFFFFF
    rng = np.random.default_rng(71)
    kinds = rng.integers(0, 8, size=(h, w))
    np.seterr(over="ignore")
    tb = np.where(kinds == 0, depth, 0)                                  # exact t
    tb = np.where(kinds == 1, depth * 0.5, tb)                           # too small: fallback
    tb = np.where(kinds == 2, depth * 3.0 + 1.0, tb)                     # too large: resolved
    tb = np.where(kinds == 3, -1.0, tb)
    tb = np.where(kinds == 4, np.inf, tb)
    tb = np.where(kinds == 5, np.nan, tb)
    tb = np.where(kinds == 6, rng.uniform(0, 10, size=(h, w)), tb)       # random
    tb = np.where(kinds == 7, np.nextafter(depth, np.float32(0)), tb)    # one ulp short
    logit = np.where(rng.random((h, w)) < 0.8, 1.0, -1.0)                # some predicted misses
    pairs = np.stack([tb, logit], axis=-1).astype(np.float32)
    r.write_proposal(tg, pairs)
    g0c, im0c, s0, g1, im1, s1 = _render_pair(r, plain, keep, tg)
    assert all(c == 0 for c in compare_gbuffers(g1, g0c).values())
    assert compare_rgba(im1, im0c)[0] == 0
    n_pred = int((logit > 0).sum())
    assert s1["learned_resolved"] + s1["learned_fallback"] == n_pred
    # exact bounds never fall back on a hit pixel; bounds one ulp short always do
    assert s1["learned_fallback"] >= int(((kinds == 7) & hit & (logit > 0)).sum())
    r.destroy_target(tg)


def test_perfect_proposal(renderer):
    """Bounds equal to the plain hit distance (margin 0): every hit pixel is
    resolved by the bounded pass, none falls back, and no more nodes are visited."""
    r = renderer
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(600, 340, seed=3)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 10))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    w, h = 640, 360
    tg = r.create_target(w, h)
    plain = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS)
    keep = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS | ffi.LEARNED | ffi.KEEP_PROPOSAL)
    g0, _ = r.render(plain, tg)
    hit = g0["object_id"] != 0xFFFFFFFF
    pairs = np.stack([np.where(hit, g0["depth"], 0), np.where(hit, 1.0, -1.0)], axis=-1).astype(np.float32)
    r.write_proposal(tg, pairs)
    g0c, im0c, s0, g1, im1, s1 = _render_pair(r, plain, keep, tg)
    assert all(c == 0 for c in compare_gbuffers(g1, g0c).values())
    assert s1["learned_fallback"] == 0 and s1["learned_resolved"] == int(hit.sum())
    print("perfect proposal: primary node visits %d -> %d (%.3f), tri tests %d -> %d" % (
        s0["node_visits"], s1["node_visits"], s1["node_visits"] / s0["node_visits"], s0["tri_tests"],
        s1["tri_tests"]))
    assert s1["node_visits"] <= s0["node_visits"]
    r.destroy_target(tg)


def test_perfect_shadow_bound(renderer):
    """Shadow rays rebuilt on the host from the G-buffer (the hit point from
    the triangle's vertices and the stored barycentrics) reproduce the frame's
    shadow bits through the any-hit query; inside a bound around each
    occluded ray's nearest occluder (closest-hit query) every occluded ray is
    still found, with fewer node visits -- the most a learned shadow bound
    could save.  Prints the counts."""
    r = renderer
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(600, 340, seed=3)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 10))
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    w, h = 640, 360
    tg = r.create_target(w, h)
    try:
        fr = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS)
        r.set_counting(True)
        g, _ = r.render(fr, tg)
        st = r.stats()
        rays, idx = shadow_rays(v, t, g, fr.light_dir)
        occ = r.any_hit([slot], rays).astype(bool)
        full = r.stats()["node_visits"]
        assert len(rays) == st["hits"] and occ.sum() == st["shadowed"]
        assert np.array_equal(occ, (g.reshape(-1)["mark"][idx] & 1) == 1)
        tocc = r.trace_rays([slot], rays[occ])["t"]
        r.any_hit([slot], rays[occ])
        occ_full = r.stats()["node_visits"]
        b = rays[occ].copy()
        b[:, 6] = np.maximum(np.float32(1e-3), tocc * np.float32(1 - 1e-4))
        b[:, 7] = tocc * np.float32(1 + 1e-4)
        found = r.any_hit([slot], b).astype(bool)
        occ_bounded = r.stats()["node_visits"]
        r.set_counting(False)
        assert found.all()
        # any-hit queries count their work like the frame's shadow phase (the
        # order of a cooperative tail may differ, so the totals only nearly agree)
        assert abs(full - st["shadow_node_visits"]) <= 0.1 * st["shadow_node_visits"]
        assert 0 < occ_bounded <= occ_full < full
        print("shadow rays %d, occluded %d; node visits: frame %d, query %d; occluded %d -> %d with a perfect bound"
              % (len(rays), occ.sum(), st["shadow_node_visits"], full, occ_full, occ_bounded))
    finally:
        r.set_counting(False)
        r.destroy_target(tg)
        r.remove_mesh(slot)


@pytest.fixture(scope="module")
def trained(renderer):
    """A torus of 408,000 triangles with a network trained on 2^22 rays."""
    r = renderer
    v, t, cam, cs, ci, pivot, diag, light = torus_scene(600, 340, seed=4)
    slot = r.upload_mesh(v, t, db_id=J.make_db_id(1, 11))
    rep = r.train_proposal(slot, n_rays=1 << 22, epochs=4, batch=8192, seed=5)
    print("training:", rep)
    return slot, (cam, cs, ci, pivot, diag, light), rep


def test_training_reduces_loss(trained):
    _, _, rep = trained
    assert rep["steps"] == 4 * (1 << 22) // 8192
    assert math.isfinite(rep["loss_last"]) and rep["loss_last"] < 0.5 * rep["loss_first"]
    # better than the constant predictor, and some held-out hits resolved at margin 0.1
    assert rep["val_hit_accuracy"] > max(rep["hit_fraction"], 1.0 - rep["hit_fraction"])
    assert rep["val_resolved"][3] > 0.05


def test_learned_frames_exact(renderer, trained):
    r = renderer
    slot, (cam, cs, ci, pivot, diag, light), _ = trained
    mc, cav = J.make_matcap(2)
    r.set_matcap(mc, cav)
    w, h = 480, 270
    tg = r.create_target(w, h)
    for k, (az, el) in enumerate([(0.3, 0.4), (2.0, 0.7), (4.0, 0.2)]):
        cs_k, ci_k = J.orbit_camera(3.0, az, el)
        for margin, rect in ((0.0, None), (0.05, None), (0.02, (17, 30, 300, 200))):
            plain = _frame(slot, w, h, cam, cs_k, ci_k, pivot, diag, light, FLAGS, rect=rect)
            learned = _frame(slot, w, h, cam, cs_k, ci_k, pivot, diag, light, FLAGS | ffi.LEARNED, margin=margin,
                             rect=rect)
            g0, im0, s0, g1, im1, s1 = _render_pair(r, plain, learned, tg)
            assert all(c == 0 for c in compare_gbuffers(g1, g0).values()), (k, margin, rect)
            assert compare_rgba(im1, im0)[0] == 0
            assert s1["learned_resolved"] > 0
            print("cam %d margin %.2f rect %s: resolved %d fallback %d hits %d, nodes %d -> %d" % (
                k, margin, rect, s1["learned_resolved"], s1["learned_fallback"], s1["hits"], s0["node_visits"],
                s1["node_visits"]))
    r.destroy_target(tg)


def test_checkpoint_roundtrip(renderer, trained, tmp_path):
    r = renderer
    slot, (cam, cs, ci, pivot, diag, light), _ = trained
    p, nrm = r.get_proposal(slot)
    path = str(tmp_path / "torus.j3dgmlp")
    r.save_proposal(slot, path)
    p2, n2 = J.read_proposal(path)
    assert np.array_equal(p, p2) and np.array_equal(nrm, n2)
    w, h = 256, 144
    tg = r.create_target(w, h)
    fr = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS | ffi.LEARNED, margin=0.03)
    r.render(fr, tg)
    a = r.read_proposal(tg)
    r.set_proposal(slot, None, None)
    r.load_proposal(slot, path)
    r.render(fr, tg)
    b = r.read_proposal(tg)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    r.destroy_target(tg)


def test_learned_gain(renderer, trained):
    """What the trained network gains over plain traversal on the same frames:
    primary node visits and time per margin (printed), with the G-buffers
    identical.  Asserted: at margin 0.1 the bounded pass resolves at least 90%
    of the hit pixels and the bounded + fallback traversal visits at most 2%
    more nodes than plain.  It does not visit fewer: front-to-back traversal
    already stops near the closest hit, and even exact bounds save about 0.1%
    of the visits (test_perfect_proposal), so the network's cost is not repaid
    on these frames."""
    r = renderer
    slot, (cam, cs, ci, pivot, diag, light), _ = trained
    mc, cav = J.make_matcap(0)
    r.set_matcap(mc, cav)
    w, h = 1280, 720
    tg = r.create_target(w, h)
    plain = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS)
    rows = {}
    for margin in (0.0, 0.02, 0.05, 0.1):
        learned = _frame(slot, w, h, cam, cs, ci, pivot, diag, light, FLAGS | ffi.LEARNED, margin=margin)
        g0, im0, s0, g1, im1, s1 = _render_pair(r, plain, learned, tg)
        assert all(c == 0 for c in compare_gbuffers(g1, g0).values())
        times = {}
        for name, fr in (("plain", plain), ("learned", learned)):
            for _ in range(3):
                r.render(fr, tg, readback=False)
            r.synchronize()
            r.reset_timings()
            for _ in range(10):
                r.render(fr, tg, readback=False)
            r.synchronize()
            tm = r.timings()
            times[name] = float(np.median(np.array(tm["trace_ms"]) + np.array(tm["mlp_ms"])))
        rows[margin] = (s0, s1)
        print("margin %.2f: primary nodes %d -> %d (%.3f), resolved %d fallback %d of %d hits, primary trace(+mlp) "
              "ms %.3f -> %.3f" % (margin, s0["node_visits"], s1["node_visits"], s1["node_visits"] / s0["node_visits"],
                                   s1["learned_resolved"], s1["learned_fallback"], s1["hits"], times["plain"],
                                   times["learned"]))
    s0, s1 = rows[0.1]
    assert s1["learned_resolved"] >= 0.9 * s1["hits"]
    assert s1["node_visits"] <= 1.02 * s0["node_visits"]
    r.destroy_target(tg)
