"""Pins the CPU oracle of the learned proposal network (oracle/mlp_oracle.cpp) -- CPU only.

The network has no counterpart in the reference, so nothing there to compare
with: the oracle is pinned by closed forms of its definition (feature values
at angles with exact sines, a network wired to copy one feature to the output,
the loss of a constant network) and by central finite differences of its own
fp32 forward pass against its analytic gradients.  A slip in the feature
order, a transposed weight in the backward pass, a dropped loss factor or a
wrong bias gradient fails one of these.
"""
import math

import numpy as np
import pytest

import oracle as O

N_IN, N_H = 78, 64
P = 13506
OFF = {"W1": 0, "b1": 4992, "W2": 5056, "b2": 9152, "W3": 9216, "b3": 13312, "W4": 13376, "b4": 13504}


def bf16(x):
    """Round-to-nearest-even to bfloat16, as float32 (NaN-free inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    def fill(name, rows, cols, s):
        lim = s * math.sqrt(6.0 / cols)
        p[OFF[name]:OFF[name] + rows * cols] = rng.uniform(-lim, lim, rows * cols)
    fill("W1", N_H, N_IN, scale)
    fill("W2", N_H, N_H, scale)
    fill("W3", N_H, N_H, scale)
    fill("W4", 2, N_H, 0.5)
    for b in ("b1", "b2", "b3"):
        p[OFF[b]:OFF[b] + N_H] = rng.uniform(-0.1, 0.1, N_H)
    p[OFF["b4"]:OFF["b4"] + 2] = [1.0, 0.2]
    return p


def random_rays(seed, n, norm):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    d = rng.normal(size=(n, 3))
    r[:, 0:3] = norm[:3] + norm[3] * 2.0 * d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3:6] = rng.normal(size=(n, 3)) * rng.uniform(0.5, 2.0, (n, 1))
    r[:, 7] = np.finfo(np.float32).max
    return r


def random_labels(seed, n):
    rng = np.random.default_rng(seed)
    lab = np.zeros((n, 2), np.float32)
    hit = rng.random(n) < 0.6
    lab[:, 1] = hit
    lab[:, 0] = np.where(hit, rng.uniform(0.5, 3.0, n), 0.0)
    return lab


NORM = np.array([0.25, -0.5, 1.0, 2.0], np.float32)


def test_param_count():
    assert O.mlp_param_count() == P


def test_features_closed_form():
    # o = c + R (0.5, 0.5, 0) and d = (1, -1, 1): o is already the line's point
    # closest to c (a = 0 exactly), so x = (0.5, 0.5, 0, u) and the sines and
    # cosines of 2^k pi x have exact values
    o = NORM[:3] + NORM[3] * np.array([0.5, 0.5, 0.0], np.float32)
    d = np.array([1, -1, 1], np.float32)
    ray = np.zeros((1, 8), np.float32)
    ray[0, :3] = o
    ray[0, 3:6] = d
    line = O.mlp_line(ray, NORM)[0]
    assert line[0] == np.sqrt(np.float32(3.0)) and line[1] == 0.0
    f = O.mlp_features(ray, NORM)[0]
    xs = np.array([0.5, 0.5, 0.0] + list(d / np.sqrt(np.float32(3.0))), np.float32)
    assert np.array_equal(f[:6], xs) and np.array_equal(line[2:], xs)
    for i in range(3):
        for k in range(6):
            sin_v, cos_v = f[6 + 12 * i + 2 * k], f[7 + 12 * i + 2 * k]
            if i < 2:   # sin / cos(2^k pi / 2)
                exp_sin, exp_cos = (1.0, 0.0) if k == 0 else (0.0, -1.0 if k == 1 else 1.0)
            else:       # x = 0
                exp_sin, exp_cos = 0.0, 1.0
            assert abs(sin_v - exp_sin) < 1e-6 and abs(cos_v - exp_cos) < 1e-6, (i, k, sin_v, cos_v)
    # direction features: sin/cos of 2^k pi / sqrt(3) against float64
    for i in range(3, 6):
        for k in range(6):
            a = (2 ** k) * math.pi * float(xs[i])
            assert abs(f[6 + 12 * i + 2 * k] - math.sin(a)) < 2e-6
            assert abs(f[7 + 12 * i + 2 * k] - math.cos(a)) < 2e-6


def test_line_input_invariants():
    # every origin on one line gives the same input (up to fp32 rounding), a is
    # the distance to the line's point closest to c, and (x0..x2) is orthogonal to u
    rng = np.random.default_rng(8)
    rays = random_rays(9, 64, NORM)
    lam = rng.uniform(-3, 3, 64).astype(np.float32)
    moved = rays.copy()
    moved[:, :3] += lam[:, None] * rays[:, 3:6]
    a0, a1 = O.mlp_line(rays, NORM), O.mlp_line(moved, NORM)
    assert np.allclose(a0[:, 2:], a1[:, 2:], atol=2e-5)
    dn = np.linalg.norm(rays[:, 3:6].astype(np.float64), axis=1)
    assert np.allclose(a1[:, 1], a0[:, 1] - lam * dn, atol=2e-5)
    assert np.allclose(np.sum(a0[:, 2:5].astype(np.float64) * a0[:, 5:8], axis=1), 0, atol=1e-5)
    u = rays[:, 3:6] / dn[:, None]
    p = rays[:, :3] + a0[:, 1:2] * u
    assert np.allclose(np.linalg.norm(p - NORM[:3], axis=1) ** 2 + 0,
                       np.linalg.norm(rays[:, :3] - NORM[:3], axis=1) ** 2 - a0[:, 1] ** 2, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("bf", [False, True])
@pytest.mark.parametrize("idx", [0, 4, 6, 7, 29, 77])
def test_forward_copies_one_feature(idx, bf):
    # W1[0][idx] = W2[0][0] = W3[0][0] = W4[0][0] = 1, every other weight and bias 0,
    # b4 = (0, 0.75): y0 = relu(f[idx]) exactly, y1 = 0.75
    p = np.zeros(P, np.float32)
    p[OFF["W1"] + idx] = 1.0
    p[OFF["W2"]] = 1.0
    p[OFF["W3"]] = 1.0
    p[OFF["W4"]] = 1.0
    p[OFF["b4"] + 1] = 0.75
    rays = random_rays(3, 256, NORM)
    f = O.mlp_features(rays, NORM)[:, idx]
    y = O.mlp_forward(p, NORM, rays, bf16=bf)
    fe = bf16(f) if bf else f
    assert np.array_equal(y[:, 0], np.maximum(fe, 0).astype(np.float32))
    assert np.all(y[:, 1] == np.float32(0.75))


def test_loss_of_constant_network():
    # W4 = 0: y = b4 for every ray, so the loss is a closed form of the labels
    p = random_params(5)
    p[OFF["W4"]:OFF["W4"] + 128] = 0.0
    c0, c1 = 1.25, -0.5
    p[OFF["b4"]:OFF["b4"] + 2] = [c0, c1]
    rays = random_rays(6, 500, NORM)
    lab = random_labels(7, 500)
    g, loss = O.mlp_gradients(p, NORM, rays, lab, bf16=False)
    s, hit = lab[:, 0].astype(np.float64), lab[:, 1].astype(np.float64)
    sp = math.log1p(math.exp(c1))
    expect = np.mean(hit * 10.0 * (c0 - s) ** 2 + sp - hit * c1)
    assert abs(loss - expect) < 1e-5 * max(1.0, abs(expect))
    # bias gradients: d/dc0 = mean(20 hit (c0 - s)), d/dc1 = mean(sigmoid(c1) - hit)
    sig = 1.0 / (1.0 + math.exp(-c1))
    assert abs(g[OFF["b4"]] - np.mean(20.0 * hit * (c0 - s))) < 1e-5
    assert abs(g[OFF["b4"] + 1] - np.mean(sig - hit)) < 1e-6
    # nothing reaches the hidden layers through W4 = 0
    assert np.all(g[:OFF["W4"]] == 0)


def _loss(p, rays, lab):
    return O.mlp_gradients(p, NORM, rays, lab, bf16=False)[1]


def test_gradients_match_finite_differences():
    p = random_params(11)
    rays = random_rays(12, 96, NORM)
    lab = random_labels(13, 96)
    g, _ = O.mlp_gradients(p, NORM, rays, lab, bf16=False)
    rng = np.random.default_rng(14)
    groups = (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "W4"), ("W4", "b4"))
    eps = 2e-3
    errs = []
    for a, b in groups:
        idx = rng.choice(np.arange(OFF[a], OFF[b]), 24, replace=False)
        ok = 0
        for k in idx:
            pp, pm = p.copy(), p.copy()
            pp[k] += np.float32(eps)
            pm[k] -= np.float32(eps)
            fd = (_loss(pp, rays, lab) - _loss(pm, rays, lab)) / (float(pp[k]) - float(pm[k]))
            err = abs(fd - g[k])
            errs.append(err)
            ok += err <= 2e-3 + 2e-2 * abs(g[k])
        # a ReLU kink within +-eps of some ray's pre-activation spoils an isolated
        # difference (a bias moves every ray); a wrong formula fails the whole group
        assert ok >= 21, (a, ok)
    for k in (OFF["b4"], OFF["b4"] + 1):
        pp, pm = p.copy(), p.copy()
        pp[k] += np.float32(eps)
        pm[k] -= np.float32(eps)
        fd = (_loss(pp, rays, lab) - _loss(pm, rays, lab)) / (float(pp[k]) - float(pm[k]))
        assert abs(fd - g[k]) <= 1e-3 + 1e-3 * abs(g[k])
    assert np.median(errs) < 3e-4


def test_bf16_gradients_close_to_fp32():
    p = random_params(21)
    rays = random_rays(22, 512, NORM)
    lab = random_labels(23, 512)
    g32, l32 = O.mlp_gradients(p, NORM, rays, lab, bf16=False)
    g16, l16 = O.mlp_gradients(p, NORM, rays, lab, bf16=True)
    cos = float(np.dot(g32, g16) / (np.linalg.norm(g32) * np.linalg.norm(g16)))
    assert cos > 0.995
    assert abs(l32 - l16) < 0.02 * abs(l32)


def test_gradient_is_mean_over_rays():
    # the gradient of a batch is the mean of the per-half gradients (loss is a mean)
    p = random_params(31)
    rays = random_rays(32, 200, NORM)
    lab = random_labels(33, 200)
    g, loss = O.mlp_gradients(p, NORM, rays, lab, bf16=True)
    ga, la = O.mlp_gradients(p, NORM, rays[:100], lab[:100], bf16=True)
    gb, lb = O.mlp_gradients(p, NORM, rays[100:], lab[100:], bf16=True)
    assert np.allclose(g, 0.5 * (ga + gb), atol=1e-6, rtol=1e-5)
    assert abs(loss - 0.5 * (la + lb)) < 1e-5 * max(1.0, loss)
