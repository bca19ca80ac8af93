"""Pins the CPU restatement of the narrow-band network (oracle/mlp.py)."""
import struct

import numpy as np
import pytest

from oracle import mlp as M


def bf16_ref(x: float) -> float:
    """Independent round-to-nearest-even to bfloat16 via integer arithmetic."""
    u = struct.unpack("<I", struct.pack("<f", x))[0]
    hi, lo = u >> 16, u & 0xFFFF
    if lo > 0x8000 or (lo == 0x8000 and (hi & 1)):
        hi += 1
    return struct.unpack("<f", struct.pack("<I", (hi << 16) & 0xFFFFFFFF))[0]


def test_bf16_rounding():
    vals = [1.0, -1.0, 0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8,
            3.14159265, 1e-30, -7.25e5, 65504.0, 1.0e38]
    rng = np.random.default_rng(0)
    vals += list(rng.normal(0, 10, 2000).astype(np.float32))
    got = M.bf16(np.array(vals, np.float32))
    want = np.array([bf16_ref(float(np.float32(v))) for v in vals], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert M.bf16(np.array([1.0 + 2 ** -8], np.float32))[0] == 1.0  # tie to even
    assert np.isnan(M.bf16(np.array([np.nan], np.float32))[0])


def test_feature_closed_forms():
    p = np.array([[0.0, 0.5, 0.25]], np.float32)
    x = M.features(p)[0]
    assert x.shape == (M.FEAT_PAD,)
    assert np.array_equal(x[:3], p[0])
    # k = 0: w = pi -> sin(0)=0, sin(pi/2)=1, cos(pi/4)=sqrt(1/2)
    assert abs(x[3] - 0.0) < 1e-7 and abs(x[4] - 1.0) < 1e-6
    assert abs(x[8] - np.sqrt(0.5)) < 1e-6
    # k = 2: w = 2 pi -> cos(2 pi * 0.5) = -1
    assert abs(x[3 + 12 + 3 + 1] + 1.0) < 1e-6
    assert np.all(x[M.N_FEAT:] == 0.0)
    # c_k = 2^(k/2) / 2 revolutions per unit: 1/2 (k = 0), sqrt(2)/2 (k = 1), 64 (k = 14)
    assert M.REVS[0] == 0.5 and M.REVS[14] == 64.0
    assert M.REVS[1] == np.float32(np.sqrt(2.0) / 2.0)
    # k = 1: w = pi sqrt(2); at p = 1/sqrt(2) the angle is pi: sin = 0, cos = -1
    q = np.array([[np.float32(1.0 / np.sqrt(2.0)), 0.0, 0.0]], np.float32)
    y = M.features(q)[0]
    assert abs(y[3 + 6]) < 1e-6 and abs(y[6 + 6] + 1.0) < 1e-6


def test_parameter_layout():
    theta = M.init(3)
    assert len(theta) == M.N_PARAMS == 15617
    P = M.unpack(theta)
    assert P["W1"].shape == (64, 112) and np.all(P["W1"][:, 99:] == 0.0)
    assert np.array_equal(M.pack(P), theta)


def test_excess_closed_forms():
    t = np.array([0.3, 0.3, 0.3, 0.3, -0.2, -0.2, -0.2, 0.0, 0.0])
    f = np.array([0.1, 0.5, -0.4, 0.3, -0.1, -0.5, 0.25, 0.7, -0.7])
    want = np.array([0.0, 0.2, -0.4, 0.0, 0.0, -0.3, 0.25, 0.7, -0.7])
    assert np.allclose(M.excess(f, t), want, atol=1e-15)


def test_overshoot_one_is_least_squares():
    rng = np.random.default_rng(4)
    theta = M.init(4)
    p = rng.random((64, 3)).astype(np.float32)
    t = rng.normal(0, 0.05, 64).astype(np.float32)
    w = np.ones(64, np.float32)
    f = M.forward64(theta.astype(np.float64), p)
    assert M.loss64(theta, p, t, w, 1.0) == pytest.approx(np.sum((f - t) ** 2) / 128, rel=1e-12)
    # the extra term only counts errors beyond the interval [min(0,t), max(0,t)]
    extra = M.loss64(theta, p, t, w, 3.0) - M.loss64(theta, p, t, w, 1.0)
    e = np.where(t > 0, np.where(f > t, f - t, np.minimum(f, 0.0)),
                 np.where(f < t, f - t, np.maximum(f, 0.0)))
    assert extra == pytest.approx(2.0 * np.sum(e ** 2) / 128, rel=1e-9)


@pytest.mark.parametrize("seed,overshoot", [(0, 1.0), (1, 1.0), (2, 1.0), (3, 5.0), (4, 5.0)])
def test_float64_backprop_matches_finite_differences(seed, overshoot):
    rng = np.random.default_rng(seed)
    theta = M.init(seed).astype(np.float64)
    p = rng.random((24, 3)).astype(np.float32)
    t = rng.normal(0, 0.05, 24)
    w = rng.random(24) + 0.5
    g = M.grad64(theta, p, t, w, overshoot)
    idx = rng.choice(np.nonzero(np.abs(g) > 1e-6)[0], 60, replace=False)
    eps = 1e-6
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (M.loss64(tp, p, t, w, overshoot) - M.loss64(tm, p, t, w, overshoot)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_rounded_gradient_is_close_to_exact():
    rng = np.random.default_rng(7)
    theta = M.init(7)
    p = rng.random((256, 3)).astype(np.float32)
    t = rng.normal(0, 0.05, 256).astype(np.float32)
    w = np.ones(256, np.float32)
    loss, g = M.loss_and_grad(theta, p, t, w)
    g64 = M.grad64(theta.astype(np.float64), p, t.astype(np.float64), w.astype(np.float64))
    # bf16 operands: sums of signed terms lose a few per cent (most in W1,
    # whose inputs are the rounded features); the device is checked against
    # this restatement, which has its rounding points, at a tight tolerance.
    o = 0
    for name, shape in M.SIZES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], g64[o:o + k]
        tol = 8e-2 if name == "W1" else 3e-2
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b) + 1e-9, name
        o += k
    assert abs(loss - M.loss64(theta, p, t, w)) <= 2e-2 * loss


def test_adam_step_moves_against_the_gradient():
    theta = M.init(1)
    g = np.random.default_rng(1).normal(0, 1, M.N_PARAMS).astype(np.float32)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    t2, m2, v2 = M.adam(theta, m, v, g, 1, 1e-3)
    # first bias-corrected step is lr * sign(g) (up to eps)
    assert np.allclose(theta - t2, 1e-3 * np.sign(g), atol=1e-6)
