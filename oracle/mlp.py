"""CPU restatement of the narrow-band network — TEST INFRASTRUCTURE ONLY.

The learned narrow band has no counterpart in the reference crate: it is
the library's own method (a Fourier-feature MLP that predicts the truncated
SDF at 4^3-brick centres).  This module restates the network the GPU runs,
with the same rounding points, so the GPU forward pass and gradients can be
checked against it; its own backward pass is pinned by finite differences
in float64 (tests/test_oracle_mlp.py).  Only the tests, the smoke check and
the benchmark's CPU-baseline leg use it.

Network (parameter vector layout identical to the device's):
  features  x = [p, sin(2 pi r_k), cos(2 pi r_k)], r_k = c_k p rounded to
            float32, c_k = 2^(k/2) / 2 (so 2 pi c_k = pi 2^(k/2)), k = 0..15,
            99 values in the order p_xyz, then per k: sin_xyz, cos_xyz;
            padded with zeros to 112.  The angle is kept in revolutions,
            r_k, whose fractional part is exact in float32: the device's
            sine and cosine take revolutions.
  layers    h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)
            (64 units each), f = w4 . h3 + b4
  rounding  x, W1..W3, h1..h3 in bfloat16 (round to nearest even);
            matrix products accumulate in float32; biases, w4, b4 float32
  loss      L = sum_s wgt_s * (r_s^2 + (a - 1) e_s^2) / (2 * n) over a batch of
            n, r = f - t, e = f - clip(f, min(0, t), max(0, t)) (the error
            beyond the interval between 0 and t), a the overshoot weight
  backward  g_s = wgt_s (r_s + (a - 1) e_s) / n; dz3 = bf16(g w4 * [h3 > 0]);
            dz2 = bf16((W3^T dz3) * [h2 > 0]); dz1 = bf16((W2^T dz2) * [h1 > 0]);
            dW_l = sum_s dz_l h_(l-1)^T (bf16 operands, float32 sums),
            db_l = sum_s dz_l, dw4 = sum_s g h3, db4 = sum_s g
"""
from __future__ import annotations

import numpy as np

HIDDEN = 64
N_FREQ = 16
N_FEAT = 3 + 6 * N_FREQ  # 99
FEAT_PAD = 112
# revolutions per unit: 2^(k/2) / 2 in float32 (exact for even k)
REVS = (2.0 ** (np.arange(N_FREQ) / 2.0) / 2.0).astype(np.float32)

# parameter vector layout
SIZES = [("W1", (HIDDEN, FEAT_PAD)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)),
         ("b2", (HIDDEN,)), ("W3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("w4", (HIDDEN,)),
         ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = theta[o:o + k].reshape(shape)
        o += k
    return out


def pack(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[n], np.float32).reshape(-1) for n, _ in SIZES])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), x, out)


def features(p: np.ndarray) -> np.ndarray:
    """(n, 3) float32 -> (n, FEAT_PAD) float32."""
    p = np.asarray(p, np.float32)
    n = len(p)
    x = np.zeros((n, FEAT_PAD), np.float32)
    x[:, 0:3] = p
    for k in range(N_FREQ):
        r = (REVS[k] * p).astype(np.float32).astype(np.float64)
        x[:, 3 + 6 * k:6 + 6 * k] = np.sin(2.0 * np.pi * r)
        x[:, 6 + 6 * k:9 + 6 * k] = np.cos(2.0 * np.pi * r)
    return x


def init(seed: int = 0) -> np.ndarray:
    """He-uniform initialisation (the device uses its own generator)."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, fan_in in (("W1", N_FEAT), ("W2", HIDDEN), ("W3", HIDDEN)):
        lim = np.sqrt(6.0 / fan_in)
        shape = (HIDDEN, FEAT_PAD) if name == "W1" else (HIDDEN, HIDDEN)
        W = rng.uniform(-lim, lim, shape).astype(np.float32)
        if name == "W1":
            W[:, N_FEAT:] = 0.0
        P[name] = W
    for b in ("b1", "b2", "b3"):
        P[b] = np.zeros(HIDDEN, np.float32)
    P["w4"] = rng.uniform(-np.sqrt(6.0 / HIDDEN), np.sqrt(6.0 / HIDDEN), HIDDEN).astype(np.float32)
    P["b4"] = np.zeros(1, np.float32)
    return pack(P)


def _mm(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """float32 result of a float64 product of (already rounded) operands."""
    return (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32)


def forward(theta: np.ndarray, p: np.ndarray, keep: bool = False):
    P = unpack(np.asarray(theta, np.float32))
    x = bf16(features(p))
    z1 = _mm(x, bf16(P["W1"]).T) + P["b1"]
    h1 = bf16(np.maximum(z1, 0.0))
    z2 = _mm(h1, bf16(P["W2"]).T) + P["b2"]
    h2 = bf16(np.maximum(z2, 0.0))
    z3 = _mm(h2, bf16(P["W3"]).T) + P["b3"]
    h3 = bf16(np.maximum(z3, 0.0))
    f = (h3.astype(np.float64) @ P["w4"].astype(np.float64)).astype(np.float32) + P["b4"][0]
    if keep:
        return f, (x, h1, h2, h3)
    return f


def excess(f: np.ndarray, t: np.ndarray) -> np.ndarray:
    """The part of f - t beyond the interval between 0 and t (0 inside it)."""
    return f - np.clip(f, np.minimum(t, 0.0), np.maximum(t, 0.0))


def loss_and_grad(theta: np.ndarray, p: np.ndarray, t: np.ndarray, wgt: np.ndarray,
                  overshoot: float = 1.0):
    """Loss and gradient with the device's rounding points (see module doc)."""
    P = unpack(np.asarray(theta, np.float32))
    n = len(p)
    f, (x, h1, h2, h3) = forward(theta, p, keep=True)
    t = np.asarray(t, np.float32)
    r = (f - t).astype(np.float32)
    e = excess(f, t).astype(np.float32)
    ow = np.float32(overshoot) - np.float32(1.0)
    loss = float(np.sum(wgt.astype(np.float64) * (r.astype(np.float64) ** 2 +
                                                  float(ow) * e.astype(np.float64) ** 2)) / (2 * n))
    g = (wgt * (r + ow * e) * (np.float32(1.0) / np.float32(n))).astype(np.float32)
    dz3 = bf16(g[:, None] * P["w4"][None, :] * (h3 > 0))
    dz2 = bf16(_mm(dz3, bf16(P["W3"])) * (h2 > 0))
    dz1 = bf16(_mm(dz2, bf16(P["W2"])) * (h1 > 0))
    G = {
        "W1": _mm(dz1.T, x), "b1": dz1.astype(np.float64).sum(0).astype(np.float32),
        "W2": _mm(dz2.T, h1), "b2": dz2.astype(np.float64).sum(0).astype(np.float32),
        "W3": _mm(dz3.T, h2), "b3": dz3.astype(np.float64).sum(0).astype(np.float32),
        "w4": _mm(g[None, :], h3)[0], "b4": np.array([g.astype(np.float64).sum()], np.float32),
    }
    G["W1"][:, N_FEAT:] = 0.0
    return loss, pack(G)


def forward64(theta: np.ndarray, p: np.ndarray) -> np.ndarray:
    """The same network in float64 without bfloat16 rounding (for finite differences)."""
    P = {k: v.astype(np.float64) for k, v in unpack(np.asarray(theta, np.float64)).items()}
    x = features(p).astype(np.float64)
    h1 = np.maximum(x @ P["W1"].T + P["b1"], 0.0)
    h2 = np.maximum(h1 @ P["W2"].T + P["b2"], 0.0)
    h3 = np.maximum(h2 @ P["W3"].T + P["b3"], 0.0)
    return h3 @ P["w4"] + P["b4"][0]


def grad64(theta: np.ndarray, p: np.ndarray, t: np.ndarray, wgt: np.ndarray,
           overshoot: float = 1.0) -> np.ndarray:
    """Exact float64 gradient of the same loss (no rounding), by backpropagation."""
    P = {k: v.astype(np.float64) for k, v in unpack(np.asarray(theta, np.float64)).items()}
    n = len(p)
    x = features(p).astype(np.float64)
    z1 = x @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ P["W3"].T + P["b3"]
    h3 = np.maximum(z3, 0.0)
    f = h3 @ P["w4"] + P["b4"][0]
    g = wgt * ((f - t) + (overshoot - 1.0) * excess(f, t)) / n
    dz3 = g[:, None] * P["w4"][None, :] * (z3 > 0)
    dz2 = (dz3 @ P["W3"]) * (z2 > 0)
    dz1 = (dz2 @ P["W2"]) * (z1 > 0)
    G = {"W1": dz1.T @ x, "b1": dz1.sum(0), "W2": dz2.T @ h1, "b2": dz2.sum(0),
         "W3": dz3.T @ h2, "b3": dz3.sum(0), "w4": g @ h3, "b4": np.array([g.sum()])}
    G["W1"][:, N_FEAT:] = 0.0
    return np.concatenate([G[nm].reshape(-1) for nm, _ in SIZES])


def loss64(theta: np.ndarray, p: np.ndarray, t: np.ndarray, wgt: np.ndarray,
           overshoot: float = 1.0) -> float:
    f = forward64(theta, p)
    return float(np.sum(wgt * ((f - t) ** 2 + (overshoot - 1.0) * excess(f, t) ** 2)) /
                 (2 * len(p)))


def adam(theta, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32, as the device: bias-corrected moments."""
    theta = theta.astype(np.float32).copy()
    m = (np.float32(beta1) * m + np.float32(1 - beta1) * grad).astype(np.float32)
    v = (np.float32(beta2) * v + np.float32(1 - beta2) * grad * grad).astype(np.float32)
    c1 = np.float32(1 - beta1 ** step)
    c2 = np.float32(1 - beta2 ** step)
    theta -= np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))
    return theta, m, v
