"""CPU oracle of the learned step-size proposer -- TEST INFRASTRUCTURE ONLY.

Nothing in the product path imports this module: it is the checker that the HIP network
kernels (red.cuda.closeencounter_amd/csrc/mlp.hip) and the step-feature kernel
(csrc/kern_events.hip, k_step_features) are compared with.  The proposer has no
counterpart in the reference (its only step control is the elementary controller of
red.cuda/src/red.cuda/int_rungekutta8.cu:776-777), so the oracle restates the network's
mathematics, not reference code, and is pinned by closed forms and finite differences
(tests/test_mlp_oracle.py).

Network: z = clip((x - mean) * inv_std, -8, 8) (FP32, as the kernels), then in float64 on the given
(already BF16-rounded) weights
    p1 = W1 z + b1, h1 = gelu(p1), p2 = W2 h1 + b2, h2 = gelu(p2), y = W3 . h2 + b3
with gelu(x) = x Phi(x) = x (1 + erf(x / sqrt 2)) / 2, and loss = mean (y - t)^2 or the
pinball loss of level tau.
Parameter vector layout: W1[64][12] b1[64] W2[64][64] b2[64] W3[64] b3.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NF, NH = 12, 64
SIZES = (("W1", NH * NF), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH), ("b3", 1))
NPARAM = sum(s for _, s in SIZES)


def offsets():
    out, k = {}, 0
    for name, s in SIZES:
        out[name] = (k, k + s)
        k += s
    return out


OFF = offsets()


def bf16_bits(w) -> np.ndarray:
    """IEEE binary32 -> bfloat16 bit patterns, round to nearest, ties to even; NaN stays NaN."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def bf16_round(w) -> np.ndarray:
    return bf16_value(bf16_bits(w))


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    g = {k: p[a:b] for k, (a, b) in OFF.items()}
    return g["W1"].reshape(NH, NF), g["b1"], g["W2"].reshape(NH, NH), g["b2"], g["W3"], g["b3"][0]


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, W3, [b3]])


Z_CLIP = np.float32(8.0)


def standardize(X, norm) -> np.ndarray:
    """FP32 clip((x - mean) * inv_std, -8, 8), one rounding per operation as in the kernels."""
    X = np.asarray(X, dtype=np.float32).reshape(-1, NF)
    mean = np.asarray(norm[:NF], dtype=np.float32)
    istd = np.asarray(norm[NF:], dtype=np.float32)
    return np.clip((X - mean) * istd, -Z_CLIP, Z_CLIP)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(params, Z):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    Z = np.asarray(Z, dtype=np.float64)
    p1 = Z @ W1.T + b1
    h1 = gelu(p1)
    p2 = h1 @ W2.T + b2
    h2 = gelu(p2)
    y = h2 @ W3 + b3
    return dict(p1=p1, h1=h1, p2=p2, h2=h2, y=y)


def loss_and_grad(params, Z, T, tau=None):
    """Loss and its gradient with respect to every parameter (float64): mean squared error,
    or with tau the pinball loss mean(max(tau (t - y), (tau - 1) (t - y)))."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    Z = np.asarray(Z, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    f = forward(params, Z)
    n = Z.shape[0]
    if tau is None:
        e = f["y"] - T
        loss = float(np.mean(e * e))
        dy = 2.0 * e / n
    else:
        e = T - f["y"]
        loss = float(np.mean(np.maximum(tau * e, (tau - 1.0) * e)))
        dy = np.where(e > 0, -tau, 1.0 - tau) / n
    gW3 = f["h2"].T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, W3) * gelu_grad(f["p2"])
    gW2 = d2.T @ f["h1"]
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * gelu_grad(f["p1"])
    gW1 = d1.T @ Z
    gb1 = d1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in FP32 in the kernels' operation order; returns (w, m, v)."""
    f = np.float32
    w = np.asarray(w, f); g = np.asarray(g, f); m = np.asarray(m, f); v = np.asarray(v, f)
    b1f, b2f, epsf, lrf = f(b1), f(b2), f(eps), f(lr)
    # bias corrections: the FP32 betas raised in double, rounded, subtracted in FP32
    c1 = f(1.0) - f(float(b1f) ** step)
    c2 = f(1.0) - f(float(b2f) ** step)
    mk = b1f * m + (f(1.0) - b1f) * g
    vk = b2f * v + (f(1.0) - b2f) * (g * g)
    wk = w - lrf * (mk / c1) / (np.sqrt(vk / c2) + epsf)
    return wk.astype(f), mk.astype(f), vk.astype(f)


def step_features(r, a, mass, ids, n_si, n_massive):
    """State features of a step start: (min d/R_H, max |a|, max |a - a_kep| / |a_kep|).

    Pairs: SI sinks 1..n_si-1 x massive sources 1..n_massive-1 and NSI sinks x SI sources
    1..n_si-1, both bodies active (id > 0), star (index 0) excluded; mutual Hill radius
    ((m_i + m_j) / (3 m_star))^(1/3) (h_i + h_j) / 2 with heliocentric distances h.
    Bodies: 1..n-1 with id > 0; a_kep = -m_star (r - r_star) / |r - r_star|^3.
    """
    r = np.asarray(r, np.float64); a = np.asarray(a, np.float64)
    mass = np.asarray(mass, np.float64); ids = np.asarray(ids)
    ms = mass[0]
    h = np.linalg.norm(r - r[0], axis=1)
    act = ids > 0
    best = np.inf
    pairs = []
    if n_si > 1:
        pairs.append((np.arange(1, n_si), np.arange(1, n_massive)))
        pairs.append((np.arange(n_si, n_massive), np.arange(1, n_si)))
    for I, J in pairs:
        if I.size == 0 or J.size == 0:
            continue
        d = np.linalg.norm(r[I][:, None, :] - r[J][None, :, :], axis=2)
        mm = mass[I][:, None] + mass[J][None, :]
        rh = np.cbrt(mm / (3.0 * ms)) * 0.5 * (h[I][:, None] + h[J][None, :])
        ok = act[I][:, None] & act[J][None, :] & (I[:, None] != J[None, :]) & (mm > 0)
        if ok.any():
            best = min(best, float(np.min((d / rh)[ok])))
    B = np.arange(1, r.shape[0])
    B = B[act[B]]
    amax = float(np.max(np.linalg.norm(a[B], axis=1))) if B.size else 0.0
    rel = r[B] - r[0]
    hb = np.linalg.norm(rel, axis=1)
    ak = -ms * rel / hb[:, None] ** 3
    pert = np.linalg.norm(a[B] - ak, axis=1) / np.linalg.norm(ak, axis=1)
    return best, amax, float(np.max(pert)) if B.size else 0.0
