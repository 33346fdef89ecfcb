"""Pins the proposer network's CPU oracle (oracle/mlp_oracle.py) by what the mathematics
fixes: BF16 rounding on hand-checked bit patterns, GELU against the normal CDF from an
independent library routine, every gradient entry against central finite differences,
a closed form of the output, and the Adam update on a one-parameter case worked by hand.
The step-feature restatement is pinned on a three-body configuration whose values are
computed by hand."""
import math

import numpy as np
import pytest
from scipy.stats import norm as normal

import mlp_oracle as mo


def test_bf16_rounding_known_patterns():
    f = np.float32
    cases = {
        1.0: 0x3F80,
        -2.0: 0xC000,
        float(f(1.0) + f(2.0 ** -8)): 0x3F80,      # exact tie: stays on the even pattern
        float(f(1.0) + f(3 * 2.0 ** -8)): 0x3F82,  # tie above an odd pattern: rounds up to even
        float(f(1.0) + f(2.0 ** -8) + f(2.0 ** -20)): 0x3F81,   # just above the tie: up
        float(np.inf): 0x7F80,
        float(np.finfo(np.float32).max): 0x7F80,   # rounds past the largest BF16 to inf
        0.0: 0x0000,
    }
    for x, bits in cases.items():
        got = int(mo.bf16_bits(np.array([x], np.float32))[0])
        assert got == bits, (x, hex(got), hex(bits))
    nan = mo.bf16_bits(np.array([np.nan], np.float32))[0]
    assert (int(nan) & 0x7F80) == 0x7F80 and (int(nan) & 0x7F) != 0


def test_gelu_against_normal_cdf_and_derivative():
    x = np.linspace(-6, 6, 241)
    np.testing.assert_allclose(mo.gelu(x), x * normal.cdf(x), rtol=1e-13, atol=1e-15)
    h = 1e-5
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(x), fd, rtol=1e-8, atol=1e-9)
    assert mo.gelu(0.0) == 0.0 and mo.gelu_grad(0.0) == 0.5


def random_net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, mo.NPARAM)
    return mo.bf16_round(p.astype(np.float32)).astype(np.float64)


def test_gradient_matches_finite_differences():
    p = random_net(1)
    rng = np.random.default_rng(2)
    Z = rng.normal(size=(7, mo.NF))
    T = rng.normal(size=7)
    loss, g = mo.loss_and_grad(p, Z, T)
    assert loss == pytest.approx(np.mean((mo.forward(p, Z)["y"] - T) ** 2))
    # every b3/W3/b2/b1 entry and a spread of W1/W2 entries
    idx = list(range(mo.OFF["b1"][0], mo.NPARAM))
    idx += list(range(0, mo.OFF["W1"][1], 5))
    idx = sorted(set(idx) - set(range(*mo.OFF["W2"]))) + list(range(mo.OFF["W2"][0], mo.OFF["W2"][1], 37))
    h = 1e-6
    for k in idx:
        q = p.copy(); q[k] += h
        lp = np.mean((mo.forward(q, Z)["y"] - T) ** 2)
        q[k] -= 2 * h
        lm = np.mean((mo.forward(q, Z)["y"] - T) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_pinball_gradient_matches_finite_differences():
    p = random_net(4)
    rng = np.random.default_rng(5)
    Z = rng.normal(size=(9, mo.NF))
    T = rng.normal(size=9)
    tau = 0.1
    loss, g = mo.loss_and_grad(p, Z, T, tau=tau)
    e = T - mo.forward(p, Z)["y"]
    assert loss == pytest.approx(np.mean(np.where(e > 0, tau * e, (tau - 1) * e)))
    h = 1e-7
    for k in list(range(mo.OFF["W3"][0], mo.NPARAM)) + list(range(0, mo.OFF["W2"][1], 97)):
        q = p.copy(); q[k] += h
        lp = mo.loss_and_grad(q, Z, T, tau=tau)[0]
        q[k] -= 2 * h
        lm = mo.loss_and_grad(q, Z, T, tau=tau)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_pinball_minimiser_is_the_quantile():
    # the constant minimising the mean pinball loss over a sample is its tau-quantile
    rng = np.random.default_rng(6)
    t = rng.normal(size=2001)
    grid = np.linspace(-3, 3, 6001)
    for tau in (0.05, 0.5, 0.9):
        L = [np.mean(np.maximum(tau * (t - c), (tau - 1) * (t - c))) for c in grid]
        assert abs(grid[int(np.argmin(L))] - np.quantile(t, tau, method="inverted_cdf")) < 2e-3


def test_output_closed_form_without_hidden_weights():
    rng = np.random.default_rng(3)
    W1 = np.zeros((mo.NH, mo.NF)); W2 = np.zeros((mo.NH, mo.NH))
    b1 = rng.normal(size=mo.NH); b2 = rng.normal(size=mo.NH); W3 = rng.normal(size=mo.NH); b3 = 0.25
    p = mo.pack(W1, b1, W2, b2, W3, b3)
    Z = rng.normal(size=(5, mo.NF))
    expect = b3 + sum(W3[u] * b2[u] * normal.cdf(b2[u]) for u in range(mo.NH))
    np.testing.assert_allclose(mo.forward(p, Z)["y"], expect, rtol=1e-13)


def test_adam_first_step_moves_by_lr():
    # at step 1 the bias-corrected update is lr * g / (|g| + eps): a step of size ~lr against g
    w = np.array([1.0, -1.0, 0.5], np.float32)
    g = np.array([0.3, -2.0, 0.0], np.float32)
    z = np.zeros(3, np.float32)
    w1, m1, v1 = mo.adam_step(w, g, z, z, 1, 0.01)
    np.testing.assert_allclose(w1, [0.99, -0.99, 0.5], rtol=0, atol=2e-7)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=5e-5)   # 1 - 0.999f is 0.99998713e-3


def test_standardize_is_fp32():
    X = np.full((1, mo.NF), 3.0, np.float32)
    nrm = np.concatenate([np.full(mo.NF, 1.0), np.full(mo.NF, 0.5)]).astype(np.float32)
    z = mo.standardize(X, nrm)
    assert z.dtype == np.float32 and np.all(z == 1.0)
    far = mo.standardize(np.array([[1e6] * 6 + [-1e6] * 6], np.float32), nrm)
    assert np.all(far[0, :6] == 8.0) and np.all(far[0, 6:] == -8.0)


def test_step_features_by_hand():
    # star at the origin, one protoplanet (SI) at 1 AU, one planetesimal (NSI) at 1.01 AU
    ms, m1, m2 = 1.0, 3e-6, 1e-9
    r = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.01, 0, 0]])
    mass = np.array([ms, m1, m2])
    ids = np.array([1, 2, 3])
    a = np.array([[0.0, 0, 0], [-1.0 + m2 / 0.01 ** 2, 0, 0], [-1 / 1.01 ** 2 - m1 / 0.01 ** 2, 0, 0]])
    q, amax, pert = mo.step_features(r, a, mass, ids, n_si=2, n_massive=3)
    rh = ((m1 + m2) / 3.0) ** (1 / 3) * 0.5 * (1.0 + 1.01)
    assert q == pytest.approx(0.01 / rh, rel=1e-12)
    assert amax == pytest.approx(1 / 1.01 ** 2 + m1 / 0.01 ** 2, rel=1e-12)
    pert_2 = (m1 / 0.01 ** 2) / (1 / 1.01 ** 2)
    assert pert == pytest.approx(pert_2, rel=1e-12)
    # an inactive body leaves the pair and body sets
    q2, _, _ = mo.step_features(r, a, mass, np.array([1, 2, -3]), n_si=2, n_massive=3)
    assert q2 == math.inf
