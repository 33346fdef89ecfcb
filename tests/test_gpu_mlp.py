"""The proposer network's HIP kernels against the CPU oracle (oracle/mlp_oracle.py):
forward output and every gradient entry at the BF16 weights (FP32 accumulation against
float64: tolerances stated per test), the Adam update and the BF16 image of the master
weights (bit-exact), the in-loop proposal kernel against the batch kernel (bit-exact), and
the device step features against their restatement.  Recording step histories must not
change a trajectory."""
import numpy as np
import pytest

import mlp_oracle as mo
from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def make_net(native_mod, seed=11, scale=0.25):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, mo.NPARAM).astype(np.float32)
    mean = rng.normal(0.0, 1.0, mo.NF).astype(np.float32)
    istd = rng.uniform(0.5, 2.0, mo.NF).astype(np.float32)
    net = native_mod.Mlp(p)
    net.set_normalizer(mean, istd)
    return net, p, np.concatenate([mean, istd])


def test_forward_matches_oracle(native_mod):
    net, p, nrm = make_net(native_mod)
    w, b, nrm_dev = net.params()
    assert_bitwise(w, p, "master")
    assert np.array_equal(b, mo.bf16_bits(p)), "BF16 image"
    assert_bitwise(nrm_dev, nrm, "normaliser")
    X = np.random.default_rng(12).normal(size=(1000, mo.NF)).astype(np.float32)
    y = net.predict(X)
    ref = mo.forward(mo.bf16_round(p), mo.standardize(X, nrm))["y"]
    # FP32 sums of at most 64 terms against float64: |dy| <= 2e-5 (1 + |y|)
    err = np.abs(y - ref) / (1.0 + np.abs(ref))
    print("forward max rel err", err.max())
    assert err.max() < 2e-5


@pytest.mark.parametrize("n", [1, 33, 1000])
def test_gradient_matches_oracle(native_mod, n):
    net, p, nrm = make_net(native_mod, seed=20 + n)
    rng = np.random.default_rng(30 + n)
    X = rng.normal(size=(n, mo.NF)).astype(np.float32)
    T = rng.normal(size=n).astype(np.float32)
    loss, g = net.gradient(X, T)
    lref, gref = mo.loss_and_grad(mo.bf16_round(p), mo.standardize(X, nrm), T)
    print(f"n={n} loss {loss} vs {lref}")
    assert abs(loss - lref) <= 1e-5 * (1.0 + abs(lref))
    # FP32 accumulation over n samples: |dg| <= 1e-4 |g| + 2e-6 max|g|
    scale = np.abs(gref).max()
    bad = np.abs(g - gref) > 1e-4 * np.abs(gref) + 2e-6 * scale
    assert not bad.any(), (np.argwhere(bad)[:5].ravel(), g[bad][:5], gref[bad][:5])


@pytest.mark.parametrize("tau", [0.01, 0.5])
def test_pinball_gradient_matches_oracle(native_mod, tau):
    net, p, nrm = make_net(native_mod, seed=70)
    net.set_loss(native_mod.Mlp.LOSS_PINBALL, tau)
    rng = np.random.default_rng(71)
    X = rng.normal(size=(500, mo.NF)).astype(np.float32)
    T = rng.normal(size=500).astype(np.float32)
    loss, g = net.gradient(X, T)
    lref, gref = mo.loss_and_grad(mo.bf16_round(p), mo.standardize(X, nrm), T, tau=tau)
    print(f"tau={tau} loss {loss} vs {lref}")
    assert abs(loss - lref) <= 1e-5 * (1.0 + abs(lref))
    scale = np.abs(gref).max()
    bad = np.abs(g - gref) > 1e-4 * np.abs(gref) + 2e-6 * scale
    assert not bad.any(), (np.argwhere(bad)[:5].ravel(), g[bad][:5], gref[bad][:5])


def test_adam_step_and_bf16_image(native_mod):
    net, p, nrm = make_net(native_mod, seed=40)
    rng = np.random.default_rng(41)
    X = rng.normal(size=(256, mo.NF)).astype(np.float32)
    T = rng.normal(size=256).astype(np.float32)
    _, g = net.gradient(X, T)
    net.train(X, T, epochs=1, batch=256, lr=1e-3)
    w, b, _ = net.params()
    m, v, step = net.adam_state()
    assert step == 1
    w_ref, m_ref, v_ref = mo.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    np.testing.assert_allclose(m, m_ref, rtol=1e-6, atol=0)
    np.testing.assert_allclose(v, v_ref, rtol=1e-6, atol=0)
    np.testing.assert_allclose(w, w_ref, rtol=1e-6, atol=1e-9)
    assert np.array_equal(b, mo.bf16_bits(w)), "BF16 image is the rounded master"


def test_training_fits_a_smooth_target(native_mod):
    rng = np.random.default_rng(50)
    X = rng.normal(size=(20000, mo.NF)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * np.abs(X[:, 3])).astype(np.float32)
    p0 = (rng.normal(size=mo.NPARAM) * 0.1).astype(np.float32)
    net = native_mod.Mlp(p0)
    losses = net.train(X, T, epochs=30, batch=512, lr=3e-3)
    print("losses", losses[[0, 9, 19, 29]])
    assert losses[-1] < 0.1 * np.var(T)
    # the reported epoch loss agrees with a fresh evaluation of the trained weights
    w, _, nrm = net.params()
    y = net.predict(X)
    ref = mo.forward(mo.bf16_round(w), mo.standardize(X, nrm))["y"]
    assert np.abs(y - ref).max() < 1e-4


def small_disk():
    return generate.to_internal(generate.disk_powerlaw(n=600, seed=77))


def test_history_records_do_not_change_the_trajectory(native_mod):
    b = small_disk()
    kw = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01)
    d1 = native_mod.Disk(b, **kw)
    d2 = native_mod.Disk(b, **kw)
    d2.history_enable(1000)
    for _ in range(25):
        d1.step(); d2.step()
    assert_bitwise(d1.state()["r"], d2.state()["r"], "r")
    assert d1.dt_try == d2.dt_try
    H = d2.history()
    assert H.shape == (25 - 4, 16)
    assert np.all(np.isfinite(H))
    assert np.all(H[:, 13] == 0.0)           # no network: no output used


def test_step_features_and_in_loop_proposal(native_mod):
    b = small_disk()
    kw = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01)
    d = native_mod.Disk(b, **kw)
    net, p, nrm = make_net(native_mod, seed=60, scale=0.05)
    d.history_enable(100)
    d.set_proposer(net, safety=0.9, max_log_change=0.5)
    for _ in range(6):
        d.step()
    # the state features of the next step start, from the state and its first force evaluation
    s = d.state()
    a = d.accel(d.t, s["r"], s["v"])
    cnt = d.counts()
    n_si = sum(cnt[:4]); n_m = n_si + sum(cnt[4:6])
    q, amax, pert = mo.step_features(s["r"], a, s["p"][:, 0], s["id"], n_si, n_m)
    d.step()
    H = d.history()
    last = H[-1]
    print("features", last[:12], "oracle", np.log10([q, amax, pert]))
    # the pair ratio comes from FP32 star-relative positions: a few 1e-4 relative in d
    assert abs(last[7] - np.log10(q)) < 2e-3
    assert abs(last[9] - np.log10(amax)) < 1e-5 * max(1.0, abs(np.log10(amax)))
    assert abs(last[10] - np.log10(pert)) < 1e-4 * max(1.0, abs(np.log10(pert)))
    # the in-loop output equals the batch kernel on the same features, bit for bit
    y = net.predict(H[:, :12])
    assert_bitwise(H[:, 13], y, "proposal output")
    st = d.proposer_stats()
    assert st["proposals"] == len(H) and 0 <= st["first_accepted"] <= st["proposals"]
