"""Ensembles of independent disks (red_ensemble_*): every member must follow exactly the
steps it takes alone.  Each disk of a small ensemble of uneven sizes is compared bit for
bit with the single-disk path (red_step in the reference loop `while t <= t_stop`) and,
for two disks, with the CPU oracle; the autotuned launch shape and graph replay must not
change any result."""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def small_ensemble(n=10, seed=404):
    return [generate.to_internal(d) for d in generate.ensemble(n_disks=n, seed=seed, n_range=(16, 700))]


def run_alone(native_mod, b, t_stop, **kw):
    d = native_mod.Disk(b, **kw)
    while d.t <= t_stop:
        d.step()
    return d


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_members_match_single_disk(native_mod, oracle_mod, graph):
    disks = small_ensemble()
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    print(f"case ensemble n_disks={len(disks)} sizes={[d.n for d in disks]} graph={graph}", flush=True)
    E = native_mod.Ensemble(disks, use_graph=graph, **kw)
    t_stop = 4.0 * 0.01720209895          # four days in internal units
    info = E.advance(t_stop)
    assert info["active_disks"] == 0 and info["supersteps"] > 0
    for k, b in enumerate(disks):
        alone = run_alone(native_mod, b, t_stop, **kw)
        r, v = E.state(k)
        s = alone.state()
        assert_bitwise(r, s["r"], f"disk {k} r")
        assert_bitwise(v, s["v"], f"disk {k} v")
        di = E.disk_info(k)
        c = alone.counters()
        assert (di["tried"], di["failed"], di["passed"]) == (c["tried"], c["failed"], c["passed"])
        assert di["t"] == alone.t and di["dt_try"] == alone.dt_try
    for k in (0, 3):
        o = oracle_mod.Oracle(disks[k], **kw)
        while o.t <= t_stop:
            o.step()
        r, _ = E.state(k)
        assert_bitwise(r, o.state()["r"], f"disk {k} vs oracle")


def test_ensemble_autotune_keeps_results(native_mod):
    disks = small_ensemble(n=8, seed=405)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    E1 = native_mod.Ensemble(disks, **kw)
    E2 = native_mod.Ensemble(disks, **kw)
    block, jtile, table = E2.autotune(reps=3)
    print("autotune", block, jtile, table.tolist())
    assert block in (64, 128, 256) and jtile in (64, 128, 256)
    assert np.all(table > 0)
    t_stop = 2.0 * 0.01720209895
    E1.advance(t_stop); E2.advance(t_stop)
    for k in range(len(disks)):
        assert_bitwise(E1.state(k)[0], E2.state(k)[0], f"disk {k}")


def test_ensemble_fixed_step_rk4(native_mod):
    disks = small_ensemble(n=6, seed=406)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5)
    E = native_mod.Ensemble(disks, **kw)
    t_stop = 3.0 * 0.01720209895
    E.advance(t_stop)
    for k, b in enumerate(disks):
        alone = run_alone(native_mod, b, t_stop, **kw)
        assert_bitwise(E.state(k)[0], alone.state()["r"], f"disk {k}")


def test_ensemble_advance_in_parts(native_mod):
    """A run to t1 and then to t2 steps every member as one run to t2 does (the loop
    condition only decides whether a step is taken), also when the first call is cut short
    by max_attempts; every member still equals the single-disk path."""
    disks = small_ensemble(n=6, seed=407)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    print("case ensemble advance in parts", flush=True)
    t1, t2 = 1.5 * 0.01720209895, 4.0 * 0.01720209895
    whole = native_mod.Ensemble(disks, **kw)
    whole.advance(t2)
    parts = native_mod.Ensemble(disks, **kw)
    info = parts.advance(t1)
    assert info["supersteps"] > 0 and info["active_disks"] == 0
    cut = parts.advance(t2, max_attempts=1)
    assert cut["supersteps"] == 1 and cut["active_disks"] > 0
    rest = parts.advance(t2)
    assert rest["supersteps"] > 0 and rest["active_disks"] == 0
    for k, b in enumerate(disks):
        assert_bitwise(parts.state(k)[0], whole.state(k)[0], f"disk {k} r")
        assert_bitwise(parts.state(k)[1], whole.state(k)[1], f"disk {k} v")
        assert parts.disk_info(k) == whole.disk_info(k)
        alone = run_alone(native_mod, b, t2, **kw)
        assert_bitwise(parts.state(k)[0], alone.state()["r"], f"disk {k} alone")
