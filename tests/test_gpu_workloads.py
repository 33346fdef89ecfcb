"""Whole fixed-step main loops on the GPU against the CPU oracle.

The library runs the iterations of a fixed-step main loop that need nothing from the host
in one work-group on the device (red_run_fixed) and hands every other iteration back: an
ejection or hit-centrum due, a step that recorded pair events, an output time.  The runs
below go through mainloop.run_main_loop (a restatement of main.cu:159-281) with the GPU
library, which takes that fused path, and with the oracle, which steps one at a time; the
snapshots (times and states), the event records and the step counts must be identical.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate, mainloop
from test_gpu_events import PRM, events_equal, planted_disk

pytestmark = pytest.mark.gpu

GAUSS = 0.01720209895


def run_loop(eng, prm, fused=True):
    snaps, events = [], []

    def snap(e):
        s = e.state()
        snaps.append((e.t, s["id"].copy(), s["r"].copy(), s["v"].copy()))

    fused_steps = [0]
    if fused and hasattr(eng, "run_fixed"):
        inner = eng.run_fixed

        def counting(*args):
            k, ps, why = inner(*args)
            fused_steps[0] += k
            return k, ps, why

        eng.run_fixed = counting
    info = mainloop.run_main_loop(eng, prm, snap, lambda evs: events.extend(evs), fused=fused)
    info["fused_steps"] = fused_steps[0]
    return snaps, events, info


def compare_runs(ref, got):
    (s0, e0, i0), (s1, e1, i1) = ref, got
    assert i0["steps"] == i1["steps"]
    assert len(s0) == len(s1)
    for k, ((t0, id0, r0, v0), (t1, id1, r1, v1)) in enumerate(zip(s0, s1)):
        assert t0 == t1, (k, t0, t1)
        assert np.array_equal(id0, id1), k
        assert_bitwise(r1, r0, f"r snapshot {k}")
        assert_bitwise(v1, v0, f"v snapshot {k}")
    events_equal(e0, e1, "events")


def loop_prm(dt_day, steps, every, **kw):
    dt = dt_day * GAUSS
    prm = dict(stop_time=(steps - 0.5) * dt, output_interval=every * dt, inner_steps=True,
               radii_enhance_factor=kw.get("radii_enhance_factor", 0.0),
               encounter_hill_factor=kw.get("encounter_hill_factor", 0.0),
               close_encounter=kw.get("close_encounter", False))
    return prm


@pytest.mark.parametrize("hill", [0.0, 3.0])
def test_small_disk_full_length_bit_exact(oracle_mod, native_mod, hill):
    """Star + 2 giants + 253 planetesimals (seed 101), RK4 at dt = 1 day, ejection at 100 AU,
    collisions at 1 x (R_i + R_j), close encounters at `hill` mutual Hill radii (0 = off):
    10^4 steps with a snapshot every 365.25 days, the GPU run fused on the device."""
    b = generate.to_internal(generate.disk_small(seed=101))
    kw = dict(integrator=2, adaptive=False, dt_day=1.0, ejection=100.0, radii_enhance_factor=1.0,
              encounter_hill_factor=hill)
    prm = dict(stop_time=9999.5 * GAUSS, output_interval=365.25 * GAUSS, inner_steps=True, radii_enhance_factor=1.0,
               encounter_hill_factor=hill)
    print(f"case 256-body RK4 10^4 steps n={b.n} hill={hill}", flush=True)
    ref = run_loop(oracle_mod.Oracle(b, **kw), prm)
    d = native_mod.Disk(b, use_graph=True, **kw)
    got = run_loop(d, prm)
    compare_runs(ref, got)
    # with the Hill test on, the pair tests run in every sweep; no planetesimal (non-self-
    # interacting: it meets only the giants at 5.2 and 9.5 AU) comes within 3 mutual Hill
    # radii in these 10^4 days, in the oracle as on the GPU, so every step stays fused
    assert got[2]["steps"] == 10000 and got[2]["fused_steps"] == 10000
    assert len(got[1]) == len(ref[1]) == 0
    assert len(got[0]) > 20


@pytest.mark.parametrize("n_pl", [1200, 1600])
def test_fused_loop_multi_tile_sweep(oracle_mod, native_mod, n_pl):
    """A sweep that needs more than one LDS tile of sources in the fused loop: the star and
    two giants against 1203 sources (few-sinks mode, 512 sources per tile), and against 1603
    sources, more than the 1536 pair terms of one tile (9609 interacting pairs per
    evaluation, within the fused loop's limit)."""
    b = generate.to_internal(generate.disk_small(n_pl=n_pl, seed=17))
    print(f"case fused multi-tile n={b.n}", flush=True)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5, radii_enhance_factor=1.0)
    prm = loop_prm(0.5, 40, 10, radii_enhance_factor=1.0)
    ref = run_loop(oracle_mod.Oracle(b, **kw), prm)
    got = run_loop(native_mod.Disk(b, **kw), prm)
    compare_runs(ref, got)
    assert got[2]["fused_steps"] == got[2]["steps"] == 40


def test_fused_loop_refuses_large_evaluations(oracle_mod, native_mod):
    """Above 16384 interacting pairs per evaluation the fused loop is refused (stepping launch
    by launch is faster there) and the main loop steps one at a time: a power-law disk of
    2100 bodies, 70 of them self-interacting (289k pairs per evaluation)."""
    b = generate.to_internal(generate.disk_powerlaw(n=2100, seed=18))
    print(f"case fused refused n={b.n}", flush=True)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5, radii_enhance_factor=1.0)
    d = native_mod.Disk(b, **kw)
    with pytest.raises(native_mod.RedError) as ex:
        d.run_fixed(10, 1e9, 1e9)
    assert ex.value.code == native_mod.ERR_UNSUPPORTED
    prm = loop_prm(0.5, 20, 10, radii_enhance_factor=1.0)
    ref = run_loop(oracle_mod.Oracle(b, **kw), prm)
    got = run_loop(native_mod.Disk(b, **kw), prm)
    compare_runs(ref, got)
    assert got[2]["fused_steps"] == 0 and got[2]["steps"] == 20


@pytest.mark.parametrize("integ", [0, 1, 2, 3, 4])
def test_fused_loop_events_bit_exact(oracle_mod, native_mod, integ):
    """The planted disk (collisions merged, two ejections, two hit-centrum merges, the
    compaction after eight events) under every fixed-step integrator: Euler, RK2, RK4,
    Cash-Karp and RKF7(8) without step control (RKF7(8) clears the event list after f0)."""
    b = planted_disk()
    kw = dict(integrator=integ, adaptive=False, dt_day=0.5, **PRM)
    prm = loop_prm(0.5, 60, 7, **PRM)
    print(f"case fused loop integ={integ}", flush=True)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, **kw)
    ref = run_loop(o, prm)
    got = run_loop(d, prm)
    compare_runs(ref, got)
    cnt = d.counters()
    assert cnt == o.counters()
    assert cnt["n_ejection"] == 2 and cnt["n_hit_centrum"] == 2 and cnt["n_collision"] > 0
    assert d.n < b.n   # compaction ran
    assert 0 < got[2]["fused_steps"] < got[2]["steps"]


def test_fused_loop_close_encounter_mode(oracle_mod, native_mod):
    b = planted_disk(seed=45)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5, close_encounter=True, radii_enhance_factor=500.0)
    prm = loop_prm(0.5, 40, 5, close_encounter=True, radii_enhance_factor=500.0)
    print("case fused loop close-encounter mode", flush=True)
    ref = run_loop(oracle_mod.Oracle(b, **kw), prm)
    got = run_loop(native_mod.Disk(b, event_capacity=20000, **kw), prm)
    compare_runs(ref, got)
    assert len(got[1]) > 0 and all(e["name"] == 3 for e in got[1])
    assert got[2]["fused_steps"] > 0


def test_fused_loop_stops_and_fallbacks(native_mod):
    """Stop reasons, and the configurations it leaves to the step-by-step loop."""
    b = generate.to_internal(generate.disk_small(seed=101))
    print("case fused loop stop reasons", flush=True)
    d = native_mod.Disk(b, integrator=2, adaptive=False, dt_day=1.0)
    dt = 1.0 * GAUSS
    k, ps, why = d.run_fixed(25, 1e9, 1e9)
    ps_seq = 0.0
    for _ in range(25):
        ps_seq += abs(dt)
    assert (k, why, ps) == (25, native_mod.FUSED_STOP_STEPS, ps_seq)
    k, ps, why = d.run_fixed(1000, 1e9, 10 * dt - 1e-12, 0.0)
    assert why == native_mod.FUSED_STOP_OUTPUT and k == 10
    k, ps, why = d.run_fixed(1000, d.t + 3.5 * dt, 1e9, 0.0)
    assert why == native_mod.FUSED_STOP_TIME and k == 4
    k, ps, why = d.run_fixed(1000, 1e9, 0.0, 0.0)   # output due before any step
    assert (k, why) == (0, native_mod.FUSED_STOP_OUTPUT)
    for kw in (dict(integrator=4, adaptive=True, tolerance=1e-10), dict(integrator=2, precision=native_mod.PREC_MIXED)):
        e = native_mod.Disk(b, dt_day=1.0, **kw)
        with pytest.raises(native_mod.RedError) as ex:
            e.run_fixed(10, 1e9, 1e9)
        assert ex.value.code == native_mod.ERR_UNSUPPORTED
        info = mainloop.run_main_loop(e, dict(stop_time=e.t + 4.5 * dt, output_interval=1e9), lambda eng: None)
        assert info["steps"] >= 1 and e.t > 4.5 * dt


def test_fused_loop_hill_events_and_every_step_output(oracle_mod, native_mod):
    """Pair events from the mutual-Hill-radius test (10 R_H, collisions at 1 x (R_i + R_j))
    handed back after the step that found them, and a snapshot after every step (the fused
    loop stops at each output time after one step)."""
    b = planted_disk(seed=44)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5, hit_centrum=0.0, ejection=0.0, radii_enhance_factor=1.0,
              encounter_hill_factor=10.0)
    print("case fused loop hill events", flush=True)
    for every in (7, 1):
        prm = loop_prm(0.5, 30, every, radii_enhance_factor=1.0, encounter_hill_factor=10.0)
        ref = run_loop(oracle_mod.Oracle(b, **kw), prm)
        d = native_mod.Disk(b, event_capacity=20000, **kw)
        got = run_loop(d, prm)
        compare_runs(ref, got)
        assert d.counters()["n_close_encounter"] > 0
        assert got[2]["fused_steps"] > 0
        if every == 1:
            assert len(got[0]) >= got[2]["steps"]
