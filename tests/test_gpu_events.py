"""The whole driver loop on the GPU against the CPU oracle, in lock step.

A planted disk exercises every event path of the reference (pp_disk.cu:683-1409):
collisions between self-interacting bodies (merged, earliest record per pair kept),
close encounters at a few Hill radii (reported), ejections beyond R_ej, hit-centrum
merges into the star, and the compaction after eight events.  After every check and
every step the event records, the state, the counts and the step sizes must be
identical.  Also: close-encounter mode, the error raised for a merge with a
super-planetesimal, and the event-capacity overflow error.
"""
import math

import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu

EV_FIELDS_I = ("name", "id1", "idx1", "id2", "idx2", "stage", "attempt", "sink", "source")
EV_FIELDS_F = ("t", "d", "p1", "p2", "ps", "r1", "v1", "r2", "v2", "rs", "vs", "oe1", "oe2", "oes")


def events_equal(eo, eg, where):
    assert len(eo) == len(eg), (where, len(eo), len(eg))
    for a, b in zip(eo, eg):
        for k in EV_FIELDS_I:
            assert a[k] == b[k], (where, k, a[k], b[k])
        for k in EV_FIELDS_F:
            assert_bitwise(np.array(a[k], dtype=np.float64), np.array(b[k], dtype=np.float64), f"{where}:{k}")


def planted_disk(seed=41):
    """star + 60 protoplanets + 50 planetesimals + 30 test particles in a narrow ring,
    plus two bodies inside the hit-centrum radius and two beyond the ejection radius."""
    B = generate._Builder()
    rng = generate._rng(seed)
    generate._add_star(B, r_star=1e-7)   # the enhanced collision radius must not swallow the ring
    rho = 2.0 * generate.GRAM_PER_CM3_TO_SOLAR_PER_AU3
    m = np.exp(rng.uniform(math.log(1e-10), math.log(1e-8), 60))
    generate._add_bodies(B, 3, "pp", m, rho, generate._elements(rng, 60, 1.0, 1.04), 1.0)
    m = np.exp(rng.uniform(math.log(1e-12), math.log(1e-10), 50))
    generate._add_bodies(B, 5, "pl", m, rho, generate._elements(rng, 50, 1.0, 1.04), 1.0)
    generate._add_bodies(B, 6, "tp", np.zeros(30), 0.0, generate._elements(rng, 30, 1.0, 1.04), 1.0)
    generate._add_bodies(B, 3, "hc", np.array([2e-10, 3e-10]), rho, generate._elements(rng, 2, 0.2, 0.25), 1.0)
    generate._add_bodies(B, 5, "ej", np.array([1e-11, 2e-11]), rho, generate._elements(rng, 2, 8.0, 9.0), 1.0)
    return generate.to_internal(B.build())


PRM = dict(hit_centrum=0.3, ejection=6.0, radii_enhance_factor=2500.0, encounter_hill_factor=0.0)


def lockstep(o, d, steps, check_rebuild=True):
    n_events = 0
    for k in range(steps):
        ho, hg = o.check_for_ejection_hit_centrum(), d.check_for_ejection_hit_centrum()
        assert ho == hg, ("ejection", k)
        if ho:
            events_equal(o.sp_events(), d.sp_events(), f"ejection step {k}")
            n_events += len(o.sp_events())
            o.clear_event_counter(); d.clear_event_counter()
        dto, dtg = o.step(), d.step()
        assert dto == dtg, ("dt", k, dto, dtg)
        ho, hg = o.check_for_collision(), d.check_for_collision()
        assert ho == hg, ("collision", k)
        if ho:
            events_equal(o.sp_events(), d.sp_events(), f"collision step {k}")
            n_events += len(o.sp_events())
            o.clear_event_counter(); d.clear_event_counter()
        if check_rebuild:
            assert o.check_for_rebuild_vectors(8) == d.check_for_rebuild_vectors(8), ("rebuild", k)
        so, sg = o.state(), d.state()
        assert_bitwise(sg["r"], so["r"], f"r step {k}")
        assert_bitwise(sg["v"], so["v"], f"v step {k}")
        assert_bitwise(sg["p"], so["p"], f"p step {k}")
        assert np.array_equal(sg["id"], so["id"]), k
        assert o.counts() == d.counts(), k
    co, cg = o.counters(), d.counters()
    assert co == cg, (co, cg)
    return n_events, co


@pytest.mark.parametrize("integ,adaptive,tol,dt,strategy", [(2, False, 0.0, 0.5, 0), (4, True, 1e-10, 0.5, 0),
                                                            (3, True, 1e-10, 0.5, 0), (4, True, 1e-10, 0.5, 4)])
def test_driver_loop_events_bit_exact(oracle_mod, native_mod, integ, adaptive, tol, dt, strategy):
    # strategy 4: the NSI x SI and NI x massive sweeps with one work-item per sink
    print(f"case loop integ={integ} adaptive={adaptive} strategy={strategy}", flush=True)
    b = planted_disk()
    kw = dict(integrator=integ, adaptive=adaptive, tolerance=tol, dt_day=dt, **PRM)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, strategy=strategy, **kw)
    n_events, cnt = lockstep(o, d, 40)
    print("events", n_events, cnt)
    assert cnt["n_collision"] > 0 and cnt["n_ejection"] == 2 and cnt["n_hit_centrum"] == 2
    assert d.n < b.n   # compaction ran


def test_driver_loop_mixed_events_match_fp64(native_mod):
    """Mixed precision: on the same state the event lists equal the FP64 path's bit for
    bit; after a step (whose stage states differ in the last digits) the same pairs are
    found at the same stages."""
    b = planted_disk(seed=43)
    kw = dict(integrator=2, dt_day=0.5, event_capacity=20000, **PRM)
    d64 = native_mod.Disk(b, **kw)
    dmx = native_mod.Disk(b, precision=native_mod.PREC_MIXED, **kw)
    for stage in (0, 3):
        d64.accel(1.5, b.r, b.v, stage=stage)
        dmx.accel(1.5, b.r, b.v, stage=stage)
    e64, emx = d64.raw_events(), dmx.raw_events()
    assert len(e64) > 0
    events_equal(e64, emx, "same state")
    d64.clear_event_counter(); dmx.clear_event_counter()
    d64.step(); dmx.step()
    key = lambda e: (e["name"], e["id1"], e["id2"], e["stage"], e["attempt"])
    assert [key(e) for e in d64.raw_events()] == [key(e) for e in dmx.raw_events()]


def test_mixed_close_encounter_mode_matches_fp64(oracle_mod, native_mod):
    """Close-encounter mode in mixed precision: on one state the CLOSE_ENCOUNTER records
    (pair order as found, stage, orbital elements) equal the FP64 path's bit for bit, and
    after the check both equal the oracle's (12 pairs within 2500 (R_i + R_j) on this
    state).  The raw records are compared GPU against GPU only: the reference's CPU sweep
    stores the pair's parameters at detection (pp_disk.cu:515-516), its GPU kernel leaves
    them to the check (pp_disk.cu:33-66), and the oracle follows the CPU sweep."""
    b = planted_disk(seed=47)
    kw = dict(integrator=2, dt_day=0.5, close_encounter=True, radii_enhance_factor=2500.0)
    o = oracle_mod.Oracle(b, **kw)
    d64 = native_mod.Disk(b, event_capacity=20000, **kw)
    dmx = native_mod.Disk(b, precision=native_mod.PREC_MIXED, event_capacity=20000, **kw)
    print(f"case mixed close-encounter mode n={b.n}", flush=True)
    for d in (o, d64, dmx):
        d.accel(0.0, b.r, b.v, stage=0)
    e64 = d64.raw_events()
    assert len(e64) == len(o.raw_events()) == 12 and all(e["name"] == 3 for e in e64)
    events_equal(e64, dmx.raw_events(), "mixed same state")
    assert o.check_for_close_encounter() and d64.check_for_close_encounter() and dmx.check_for_close_encounter()
    sp = o.sp_events()
    assert len(sp) == 12 and all(e["oe1"][0] != 0.0 for e in sp)
    events_equal(sp, d64.sp_events(), "fp64 sp events")
    events_equal(sp, dmx.sp_events(), "mixed sp events")


@pytest.mark.parametrize("strategy", [0, 4])
def test_hill_close_encounters_reported(oracle_mod, native_mod, strategy):
    # pairs within 10 mutual Hill radii in the ring (pairs with the star are never tested)
    print(f"case hill strategy={strategy}", flush=True)
    b = planted_disk(seed=44)
    kw = dict(integrator=2, dt_day=0.5, hit_centrum=0.0, ejection=0.0, radii_enhance_factor=1.0,
              encounter_hill_factor=10.0)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, strategy=strategy, **kw)
    n_events, cnt = lockstep(o, d, 20)
    assert cnt["n_close_encounter"] > 0


def test_close_encounter_mode(oracle_mod, native_mod):
    b = planted_disk(seed=45)
    kw = dict(integrator=2, dt_day=0.5, close_encounter=True, radii_enhance_factor=500.0)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, **kw)
    seen = 0
    for k in range(10):
        assert o.step() == d.step()
        ho, hg = o.check_for_close_encounter(), d.check_for_close_encounter()
        assert ho == hg
        if ho:
            eo, eg = o.sp_events(), d.sp_events()
            events_equal(eo, eg, f"ce step {k}")
            assert all(e["name"] == 3 for e in eg) and all(e["oe1"][0] != 0.0 for e in eg)
            seen += len(eg)
            o.clear_event_counter(); d.clear_event_counter()
    assert seen > 0
    assert_bitwise(d.state()["r"], o.state()["r"], "r")


def test_superplanetesimal_merge_raises(native_mod):
    b = planted_disk(seed=46)
    b.types[b.types == 5] = 4          # planetesimals become super-planetesimals
    b.counts[4], b.counts[5] = b.counts[5], 0
    d = native_mod.Disk(b, integrator=2, dt_day=0.5, radii_enhance_factor=5000.0, event_capacity=20000)
    d.step()
    with pytest.raises(native_mod.RedError) as ei:
        d.check_for_collision()
    assert "super-planetesimal" in str(ei.value)


def test_event_capacity_overflow_raises(native_mod):
    b = planted_disk(seed=47)
    d = native_mod.Disk(b, integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5,
                        radii_enhance_factor=20000.0, event_capacity=16)
    with pytest.raises(native_mod.RedError) as ei:
        d.step()
    assert ei.value.code == 3
