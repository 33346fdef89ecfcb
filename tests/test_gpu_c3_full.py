"""Parity at the full size of the mixed-precision benchmark disk.

The disk: 1 star + 2 giants + 65536 planetesimals (m log-uniform in [1e-12, 1e-9] M_sun,
non-self-interacting) + 65533 test particles, N = 131072, seed 303 -- the workload the
benchmark quotes pair rates on.  Its three sweeps are the reference's launches
(pp_disk.cu:609-681) over the bounds of number_of_bodies.cpp:196-258: 3 SI sinks x 65539
sources, 65536 NSI sinks x 3 SI sources, 65533 NI sinks x 65539 sources (4,295,360,509
pairs per evaluation).

Three pairs are planted so that every evaluation reports collisions: a test particle
inside the largest planetesimal (NI sweep, at half its radius), a planetesimal
inside the first giant's radius (NSI sweep) and a second test particle inside the second
giant's radius (NI sweep against an FP64 source of the mixed path).

Checked here, with the oracle run on up to 16 host threads (its OpenMP sink loop gives
the serial loop's results bit for bit):
* FP64: one force evaluation and one fixed Cash-Karp step of the whole disk equal the
  oracle bit for bit, events included;
* mixed: every acceleration within 2e-6 relative of the oracle's (FP32 r^-3 carries about
  1e-7 relative rounding per pair), raw event lists identical to the FP64 path's and the
  oracle's on the same state;
* mixed trajectories: after two fixed 1-day Cash-Karp steps positions agree with the
  oracle's to 1e-10 and velocities to 1e-9 relative per body; one adaptive RK5(4) step at
  tolerance 1e-10 takes the oracle's number of attempts and a step within 1e-6 relative.
"""
import os

import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
EV_KEYS = ("name", "id1", "idx1", "id2", "idx2", "stage", "sink", "source")


@pytest.fixture(scope="module")
def disk():
    b = generate.to_internal(generate.disk_mixed(n_pl=65536, n_tp=65533, seed=303))
    assert b.n == 131072 and list(b.counts) == [1, 2, 0, 0, 0, 65536, 65533]
    nm = 3 + 65536
    # test particle 100 inside the largest planetesimal: d = R / 2
    pl, tp = 3 + int(np.argmax(b.mass[3:nm])), nm + 100
    b.r[tp] = b.r[pl] + np.array([0.5 * b.radius[pl], 0.0, 0.0])
    b.v[tp] = b.v[pl]
    # planetesimal 1000 inside the first giant
    g1 = 1
    b.r[3 + 1000] = b.r[g1] + np.array([0.0, 0.5 * b.radius[g1], 0.0])
    b.v[3 + 1000] = b.v[g1]
    # test particle 2000 inside the second giant (an FP64 source of the mixed sweep)
    g2 = 2
    b.r[nm + 2000] = b.r[g2] + np.array([0.0, 0.0, 0.25 * b.radius[g2]])
    b.v[nm + 2000] = b.v[g2]
    return b


@pytest.fixture(scope="module")
def oracle_accel(oracle_mod, disk):
    o = oracle_mod.Oracle(disk, radii_enhance_factor=1.0, n_threads=THREADS)
    a = o.accel(0.25, disk.r, disk.v, stage=2)
    return a, o.raw_events()


def _events_equal(ev_a, ev_b, what):
    assert len(ev_a) == len(ev_b), (what, len(ev_a), len(ev_b))
    for ea, eb in zip(ev_a, ev_b):
        for k in EV_KEYS:
            assert ea[k] == eb[k], (what, k, ea[k], eb[k])
        for k in ("t", "d", "r1", "v1", "r2", "v2"):
            assert_bitwise(np.array(ea[k]), np.array(eb[k]), f"{what} {k}")


def test_c3_fp64_accel_bit_exact(native_mod, disk, oracle_accel):
    print(f"case c3 full fp64 accel n={disk.n}", flush=True)
    a_o, ev_o = oracle_accel
    d = native_mod.Disk(disk, radii_enhance_factor=1.0, event_capacity=4096)
    a_g = d.accel(0.25, disk.r, disk.v, stage=2)
    assert_bitwise(a_g, a_o, "accel")
    ev_g = d.raw_events()
    names = sorted((e["sink"], e["source"]) for e in ev_o)
    pl = 3 + int(np.argmax(disk.mass[3:3 + 65536]))
    assert names == [(3 + 1000, 1), (3 + 65536 + 100, pl), (3 + 65536 + 2000, 2)], names
    _events_equal(ev_o, ev_g, "fp64 vs oracle")
    d.close()


def test_c3_fp64_cash_karp_step_bit_exact(oracle_mod, native_mod, disk):
    print(f"case c3 full fp64 cash-karp step n={disk.n}", flush=True)
    kw = dict(integrator=3, adaptive=False, dt_day=1.0, radii_enhance_factor=1.0)
    o = oracle_mod.Oracle(disk, n_threads=THREADS, **kw)
    d = native_mod.Disk(disk, event_capacity=4096, **kw)
    assert o.step() == d.step()
    so, sg = o.state(), d.state()
    assert_bitwise(sg["r"], so["r"], "r")
    assert_bitwise(sg["v"], so["v"], "v")
    # the planted pairs are reported by the stage evaluations that still find them close
    # (the first two at least: their stage positions move both members with the same velocity)
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert len(ev_o) >= 6, len(ev_o)
    _events_equal(ev_o, ev_g, "step events")
    # and the collision pass merges them identically
    assert o.check_for_collision() and d.check_for_collision()
    so, sg = o.state(), d.state()
    for k in ("r", "v", "p"):
        assert_bitwise(sg[k], so[k], f"after merge {k}")
    assert np.array_equal(sg["id"], so["id"])
    d.close()


def test_c3_mixed_within_bound_and_same_events(native_mod, disk, oracle_accel):
    print(f"case c3 full mixed accel n={disk.n}", flush=True)
    a_o, ev_o = oracle_accel
    d = native_mod.Disk(disk, precision=native_mod.PREC_MIXED, radii_enhance_factor=1.0, event_capacity=4096)
    a_m = d.accel(0.25, disk.r, disk.v, stage=2)
    rel = np.linalg.norm(a_m - a_o, axis=1) / np.maximum(np.linalg.norm(a_o, axis=1), 1e-300)
    print("mixed rel err max", rel.max(), "median", np.median(rel), flush=True)
    assert np.all(np.isfinite(a_m))
    assert rel.max() < 2e-6
    _events_equal(ev_o, d.raw_events(), "mixed vs oracle")
    d.close()


def test_c3_mixed_fixed_steps_within_tolerance(oracle_mod, native_mod):
    print("case c3 full mixed two fixed cash-karp steps", flush=True)
    b = generate.to_internal(generate.disk_mixed(n_pl=65536, n_tp=65533, seed=303))
    kw = dict(integrator=3, adaptive=False, dt_day=1.0, radii_enhance_factor=1.0)
    o = oracle_mod.Oracle(b, n_threads=THREADS, **kw)
    d = native_mod.Disk(b, precision=native_mod.PREC_MIXED, **kw)
    for _ in range(2):
        assert o.step() == d.step()
    so, sg = o.state(), d.state()
    dr = np.linalg.norm(sg["r"] - so["r"], axis=1) / np.linalg.norm(so["r"], axis=1)
    dv = np.linalg.norm(sg["v"] - so["v"], axis=1) / np.linalg.norm(so["v"], axis=1)
    print("pos rel err max", dr.max(), "vel rel err max", dv.max(), flush=True)
    assert dr.max() < 1e-10 and dv.max() < 1e-9
    assert o.n_event() == 0 and d.n_event() == 0
    d.close()


def test_c3_mixed_adaptive_step_matches_oracle_control(oracle_mod, native_mod):
    print("case c3 full mixed adaptive rk5(4) step", flush=True)
    b = generate.to_internal(generate.disk_mixed(n_pl=65536, n_tp=65533, seed=303))
    kw = dict(integrator=3, adaptive=True, tolerance=1e-10, dt_day=1.0, radii_enhance_factor=1.0)
    o = oracle_mod.Oracle(b, n_threads=THREADS, **kw)
    d = native_mod.Disk(b, precision=native_mod.PREC_MIXED, **kw)
    dt_o, dt_g = o.step(), d.step()
    assert abs(dt_g - dt_o) <= 1e-6 * abs(dt_o), (dt_g, dt_o)
    assert d.counters()["tried"] == o.counters()["tried"]
    assert abs(d.dt_try - o.dt_try) <= 1e-6 * abs(o.dt_try)
    d.close()
