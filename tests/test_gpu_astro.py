"""Astro-centric frame on the GPU against the CPU oracle, FP64, bit for bit.

The reference computes this frame on its CPU path only (pp_disk.cu:409-454): the star is
body 0 at the origin, outside the sweep bounds (number_of_bodies.cpp:218-221), every sink
starts from -(m_star + m_i) r_i / |r_i|^3 and every active source adds its direct term and
subtracts its indirect term.  Forces, pair events, integrator steps (with and without graph
replay) and the driver loop with merges must equal the oracle's.  A disk with
non-self-interacting bodies, or mixed precision, is refused: the reference's NSI and NI
sweeps keep the star as a source at the origin, where the indirect term is 0/0.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate
from test_gpu_events import lockstep

pytestmark = pytest.mark.gpu


def _helio(b):
    return generate.to_internal(b, transform_bc=False)


def _allsi(n, seed, **kw):
    return _helio(generate.disk_allsi(n, seed=seed, **kw))


def test_astro_accel_bit_exact(oracle_mod, native_mod):
    b = _allsi(3001, seed=11)
    b.ids[[5, 77, 300]] *= -1
    print(f"case astro accel n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, frame_center=0)
    d = native_mod.Disk(b, frame_center=0)
    a_o = o.accel(0.5, b.r, b.v)
    a_g = d.accel(0.5, b.r, b.v)
    assert_bitwise(a_g, a_o, "accel")
    assert np.all(a_g[0] == 0.0)


@pytest.mark.parametrize("kind", ["collision", "hill"])
def test_astro_events_bit_exact(oracle_mod, native_mod, kind):
    if kind == "collision":
        b = _allsi(1500, seed=21, a_range=(0.9, 1.1))
        kw = dict(radii_enhance_factor=3000.0)
    else:
        b = _allsi(600, seed=8, a_range=(1.0, 1.2), m_range=(1e-11, 1e-9))
        kw = dict(radii_enhance_factor=1.0, encounter_hill_factor=30.0)
    print(f"case astro events {kind} n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, frame_center=0, **kw)
    d = native_mod.Disk(b, frame_center=0, event_capacity=200000, **kw)
    assert_bitwise(d.accel(3.0, b.r, b.v, stage=2), o.accel(3.0, b.r, b.v, stage=2), "accel")
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert len(ev_o) > 0
    assert len(ev_o) == len(ev_g)
    for eo, eg in zip(ev_o, ev_g):
        for k in ("name", "id1", "idx1", "id2", "idx2", "stage", "sink", "source"):
            assert eo[k] == eg[k], (k, eo, eg)
        for k in ("t", "d", "r1", "v1", "r2", "v2"):
            assert_bitwise(np.array(eo[k], dtype=np.float64), np.array(eg[k], dtype=np.float64), k)


CASES = [
    ("euler", 0, False, 0.0, 0.5),
    ("rk4a", 2, True, 1e-9, 1.0),
    ("rk5a", 3, True, 1e-10, 1.0),
    ("rkf78a", 4, True, 1e-12, 2.0),
]


@pytest.mark.parametrize("name,integ,adaptive,tol,dt_day", CASES)
@pytest.mark.parametrize("graph", [False, True])
def test_astro_steps_bit_exact(oracle_mod, native_mod, name, integ, adaptive, tol, dt_day, graph):
    print(f"case astro {name} graph={graph}", flush=True)
    b = _allsi(200, seed=31)
    kw = dict(integrator=integ, adaptive=adaptive, tolerance=tol, dt_day=dt_day, frame_center=0)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, use_graph=graph, **kw)
    for k in range(12):
        dt_o, dt_g = o.step(), d.step()
        assert dt_o == dt_g, (k, dt_o, dt_g)
        assert o.dt_try == d.dt_try, (k, o.dt_try, d.dt_try)
    so, sg = o.state(), d.state()
    assert_bitwise(sg["r"], so["r"], "r")
    assert_bitwise(sg["v"], so["v"], "v")
    assert np.all(sg["r"][0] == 0.0) and np.all(sg["v"][0] == 0.0)
    assert o.counters() == d.counters()


def test_astro_driver_loop_bit_exact(oracle_mod, native_mod):
    """Collisions, a hit-centrum merge into the star (which then grows and moves) and
    ejections in the astro-centric frame, in lock step with the oracle."""
    import math
    B = generate._Builder()
    rng = generate._rng(43)
    generate._add_star(B, r_star=1e-7)
    rho = 2.0 * generate.GRAM_PER_CM3_TO_SOLAR_PER_AU3
    m = np.exp(rng.uniform(math.log(1e-10), math.log(1e-8), 100))
    generate._add_bodies(B, 3, "pp", m, rho, generate._elements(rng, 100, 1.0, 1.04), 1.0)
    generate._add_bodies(B, 3, "hc", np.array([2e-10, 3e-10]), rho, generate._elements(rng, 2, 0.2, 0.25), 1.0)
    generate._add_bodies(B, 3, "ej", np.array([1e-11, 2e-11]), rho, generate._elements(rng, 2, 8.0, 9.0), 1.0)
    b = _helio(B.build())
    kw = dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5, frame_center=0, hit_centrum=0.3,
              ejection=6.0, radii_enhance_factor=2500.0)
    print(f"case astro loop n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, **kw)
    n_events, cnt = lockstep(o, d, 40)
    print("events", n_events, cnt, flush=True)
    assert cnt["n_collision"] > 0 and cnt["n_ejection"] >= 2 and cnt["n_hit_centrum"] == 2


def test_astro_refuses_nsi_and_mixed(native_mod):
    b = _helio(generate.disk_small(60, seed=3))           # planetesimals: NSI
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Disk(b, frame_center=0)
    assert e.value.code == 5   # RED_ERR_UNSUPPORTED
    b = _allsi(100, seed=3)
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Disk(b, frame_center=0, precision=native_mod.PREC_MIXED)
    assert e.value.code == 5
    native_mod.Disk(b, frame_center=0).close()
