"""Padded storage (-ups, options.cpp:114) on the GPU against the CPU oracle.

The reference's padded kernel (pp_disk.cu:144-215) tests no id: a body made inactive by an
ejection keeps its mass, so until the next compaction it still attracts every sink, is
itself accelerated, and meets pair tests.  Merged bodies are massless and parked at 1e20 AU,
so for them (and for the massless padding particles of that layout) both kernels agree.
Every sweep (producer/consumer, broadcast, LDS tiles, split, fused fixed-step loop) must
equal the oracle's padded rule bit for bit; mixed precision within its stated 2e-6 per body;
the driver loop with planted ejections, close encounters and compaction stays in lock step;
two parked mergers turn the state NaN in the same places as the oracle's.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate, mainloop
from test_gpu_events import PRM, events_equal, lockstep, planted_disk

pytestmark = pytest.mark.gpu


def _inactive(b, idx):
    b.ids[idx] *= -1   # as after an ejection: mass, radius and position kept
    return b


@pytest.mark.parametrize("strategy,kw", [(3, {}), (1, {}), (2, {}), (1, dict(sinks_per_thread=2)),
                                         (3, dict(fp64_split=True))])
def test_padded_accel_bit_exact(oracle_mod, native_mod, strategy, kw):
    print(f"case padded allsi n=600 strategy={strategy} {kw}", flush=True)
    b = _inactive(generate.to_internal(generate.disk_allsi(600, seed=3)), [5, 77, 300])
    o = oracle_mod.Oracle(b, padded=True)
    d = native_mod.Disk(b, padded_storage=True, strategy=strategy, **kw)
    a_g = d.accel(0.0, b.r, b.v)
    assert_bitwise(a_g, o.accel(0.0, b.r, b.v), "accel")
    assert np.all(np.abs(a_g[[5, 77, 300]]).sum(axis=1) > 0.0)   # inactive sinks are accelerated
    plain = native_mod.Disk(b, strategy=strategy, **kw).accel(0.0, b.r, b.v)
    assert not np.array_equal(plain, a_g)   # and inactive sources pull


def test_padded_split_and_classes_bit_exact(oracle_mod, native_mod):
    # 3 SI sinks against 6003 sources (the split path with the broadcast strategy), NSI and NI
    # sinks, inactive bodies in every class; strategy 4 sweeps the NSI and NI sinks with one
    # work-item per sink (every source packed as it is: padded storage parks none)
    print("case padded disk_mixed 3+3000+3000", flush=True)
    b = generate.to_internal(generate.disk_mixed(n_pl=3000, n_tp=3000, seed=5))
    b = _inactive(b, [1, 10, 2000, 4000, 5000])
    o = oracle_mod.Oracle(b, padded=True)
    for strategy in (1, 3, 4):
        d = native_mod.Disk(b, padded_storage=True, strategy=strategy)
        assert_bitwise(d.accel(2.0, b.r, b.v), o.accel(2.0, b.r, b.v), f"accel strategy {strategy}")
        d.close()


def test_padded_events_bit_exact(oracle_mod, native_mod):
    """Pairs with an inactive body are tested (and recorded with its negated id)."""
    b = generate.to_internal(generate.disk_allsi(1500, seed=21, a_range=(0.9, 1.1)))
    b = _inactive(b, list(range(100, 1500, 7)))
    print("case padded events n=1500", flush=True)
    kw = dict(radii_enhance_factor=3000.0)
    o = oracle_mod.Oracle(b, padded=True, **kw)
    d = native_mod.Disk(b, padded_storage=True, event_capacity=200000, **kw)
    assert_bitwise(d.accel(3.0, b.r, b.v, stage=2), o.accel(3.0, b.r, b.v, stage=2), "accel")
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert sum(1 for e in ev_o if e["id1"] < 0 or e["id2"] < 0) > 0
    events_equal(ev_o, ev_g, "padded raw events")


def test_padded_mixed_within_bound(oracle_mod, native_mod):
    b = _inactive(generate.to_internal(generate.disk_mixed(n_pl=3000, n_tp=3000, seed=5)), [10, 2000, 4000])
    print("case padded mixed 3+3000+3000", flush=True)
    a_o = oracle_mod.Oracle(b, padded=True).accel(1.0, b.r, b.v)
    a_g = native_mod.Disk(b, padded_storage=True, precision=native_mod.PREC_MIXED).accel(1.0, b.r, b.v)
    rel = np.linalg.norm(a_g - a_o, axis=1) / np.maximum(np.linalg.norm(a_o, axis=1), 1e-300)
    assert rel.max() < 2e-6, rel.max()


# ejections (the bodies that keep pulling), close encounters at 30 mutual Hill radii and the
# rebuilds after eight events; no merges, which would park bodies at one point (see below)
PRM_EJ = dict(ejection=6.0, radii_enhance_factor=1.0, encounter_hill_factor=30.0)


@pytest.mark.parametrize("integ,adaptive,tol", [(2, False, 0.0), (4, True, 1e-10), (3, True, 1e-10)])
def test_padded_driver_loop_bit_exact(oracle_mod, native_mod, integ, adaptive, tol):
    """Ejected bodies keep pulling and moving until the compaction after eight events."""
    print(f"case padded loop integ={integ}", flush=True)
    b = planted_disk()
    kw = dict(integrator=integ, adaptive=adaptive, tolerance=tol, dt_day=0.5, **PRM_EJ)
    o = oracle_mod.Oracle(b, padded=True, **kw)
    d = native_mod.Disk(b, padded_storage=True, event_capacity=20000, **kw)
    n_events, cnt = lockstep(o, d, 40)
    print("events", n_events, cnt, flush=True)
    assert cnt["n_ejection"] == 2 and cnt["n_close_encounter"] > 0
    assert d.n < b.n
    assert np.all(np.isfinite(d.state()["r"]))
    # the same run without padding differs: the ejected bodies' pull has acted
    p = native_mod.Disk(b, event_capacity=20000, **kw)
    lockstep(oracle_mod.Oracle(b, **kw), p, 40)
    assert not np.array_equal(p.state()["r"], d.state()["r"])


def _same_allowing_nan(a, b, what):
    """Equal bit for bit where finite; NaN in the same places (a 0/0 on the host and on the
    GPU need not give the same NaN bits)."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), what
    assert_bitwise(np.where(na, 0.0, a), np.where(nb, 0.0, b), what)


def test_padded_parked_mergers_meet(oracle_mod, native_mod):
    """A quirk of the padded kernel kept as it is: merged bodies are parked massless at
    exactly (1e20, 0, 0) (pp_disk.cu:1180-1188), and with no id test two of them meet as sink
    and source at distance 0, a 0/0 term.  Their stage velocities turn NaN, then their stage
    positions, and as sources they carry the NaN into every sink.  In lock step with the
    oracle until the first step with two parked bodies: there the oracle's state turns NaN
    (its error maximum passes NaN over, as the reference's CPU loop does), the fixed-step RK4
    on the GPU gives NaN in the same places, and an adaptive GPU step reports the NaN error
    estimate instead."""
    b = planted_disk()
    for integ, adaptive in ((2, False), (4, True)):
        print(f"case padded parked integ={integ}", flush=True)
        kw = dict(integrator=integ, adaptive=adaptive, tolerance=1e-10, dt_day=0.5, **PRM)
        o = oracle_mod.Oracle(b, padded=True, **kw)
        d = native_mod.Disk(b, padded_storage=True, event_capacity=20000, **kw)
        met = False
        for k in range(40):
            for e in (o, d):
                if e.check_for_ejection_hit_centrum():
                    e.clear_event_counter()
            s = o.state()
            parked = (s["id"] < 0) & (s["r"][:, 0] == 1.0e20)
            o.step()
            if parked.sum() >= 2:
                so = o.state()
                assert np.all(np.isnan(so["r"][parked])), k
                if adaptive:
                    with pytest.raises(native_mod.RedError) as err:
                        d.step()
                    assert err.value.code == 4   # RED_ERR_NAN
                else:
                    d.step()
                    _same_allowing_nan(d.state()["r"], so["r"], f"r step {k}")
                met = True
                break
            d.step()
            for e in (o, d):
                if e.check_for_collision():
                    e.clear_event_counter()
                e.check_for_rebuild_vectors(8)
            assert_bitwise(d.state()["r"], o.state()["r"], f"r step {k}")
        assert met, integ


def test_padded_fused_loop_bit_exact(oracle_mod, native_mod):
    """The fused fixed-step loop applies the same rule (RK4, ejections, close encounters)."""
    b = planted_disk(seed=44)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5, **PRM_EJ)
    prm = dict(stop_time=20.0 * 0.01720209895, output_interval=1e300, inner_steps=True,
               radii_enhance_factor=PRM_EJ["radii_enhance_factor"],
               encounter_hill_factor=PRM_EJ["encounter_hill_factor"])
    print(f"case padded fused n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, padded=True, **kw)
    d = native_mod.Disk(b, padded_storage=True, event_capacity=20000, **kw)
    ev_o, ev_g = [], []
    mainloop.run_main_loop(o, prm, lambda e: None, on_events=ev_o.extend)
    info = mainloop.run_main_loop(d, prm, lambda e: None, on_events=ev_g.extend)
    print("loop", info, flush=True)
    so, sg = o.state(), d.state()
    assert_bitwise(sg["r"], so["r"], "r")
    assert_bitwise(sg["v"], so["v"], "v")
    assert np.array_equal(sg["id"], so["id"])
    assert o.counters() == d.counters()
    events_equal(ev_o, ev_g, "fused padded events")
