"""GPU force sweeps through the C ABI (red_accel) against the CPU oracle.

FP64 mode must equal the oracle bit for bit (same operation order, no fused
multiply-adds, correctly rounded sqrt and division on gfx950), for every launch path:
broadcast and LDS-tile sweeps, 1/2/4 sinks per work-item, the split path used for
sweeps with few sinks, and the producer/consumer sweep (the FP64 default).  Mixed mode must stay within a relative tolerance of 2e-6 per
body of the FP64 acceleration (FP32 r^-3 carries about 1e-7 relative error per pair)
and must produce the same event list as FP64 on the same state.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, native_mod, b, **kw):
    okw = dict(radii_enhance_factor=kw.get("radii_enhance_factor", 0.0),
               encounter_hill_factor=kw.get("encounter_hill_factor", 0.0),
               close_encounter=kw.get("close_encounter", False))
    o = oracle_mod.Oracle(b, **okw)
    d = native_mod.Disk(b, **kw)
    return o, d


def _events_equal(ev_o, ev_g):
    assert len(ev_o) == len(ev_g), (len(ev_o), len(ev_g))
    for eo, eg in zip(ev_o, ev_g):
        for k in ("name", "id1", "idx1", "id2", "idx2", "stage", "sink", "source"):
            assert eo[k] == eg[k], (k, eo, eg)
        for k in ("t", "d", "r1", "v1", "r2", "v2"):
            assert_bitwise(np.array(eo[k], dtype=np.float64), np.array(eg[k], dtype=np.float64), k)


@pytest.mark.parametrize("spt,strategy,block", [(1, 1, 256), (2, 1, 256), (4, 1, 128), (1, 2, 256), (2, 2, 64),
                                                (1, 3, 256), (4, 3, 512), (8, 3, 1024), (16, 3, 256),
                                                (16, 3, 1024)])
def test_fp64_allsi_bit_exact(oracle_mod, native_mod, spt, strategy, block):
    print(f"case allsi n=3001 spt={spt} strategy={strategy} block={block}", flush=True)
    b = generate.to_internal(generate.disk_allsi(3001, seed=11))
    o, d = _pair(oracle_mod, native_mod, b, sinks_per_thread=spt, strategy=strategy, block=block)
    a_o = o.accel(0.5, b.r, b.v)
    a_g = d.accel(0.5, b.r, b.v)
    assert_bitwise(a_g, a_o, "accel")


def test_fp64_typed_disk_bit_exact(oracle_mod, native_mod):
    print("case disk_small n=256", flush=True)
    b = generate.to_internal(generate.disk_small(253, seed=101))
    o, d = _pair(oracle_mod, native_mod, b)
    assert_bitwise(d.accel(1.0, b.r, b.v), o.accel(1.0, b.r, b.v), "accel")


@pytest.mark.parametrize("strategy", [1, 3, 4])
def test_fp64_split_path_bit_exact(oracle_mod, native_mod, strategy):
    # 3 SI sinks against 6003 sources and 3000 NI sinks: with the broadcast strategy the SI
    # sweep takes the split path; the producer/consumer strategy takes every sweep; strategy 4
    # sweeps the NSI and NI sinks with one work-item per sink
    print(f"case disk_mixed 3+3000+3000 strategy={strategy}", flush=True)
    b = generate.to_internal(generate.disk_mixed(n_pl=3000, n_tp=3000, seed=5))
    o, d = _pair(oracle_mod, native_mod, b, strategy=strategy)
    assert_bitwise(d.accel(2.0, b.r, b.v), o.accel(2.0, b.r, b.v), "accel")


def test_fp64_inactive_sources_skipped(oracle_mod, native_mod):
    b = generate.to_internal(generate.disk_allsi(600, seed=3))
    b.ids[[5, 77, 300]] *= -1           # inactive bodies keep their mass and position
    print("case inactive n=600", flush=True)
    o, d = _pair(oracle_mod, native_mod, b)
    a_g = d.accel(0.0, b.r, b.v)
    assert_bitwise(a_g, o.accel(0.0, b.r, b.v), "accel")
    assert np.all(a_g[[5, 77, 300]] == 0.0)


def test_fp64_events_bit_exact(oracle_mod, native_mod):
    # a large collision factor makes many pairs collide; the list must match in canonical order
    b = generate.to_internal(generate.disk_allsi(1500, seed=21, a_range=(0.9, 1.1)))
    print("case events n=1500", flush=True)
    o, d = _pair(oracle_mod, native_mod, b, radii_enhance_factor=3000.0, event_capacity=200000)
    a_o = o.accel(3.0, b.r, b.v, stage=2)
    a_g = d.accel(3.0, b.r, b.v, stage=2)
    assert_bitwise(a_g, a_o, "accel")
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert len(ev_o) > 50
    _events_equal(ev_o, ev_g)


def test_fp64_hill_events_bit_exact(oracle_mod, native_mod):
    # protoplanets are self-interacting, so their pairs meet the Hill test inside the sweep
    b = generate.to_internal(generate.disk_allsi(600, seed=8, a_range=(1.0, 1.2), m_range=(1e-11, 1e-9)))
    print("case hill n=600", flush=True)
    o, d = _pair(oracle_mod, native_mod, b, radii_enhance_factor=1.0, encounter_hill_factor=30.0,
                 event_capacity=100000)
    assert_bitwise(d.accel(0.0, b.r, b.v), o.accel(0.0, b.r, b.v), "accel")
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert len(ev_o) > 0
    _events_equal(ev_o, ev_g)


def test_mixed_hill_events_match_fp64(oracle_mod, native_mod):
    # the mixed screen must keep every pair the FP64 Hill test reports (protoplanets, so
    # that mutual Hill radii are large) and recheck it exactly
    b = generate.to_internal(generate.disk_allsi(3001, seed=12, m_range=(1e-9, 1e-7)))
    kw = dict(radii_enhance_factor=1.0, encounter_hill_factor=3.0)
    print(f"case mixed hill n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, precision=native_mod.PREC_MIXED, event_capacity=100000, **kw)
    d.accel(1.0, b.r, b.v)
    o.accel(1.0, b.r, b.v)
    ev_o, ev_g = o.raw_events(), d.raw_events()
    print("events", len(ev_o), flush=True)
    assert len(ev_o) > 0
    _events_equal(ev_o, ev_g)


@pytest.mark.parametrize("parts", [0, 3, 16])
def test_fp64_split_sums_close(oracle_mod, native_mod, parts):
    """FP64 with source-partitioned sums: same terms, different summation order."""
    b = generate.to_internal(generate.disk_mixed(n_pl=3000, n_tp=5000, seed=6))
    print(f"case fp64_split parts={parts}", flush=True)
    o = oracle_mod.Oracle(b)
    d = native_mod.Disk(b, fp64_split=True, source_parts=parts)
    a_o = o.accel(0.0, b.r, b.v)
    a_g = d.accel(0.0, b.r, b.v)
    rel = np.linalg.norm(a_g - a_o, axis=1) / np.maximum(np.linalg.norm(a_o, axis=1), 1e-300)
    assert rel.max() < 1e-13, rel.max()


@pytest.mark.parametrize("kahan", [False, True])
def test_mixed_accuracy_and_events(oracle_mod, native_mod, kahan):
    b = generate.to_internal(generate.disk_mixed(n_pl=4000, n_tp=4000, seed=9, a_range=(0.9, 1.3)))
    print(f"case mixed kahan={kahan} n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, radii_enhance_factor=400.0)
    d = native_mod.Disk(b, precision=native_mod.PREC_MIXED, kahan_pairs=kahan, radii_enhance_factor=400.0,
                        event_capacity=100000)
    a_o = o.accel(1.0, b.r, b.v)
    a_g = d.accel(1.0, b.r, b.v)
    rel = np.linalg.norm(a_g - a_o, axis=1) / np.maximum(np.linalg.norm(a_o, axis=1), 1e-300)
    print("max rel err", rel.max(), "median", np.median(rel))
    assert rel.max() < 2e-6
    ev_o, ev_g = o.raw_events(), d.raw_events()
    assert len(ev_o) > 0
    _events_equal(ev_o, ev_g)



def _extreme_disk(m_lo, m_hi, mirrored=False, seed=77):
    """Star + four groups of 40 bodies at separations of 1e-45, 1e-30, 1 and 1.5e45 AU with
    log-uniform masses in [m_lo, m_hi].  mirrored: every body has a twin at -r with the
    same mass next to it, so the barycentre is exactly the origin and the barycentric
    transform leaves the positions unchanged."""
    from rednative import files
    rng = np.random.default_rng(seed)
    groups = [
        (1e-45, 40),    # mutual r^2 ~ 1e-90, around 2^-300
        (1e-30, 40),    # r^2 ~ 1e-60: inside
        (1.0, 40),      # ordinary
        (1.5e45, 40),   # r^2 ~ 1e90 .. 1e91, around 2^300
    ]
    r = [np.array([0.0, 0.0, 0.0]) if mirrored else np.array([3.0, 0.0, 0.0])]
    for scale, k in groups:
        if mirrored:
            for x in rng.uniform(-1.0, 1.0, (k // 2, 3)) * scale:
                r += [x, -x]
        else:
            r += list(rng.uniform(-1.0, 1.0, (k, 3)) * scale)
    r = np.array(r)
    n = len(r)
    m = np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), n))
    if mirrored:
        m[2::2] = m[1::2]
    m[0] = 1.0
    types = np.full(n, 3, dtype=np.int32); types[0] = 0
    counts = [1, 0, 0, n - 1, 0, 0, 0]
    return files.BodySet(counts, np.arange(1, n + 1, dtype=np.int32), [f"b{k}" for k in range(n)], types,
                         np.zeros(n), m, np.full(n, 1e-300), np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32),
                         np.zeros(n), r, np.zeros((n, 3)))


@pytest.mark.parametrize("m_lo", [1e-95, 1e-20])
def test_fp64_extreme_scales_step_bit_exact(oracle_mod, native_mod, m_lo):
    """An Euler step sweeps the context's own state, whose source masses the host checks
    once: with masses down to 1e-95 (below 2^-300) every pair tests its mass, with masses
    in [1e-20, 1e20] only the separations are tested.  Either way the step equals the
    oracle's bit for bit."""
    b = _extreme_disk(m_lo, 1e20, mirrored=True)
    print(f"case extreme scales step m_lo={m_lo} n={b.n}", flush=True)
    o = oracle_mod.Oracle(b, integrator=0, dt_day=0.01)
    so0 = o.state()
    assert_bitwise(so0["r"], b.r, "barycentric transform is the identity")
    for spt, block in ((4, 512), (16, 1024)):
        o = oracle_mod.Oracle(b, integrator=0, dt_day=0.01)
        d = native_mod.Disk(b, integrator=0, dt_day=0.01, strategy=3, sinks_per_thread=spt, block=block)
        assert o.step() == d.step()
        so, sg = o.state(), d.state()
        assert np.all(np.isfinite(so["v"])) and np.any(so["v"] != 0.0)
        assert_bitwise(sg["v"], so["v"], f"v sg={spt}")
        assert_bitwise(sg["r"], so["r"], f"r sg={spt}")
        d.close()


def test_fp64_extreme_scales_bit_exact(oracle_mod, native_mod):
    """Separations and masses across and around 2^-300 .. 2^300, where the producer/consumer
    sweep switches between its core square-root / quotient sequences and the full IEEE ones
    (per wave): every result must still equal the oracle's bit for bit."""
    from rednative import files
    rng = np.random.default_rng(77)
    groups = [
        (1e-45, 40),    # mutual r^2 ~ 1e-90, around 2^-300
        (1e-30, 40),    # r^2 ~ 1e-60: inside
        (1.0, 40),      # ordinary
        (1.5e45, 40),   # r^2 ~ 1e90 .. 1e91, around 2^300
    ]
    r = [np.array([3.0, 0.0, 0.0])]
    for scale, k in groups:
        r += list(rng.uniform(-1.0, 1.0, (k, 3)) * scale)
    r = np.array(r)
    n = len(r)
    m = np.exp(rng.uniform(np.log(1e-95), np.log(1e20), n))   # some below 2^-300
    m[0] = 1.0
    types = np.full(n, 3, dtype=np.int32); types[0] = 0
    counts = [1, 0, 0, n - 1, 0, 0, 0]
    b = files.BodySet(counts, np.arange(1, n + 1, dtype=np.int32), [f"b{k}" for k in range(n)], types,
                      np.zeros(n), m, np.full(n, 1e-300), np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32),
                      np.zeros(n), r, np.zeros((n, 3)))
    print(f"case extreme scales n={n}", flush=True)
    o = oracle_mod.Oracle(b)
    a_o = o.accel(0.0, b.r, b.v)
    assert np.all(np.isfinite(a_o))
    for spt, block in ((4, 512), (8, 512), (16, 1024)):
        d = native_mod.Disk(b, sinks_per_thread=spt, block=block, strategy=3)
        assert_bitwise(d.accel(0.0, b.r, b.v), a_o, f"accel sg={spt}")
        d.close()


def test_fp64_one_per_sink_inactive_sources(oracle_mod, native_mod):
    """One work-item per sink (strategy 4) with inactive sources (a giant and planetesimals,
    id < 0, mass and position kept) in the groups of four that the sweep takes at once, and
    the same sweeps through the producer/consumer kernels: both equal the oracle."""
    b = generate.to_internal(generate.disk_mixed(n_pl=2500, n_tp=3001, seed=9))
    b.ids[[1, 10, 11, 12, 13, 200, 1999]] *= -1
    print("case one-per-sink inactive sources", flush=True)
    o = oracle_mod.Oracle(b)
    a_o = o.accel(0.25, b.r, b.v)
    for strategy in (4, 3):
        d = native_mod.Disk(b, strategy=strategy)
        assert_bitwise(d.accel(0.25, b.r, b.v), a_o, f"accel strategy={strategy}")
        d.close()


def test_fp64_one_per_sink_extreme_scales(oracle_mod, native_mod):
    """One work-item per sink across and around 2^-300 .. 2^300: massive sources and test
    particles at the same scales, source masses from 1e-95 to 1e20, so that chunks leave the
    cores' range through r^2 and through masses and are formed again with sqrt() and the
    division operator.  Equals the oracle bit for bit."""
    from rednative import files
    rng = np.random.default_rng(78)
    groups = [(1e-45, 40), (1e-30, 40), (1.0, 40), (1.5e45, 40)]
    r_src, r_tp = [np.array([3.0, 0.0, 0.0])], []
    for scale, k in groups:
        r_src += list(rng.uniform(-1.0, 1.0, (k, 3)) * scale)
        r_tp += list(rng.uniform(-1.0, 1.0, (k, 3)) * scale)
    r = np.array(r_src + r_tp)
    n, ns = len(r), len(r_src)
    m = np.zeros(n)
    m[:ns] = np.exp(rng.uniform(np.log(1e-95), np.log(1e20), ns))
    m[0] = 1.0
    types = np.full(n, 6, dtype=np.int32)
    types[0] = 0
    types[1:ns] = 3
    counts = [1, 0, 0, ns - 1, 0, 0, n - ns]
    b = files.BodySet(counts, np.arange(1, n + 1, dtype=np.int32), [f"b{k}" for k in range(n)], types,
                      np.zeros(n), m, np.full(n, 1e-300), np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32),
                      np.zeros(n), r, np.zeros((n, 3)))
    print(f"case one-per-sink extreme scales n={n}", flush=True)
    o = oracle_mod.Oracle(b)
    a_o = o.accel(0.0, b.r, b.v)
    assert np.all(np.isfinite(a_o))
    d = native_mod.Disk(b, strategy=4)
    assert_bitwise(d.accel(0.0, b.r, b.v), a_o, "accel")
    d.close()
