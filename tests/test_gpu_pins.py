"""Hand-built answers for the library's event pipeline, independent of the oracle.

The same inputs as the oracle's pins (test_oracle_pins.py), run through the C ABI:
compaction of inactive bodies of every class (red_check_rebuild, pp_disk.cu:1324-1409,
number_of_bodies.cpp:27-104); a planted collision merged with mass, momentum and volume
conserved and the merger parked (red_check_collision, pp_disk.cu:1151-1202), its raw records
reduced to the earliest one (create_sp_events, pp_disk.cu:1111-1149); and a non-finite error
estimate reported as RED_ERR_NAN (the reference's CPU error loop passes NaN over,
integrator.cu:237-253; the library refuses to step on it).
"""
import math

import numpy as np
import pytest

from test_oracle_pins import COMPACTED_COUNTS, COMPACTED_IDS, compaction_disk

pytestmark = pytest.mark.gpu


def test_compaction_by_hand(native_mod):
    print("case compaction by hand", flush=True)
    d = native_mod.Disk(compaction_disk())
    assert d.check_for_rebuild_vectors(0)
    s = d.state()
    assert d.counts() == COMPACTED_COUNTS
    assert list(s["id"]) == COMPACTED_IDS
    assert list(s["r"][:, 0]) == [float(k) for k in COMPACTED_IDS]
    assert list(s["type"]) == [0, 1, 3, 3, 4, 5, 5, 5, 6, 6]
    assert list(d.rebuild_map()) == [k - 1 for k in COMPACTED_IDS]
    # the compacted disk still steps
    d.step()
    assert d.check_for_rebuild_vectors(0)
    assert list(d.state()["id"]) == COMPACTED_IDS
    d.close()


def _pair_disk():
    """Star + two protoplanets 1e-6 AU apart, moving together (their radii 3e-6 and 5e-6 AU
    overlap), on a 1 AU orbit."""
    from rednative import files
    v = math.sqrt(1.0)
    m1, m2 = 2.0 ** -20, 3.0 * 2.0 ** -20
    r = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0 + 1e-6, 0, 0]])
    vv = np.array([[0.0, 0, 0], [0, v, 0], [0, v, 0]])
    return files.BodySet([1, 0, 0, 2, 0, 0, 0], np.array([1, 2, 3], dtype=np.int32), ["s", "a", "b"],
                         np.array([0, 3, 3], dtype=np.int32), np.zeros(3), np.array([1.0, m1, m2]),
                         np.array([0.00465, 3.0e-6, 5.0e-6]), np.array([1.0e6, 2.0e6, 3.0e6]),
                         np.array([0.0, 0.7, 0.2]), np.zeros(3, dtype=np.int32), np.zeros(3), r, vv)


def test_collision_merge_by_hand(native_mod):
    print("case planted collision merge by hand", flush=True)
    b = _pair_disk()
    d = native_mod.Disk(b, integrator=2, adaptive=False, dt_day=0.01, radii_enhance_factor=1.0)
    t0 = d.t
    d.step()
    raw = d.raw_events()
    assert len(raw) >= 2 and all((e["id1"], e["id2"]) == (3, 2) for e in raw)   # heavier body survives
    assert d.check_for_collision()
    sp = d.sp_events()
    assert len(sp) == 1
    e = sp[0]
    assert e["t"] == min(x["t"] for x in raw) == t0 and e["stage"] == 0
    s = d.state()
    m1, m2 = b.mass[1], b.mass[2]
    M, R, rho, cd = s["p"][2]
    assert M == m1 + m2 and cd == b.cd[2]
    for k in range(3):
        assert s["r"][2][k] == (m2 * e["r1"][k] + m1 * e["r2"][k]) / M
        assert s["v"][2][k] == (m2 * e["v1"][k] + m1 * e["v2"][k]) / M
    assert abs(R ** 3 / (b.radius[1] ** 3 + b.radius[2] ** 3) - 1.0) < 1e-14
    assert abs(rho * 4.0 / 3.0 * math.pi * R ** 3 / M - 1.0) < 1e-14
    assert list(s["r"][1]) == [1.0e20, 0.0, 0.0] and list(s["v"][1]) == [0.0, 0.0, 0.0]
    assert list(s["p"][1][:3]) == [0.0, 0.0, 0.0] and list(s["id"]) == [1, -2, 3]
    assert d.counters()["n_collision"] == 1
    d.close()


def test_nan_error_estimate_reported(native_mod):
    print("case NaN error estimate", flush=True)
    b = _pair_disk()
    b.r[2] = [1.5, 0.0, 0.0]
    b.v[1] = [0.0, float("nan"), 0.0]
    d = native_mod.Disk(b, integrator=2, adaptive=True, tolerance=1e-10, dt_day=0.01)
    with pytest.raises(native_mod.RedError) as ex:
        d.step()
    assert ex.value.code == 4   # RED_ERR_NAN
    d.close()
