"""Hand-built answers for the oracle's event pipeline, error norm, step controller and
compaction.

Each function below is restated in the oracle from the reference and the product follows
the same rules; lock-step equality of the two cannot catch a rule both get wrong, so these
tests fix the rules themselves with inputs whose answers are written out by hand:

* create_sp_events (pp_disk.cu:1111-1149): one record per ordered (id1, id2), the earliest t
  kept, the first of equal times, in first-occurrence order;
* the error norm (integrator.cu:211-256, CPU branch): the largest signed component of either
  vector starting from 0, scaled by |dt * lambda|; NaN components never compare larger;
* the step controller dt <- 0.9 dt (tol/err)^(1/e) (int_rungekutta4.cu:253, RK5 exponent
  1/5, int_rungekutta8.cu:777): on a circular orbit its fixed point is err = 0.9^e tol, so
  the accepted errors settle there (a wrong safety factor or exponent moves it by >= 10 %);
* handle_collision_pair (pp_disk.cu:1151-1202, tools.cpp:238-290): mass, momentum and
  volume conserved, the merger parked at (1e20, 0, 0) at rest, massless, id negated;
* compaction (pp_disk.cu:1324-1409, number_of_bodies.cpp:27-104): inactive bodies of every
  class removed, the rest kept in order, class counts updated.
"""
import math

import numpy as np
import pytest

from rednative import files

EV_COLLISION = 4


def _event(id1, id2, t, tag, name=EV_COLLISION):
    import pyoracle
    e = pyoracle.OrcEvent()
    e.name, e.id1, e.id2, e.t, e.sink = name, id1, id2, t, tag
    return e


def test_create_sp_events_by_hand(oracle_mod):
    raw = [
        _event(5, 9, 0.30, 0),
        _event(2, 7, 0.10, 1),
        _event(5, 9, 0.20, 2),   # a later record with an earlier t replaces record 0 in place
        _event(9, 5, 0.05, 3),   # the reversed pair is another key
        _event(2, 7, 0.10, 4),   # equal t: record 1 stays
        _event(5, 9, 0.20, 5),   # equal t: record 2 stays
        _event(3, 4, 0.50, 6),
    ]
    sp = oracle_mod.create_sp_events(raw)
    assert [(e.id1, e.id2, e.t, e.sink) for e in sp] == [(5, 9, 0.20, 2), (2, 7, 0.10, 1), (9, 5, 0.05, 3),
                                                         (3, 4, 0.50, 6)]
    assert oracle_mod.create_sp_events([]) == []
    one = oracle_mod.create_sp_events([_event(1, 2, -3.0, 9)])
    assert [(e.id1, e.id2, e.t, e.sink) for e in one] == [(1, 2, -3.0, 9)]


def test_max_error_by_hand(oracle_mod):
    me = oracle_mod.max_error
    # signed maximum from 0: all-negative vectors give 0
    assert me([-1.0, -2.0, -3.0], [-0.5, -7.0, -1e300], 3, 1.0, 1.0) == 0.0
    # the largest signed component (2.0), not the largest magnitude (-3.0, -7.0);
    # |dt * lambda| with a negative dt: 0.25 * 2 = 0.5
    assert me([-3.0, 2.0, -1.0, 0.5], [1.5, -7.0, 0.0, 0.0], 4, -0.5, 0.5) == 0.5
    # the velocity vector can hold the maximum
    assert me([0.25, 0.0], [0.0, 4.0], 2, 1.0, 0.125) == 0.5
    # only the first n_var components count
    assert me([1.0, 100.0], [0.0, 100.0], 1, 1.0, 1.0) == 1.0
    # NaN components never compare larger: they are passed over
    assert me([float("nan"), 1.0], [0.5, float("nan")], 2, 2.0, 1.0) == 2.0
    assert me([float("nan")], [float("nan")], 1, 1.0, 1.0) == 0.0


def _two_body(m2=1.0e-3, a=1.0):
    """Star and one planet on a circular orbit, barycentric, internal units (G = 1)."""
    v = math.sqrt((1.0 + m2) / a)
    r = np.array([[0.0, 0.0, 0.0], [a, 0.0, 0.0]])
    vv = np.array([[0.0, 0.0, 0.0], [0.0, v, 0.0]])
    M = 1.0 + m2
    r -= m2 * r[1] / M
    vv -= m2 * vv[1] / M
    return files.BodySet([1, 1, 0, 0, 0, 0, 0], np.array([1, 2], dtype=np.int32), ["star", "planet"],
                         np.array([0, 1], dtype=np.int32), np.zeros(2), np.array([1.0, m2]),
                         np.array([0.00465, 1e-4]), np.zeros(2), np.zeros(2), np.zeros(2, dtype=np.int32),
                         np.zeros(2), r, vv)


@pytest.mark.parametrize("integ,tol,expo", [(2, 1e-10, 4), (3, 1e-12, 5), (4, 1e-14, 8)])
def test_controller_fixed_point_on_circular_orbit(oracle_mod, integ, tol, expo):
    o = oracle_mod.Oracle(_two_body(), integrator=integ, adaptive=True, tolerance=tol, dt_day=1.0)
    ratios = []
    for k in range(200):
        o.step()
        if k >= 50:
            ratios.append(o.last_err() / tol)
    med = float(np.median(ratios))
    assert abs(med / 0.9 ** expo - 1.0) < 0.03, (med, 0.9 ** expo)
    # neighbouring exponents and a 0.95 safety factor are excluded
    for other in (0.9 ** (expo - 1), 0.9 ** (expo + 1), 0.95 ** expo):
        assert abs(med / other - 1.0) > 0.05


def _merge_disk(second_type=3):
    """Star + two bodies; body 1 (2^-20 M_sun) will absorb body 2 (3 * 2^-20 M_sun)."""
    m1, m2 = 2.0 ** -20, 3.0 * 2.0 ** -20
    types = np.array([0, 3, second_type], dtype=np.int32)
    counts = [1, 0, 0, 2, 0, 0, 0] if second_type == 3 else [1, 0, 0, 1, 1, 0, 0]
    return files.BodySet(counts, np.array([1, 2, 3], dtype=np.int32), ["s", "a", "b"], types, np.zeros(3),
                         np.array([1.0, m1, m2]), np.array([0.00465, 3.0e-6, 5.0e-6]),
                         np.array([1.0e6, 2.0e6, 3.0e6]), np.array([0.0, 0.7, 0.2]),
                         np.zeros(3, dtype=np.int32), np.zeros(3), np.array([[0.0, 0, 0], [5, 5, 5], [6, 6, 6]]),
                         np.zeros((3, 3)))


def _collision_record():
    e = _event(2, 3, 0.0, 0)
    e.idx1, e.idx2 = 1, 2
    e.r1[:] = [1.0, 0.0, 0.0]
    e.r2[:] = [0.0, 1.0, 0.0]
    e.v1[:] = [0.0, 2.0, 0.0]
    e.v2[:] = [4.0, 0.0, 0.0]
    return e


def test_merge_by_hand(oracle_mod):
    b = _merge_disk()
    o = oracle_mod.Oracle(b)
    e = o.handle_collision_pair(_collision_record())
    m1, m2 = b.mass[1], b.mass[2]
    R1, R2 = b.radius[1], b.radius[2]
    s = o.state()
    # survivor: centre of mass of the recorded stage states (dyadic values: exact)
    assert list(s["r"][1]) == [0.25, 0.75, 0.0] and list(e.rs) == [0.25, 0.75, 0.0]
    assert list(s["v"][1]) == [3.0, 0.5, 0.0] and list(e.vs) == [3.0, 0.5, 0.0]
    M, R, rho, cd = s["p"][1]
    assert M == m1 + m2 == 4.0 * 2.0 ** -20
    # momentum: M v_s = m1 v1 + m2 v2
    assert list(M * s["v"][1]) == [m2 * 4.0, m1 * 2.0, 0.0]
    # volume: R^3 = R1^3 + R2^3; density = M / volume
    assert abs(R ** 3 / (R1 ** 3 + R2 ** 3) - 1.0) < 1e-14
    assert abs(rho * 4.0 / 3.0 * math.pi * (R1 ** 3 + R2 ** 3) / M - 1.0) < 1e-14
    assert cd == b.cd[1]
    assert list(e.p1) == [m1, R1, b.density[1], b.cd[1]] and list(e.p2) == [m2, R2, b.density[2], b.cd[2]]
    # merger: parked far away at rest, massless, inactive
    assert list(s["r"][2]) == [1.0e20, 0.0, 0.0] and list(s["v"][2]) == [0.0, 0.0, 0.0]
    assert list(s["p"][2][:3]) == [0.0, 0.0, 0.0]
    assert list(s["id"]) == [1, 2, -3]
    # the star is untouched
    assert list(s["r"][0]) == [0.0, 0.0, 0.0] and s["p"][0][0] == 1.0


@pytest.mark.parametrize("reactivate,expect", [(False, -3), (True, 3)])
def test_merge_of_an_already_negated_merger(oracle_mod, reactivate, expect):
    """A hit-centrum merger's id was negated by the detection pass: the GPU reference's
    merge negates a stale host copy and leaves it inactive; its CPU path negates twice."""
    b = _merge_disk()
    b.ids[2] = -3
    o = oracle_mod.Oracle(b, hc_reactivate=reactivate)
    o.handle_collision_pair(_collision_record(), kernel_negated=True)
    assert o.state()["id"][2] == expect


def test_merge_with_super_planetesimal_refused(oracle_mod):
    o = oracle_mod.Oracle(_merge_disk(second_type=4))
    with pytest.raises(RuntimeError, match="super-planetesimal is not yet implemented"):
        o.handle_collision_pair(_collision_record())


def compaction_disk():
    """16 bodies over all seven classes (counts 1 2 1 3 2 4 3), ids 1..16, body k at
    (k, 0, 0); bodies 3, 4, 6, 9, 12 and 16 inactive (one of each non-star class, the
    whole rocky class and the last body)."""
    counts = [1, 2, 1, 3, 2, 4, 3]
    types = np.repeat(np.arange(7, dtype=np.int32), counts)
    n = 16
    ids = np.arange(1, n + 1, dtype=np.int32)
    ids[[2, 3, 5, 8, 11, 15]] *= -1
    r = np.zeros((n, 3)); r[:, 0] = np.arange(1, n + 1)
    m = np.where(types == 6, 0.0, 1e-9); m[0] = 1.0
    return files.BodySet(counts, ids, [f"b{k}" for k in range(n)], types, np.zeros(n), m, np.full(n, 1e-6),
                         np.ones(n), np.zeros(n), np.zeros(n, dtype=np.int32), np.zeros(n), r, np.zeros((n, 3)))


COMPACTED_IDS = [1, 2, 5, 7, 8, 10, 11, 13, 14, 15]
COMPACTED_COUNTS = [1, 1, 0, 2, 1, 3, 2]


def test_compaction_by_hand(oracle_mod):
    o = oracle_mod.Oracle(compaction_disk())
    assert o.check_for_rebuild_vectors(0)
    s = o.state()
    assert o.counts() == COMPACTED_COUNTS
    assert list(s["id"]) == COMPACTED_IDS
    assert list(s["r"][:, 0]) == [float(k) for k in COMPACTED_IDS]
    assert list(s["type"]) == [0, 1, 3, 3, 4, 5, 5, 5, 6, 6]
    # nothing left to remove: a second pass keeps everything
    assert o.check_for_rebuild_vectors(0)
    assert list(o.state()["id"]) == COMPACTED_IDS


# ---- padded storage (-ups, pp_disk.cu:144-215) -------------------------------------------

def _padded_disk(x_pos, x_radius=1.0e-6, x_mass=3.0e-6, x_id=-3):
    """Star, protoplanet P, inactive protoplanet X (its id negated, as after an ejection: mass
    and radius kept), test particle T."""
    types = np.array([0, 3, 3, 6], dtype=np.int32)
    r = np.array([[0.0, 0, 0], [1.0, 0, 0], x_pos, [0.0, -1.0, 0]])
    return files.BodySet([1, 0, 0, 2, 0, 0, 1], np.array([1, 2, x_id, 4], dtype=np.int32), ["s", "p", "x", "t"],
                         types, np.zeros(4), np.array([1.0, 1.0e-6, x_mass, 0.0]),
                         np.array([0.00465, 1.0e-6, x_radius, 0.0]), np.zeros(4), np.zeros(4),
                         np.zeros(4, dtype=np.int32), np.zeros(4), r, np.zeros((4, 3)))


def _pull(src_m, src_r, at):
    d = np.asarray(src_r, float) - np.asarray(at, float)
    return src_m * d / np.linalg.norm(d) ** 3


def test_padded_inactive_body_keeps_pulling_and_being_pulled(oracle_mod):
    """Under padded storage an inactive body's stored mass still attracts every sink and the
    body itself is accelerated by the other sources; without it the body neither pulls nor
    moves.  So padded - plain is X's pull on S, P and T, and X's own acceleration is the pull
    of S and P (a dropped source or sink term, or a zeroed mass, fails)."""
    b = _padded_disk([0.0, 1.5, 0.0])
    a_plain = oracle_mod.Oracle(b).accel(0.0, b.r, b.v)
    a_pad = oracle_mod.Oracle(b, padded=True).accel(0.0, b.r, b.v)
    assert np.array_equal(a_plain[2], np.zeros(3))
    for i in (0, 1, 3):
        np.testing.assert_allclose(a_pad[i] - a_plain[i], _pull(3.0e-6, b.r[2], b.r[i]), rtol=1e-9, atol=1e-21)
    np.testing.assert_allclose(a_pad[2], _pull(1.0, b.r[0], b.r[2]) + _pull(1.0e-6, b.r[1], b.r[2]), rtol=1e-14)


def test_padded_inactive_body_meets_pair_tests(oracle_mod):
    """The padded kernel tests pairs without ids: X 0.5e-6 AU from P (within 1 x (R_P + R_X))
    is recorded as a collision, the heavier X the survivor with its negated id; the plain
    kernel skips the inactive X."""
    b = _padded_disk([1.0, 0.5e-6, 0.0])
    o = oracle_mod.Oracle(b, radii_enhance_factor=1.0)
    o.accel(0.0, b.r, b.v)
    assert o.raw_events() == []
    o = oracle_mod.Oracle(b, radii_enhance_factor=1.0, padded=True)
    o.accel(0.0, b.r, b.v)
    ev = o.raw_events()
    assert [(e["name"], e["id1"], e["id2"], e["sink"], e["source"]) for e in ev] == [(EV_COLLISION, -3, 2, 2, 1)]
    assert ev[0]["d"] == 0.5e-6


def test_padded_merged_body_adds_nothing(oracle_mod):
    """A merged body (massless, zero radius, parked at (1e20, 0, 0)) changes no active sink's
    acceleration in either mode, bit for bit: zero-mass terms add exact zeros.  This is why
    the massless padding particles of the padded layout need not exist here."""
    b = _padded_disk([1.0e20, 0.0, 0.0], x_radius=0.0, x_mass=0.0)
    a_plain = oracle_mod.Oracle(b).accel(0.0, b.r, b.v)
    a_pad = oracle_mod.Oracle(b, padded=True).accel(0.0, b.r, b.v)
    for i in (0, 1, 3):
        assert np.array_equal(a_pad[i], a_plain[i]), i
