"""The oracle's astro-centric frame against what the mechanics fixes, and the library's
host-side shift into that frame.

Astro-centric force (pp_disk.cu:409-454 with the bounds of number_of_bodies.cpp:196-222):
in coordinates centred on the star, body i accelerates by its inertial acceleration minus
the star's, both from the plain all-pairs sum that includes the star.  The barycentric-frame
sweep of the oracle computes that plain sum, so a_astro[i] = a_bary[i] - a_bary[0] must hold
to rounding.  A dropped indirect term, a wrong sign, m_i missing from the star's pull or the
star left inside the sweep bounds each break it by far more than rounding.

Two-body check: with the star and one planet the astro-centric equations are Kepler's
problem with mu = m_star + m_planet, so the orbital energy and angular momentum stay
constant along an RKF7(8) run.

Shift (tools.cpp:180-236): every body minus body 0's position and velocity, body 0 being
the star; the library's red_host_transform_to_ac must equal the oracle's bit for bit and
refuse a list whose body 0 is not the star.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate


def _helio_disk(n, seed, **kw):
    # the generator places the star at rest at the origin: astro-centric coordinates already
    return generate.to_internal(generate.disk_allsi(n, seed=seed, **kw), transform_bc=False)


def test_astro_accel_is_relative_inertial_accel(oracle_mod):
    b = _helio_disk(400, seed=5, m_range=(1e-9, 1e-6))
    a_astro = oracle_mod.Oracle(b, frame_center=0).accel(0.0, b.r, b.v)
    a_bary = oracle_mod.Oracle(b, frame_center=1).accel(0.0, b.r, b.v)
    expect = a_bary - a_bary[0]
    assert np.all(a_astro[0] == 0.0)
    scale = np.linalg.norm(a_bary[1:], axis=1)
    rel = np.linalg.norm(a_astro[1:] - expect[1:], axis=1) / scale
    # the indirect terms alone are ~1e-5 of the star's pull here
    ind = np.linalg.norm(a_bary[0]) / scale.max()
    print("max rel", rel.max(), "indirect share", ind)
    assert ind > 1e-7
    assert rel.max() < 1e-12


def test_astro_inactive_body_keeps_star_pull(oracle_mod):
    b = _helio_disk(200, seed=6)
    b.ids[[3, 50]] *= -1
    a = oracle_mod.Oracle(b, frame_center=0).accel(0.0, b.r, b.v)
    for i in (3, 50):
        mu = b.mass[0] + b.mass[i]
        r2 = (b.r[i, 0] * b.r[i, 0] + b.r[i, 1] * b.r[i, 1]) + b.r[i, 2] * b.r[i, 2]
        assert_bitwise(a[i], -mu / (r2 * np.sqrt(r2)) * b.r[i], f"inactive {i}")


def test_astro_two_body_is_kepler(oracle_mod):
    b = _helio_disk(2, seed=9, m_range=(1e-3, 1e-3))     # star + one Jupiter-mass body
    assert b.n == 2
    mu = b.mass[0] + b.mass[1]
    o = oracle_mod.Oracle(b, frame_center=0, integrator=4, adaptive=True, tolerance=1e-13, dt_day=1.0)

    def invariants():
        s = o.state()
        r, v = s["r"][1], s["v"][1]
        return 0.5 * v @ v - mu / np.linalg.norm(r), np.cross(r, v)

    e0, h0 = invariants()
    for _ in range(400):
        o.step()
    s = o.state()
    assert np.all(s["r"][0] == 0.0) and np.all(s["v"][0] == 0.0)
    e1, h1 = invariants()
    # integration error leaves ~1e-11; taking mu = m_star alone would swing the energy by
    # ~m_planet e / a ~ 1e-5 of it along this e ~ 0.01 orbit
    print("energy", (e1 - e0) / e0, "angular momentum", np.linalg.norm(h1 - h0) / np.linalg.norm(h0))
    assert abs((e1 - e0) / e0) < 1e-9
    assert np.linalg.norm(h1 - h0) / np.linalg.norm(h0) < 1e-9


def test_host_transform_to_ac_matches_oracle(oracle_mod, native_mod):
    b = generate.disk_small(60, seed=8)
    rng = np.random.default_rng(1)
    b.r = b.r + rng.normal(0.0, 1e-2, 3)                  # star off the origin
    b.v = b.v + rng.normal(0.0, 1e-3, 3)
    o = oracle_mod.Oracle(b, frame_center=0)
    o.transform_to_ac()
    so = o.state()
    r, v = b.r.copy(), b.v.copy()
    native_mod.host_transform_to_ac(b.types, r, v)
    assert_bitwise(r, so["r"], "r")
    assert_bitwise(v, so["v"], "v")
    assert np.all(r[0] == 0.0) and np.all(v[0] == 0.0)
    assert_bitwise(r[7], b.r[7] - b.r[0], "r7")


def test_host_transform_to_ac_needs_star_first(native_mod):
    b = generate.disk_small(10, seed=8)
    t = b.types.copy()
    t[[0, 1]] = t[[1, 0]]
    with pytest.raises(native_mod.RedError) as e:
        native_mod.host_transform_to_ac(t, b.r.copy(), b.v.copy())
    assert e.value.code == 6   # RED_ERR_NO_STAR
