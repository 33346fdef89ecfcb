"""Known answers and mathematical checks that pin the oracle.

* red.cuda/src/red.cuda.utility/red_test.cpp:483-840: barycentre of the 4-body unit
  square, radius from mass and density, merge position / velocity / properties,
  Kepler equation (e = 0.5, M = 27 deg -> E = 48.43417991487915 deg,
  M = 225 deg -> 210.47211284530107 deg).
* Convergence order of every integrator on a two-body orbit against the analytic
  Kepler solution, and the order of each embedded error estimate: a dropped or wrong
  tableau entry changes the measured order.
* Gravity against a brute-force numpy sweep written from the reference loop
  (pp_disk.cu:404-572) on tiny disks of every body class, and momentum balance.
"""
import math

import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import files, generate


def _bodies(r, v, m, types=None, radius=None):
    n = len(m)
    types = np.zeros(n, dtype=np.int32) if types is None else np.asarray(types, dtype=np.int32)
    counts = [int((types == t).sum()) for t in range(7)]
    return files.BodySet(counts, np.arange(1, n + 1, dtype=np.int32), [f"b{k}" for k in range(n)], types,
                         np.zeros(n), np.asarray(m, float), np.ones(n) if radius is None else np.asarray(radius, float),
                         np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32), np.zeros(n),
                         np.asarray(r, float), np.asarray(v, float))


def test_red_test_barycentre(oracle_mod):
    r = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    v = [[1, 0, 0], [2, 0, 0], [2, 0, 0], [1, 0, 0]]
    o = oracle_mod.Oracle(_bodies(r, v, [1.0, 1.0, 1.0, 1.0], types=[0, 1, 1, 3]))
    o.transform_to_bc()
    s = o.state()
    np.testing.assert_allclose(s["r"], [[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]], atol=1e-15)
    np.testing.assert_allclose(s["v"], [[-0.5, 0, 0], [0.5, 0, 0], [0.5, 0, 0], [-0.5, 0, 0]], atol=1e-15)


def test_red_test_radius_and_merge(oracle_mod):
    L = oracle_mod.lib()
    assert abs(L.orc_calculate_radius(1.0, 1.0) - 0.62035049089940001666800681204778) <= 1e-15
    out = np.zeros(3)
    dp = lambda a: a.ctypes.data_as(oracle_mod.ctypes.POINTER(oracle_mod.ctypes.c_double))
    L.orc_calc_position_after_collision(1.0, 1.0, dp(np.array([1.0, 1, 1])), dp(np.array([4.0, 3, 1])), dp(out))
    np.testing.assert_allclose(out, [2.5, 2.0, 1.0], atol=1e-15)
    L.orc_calc_position_after_collision(1.0, 1.0, dp(np.array([1.0, 0, 0])), dp(np.array([-1.0, 0, 0])), dp(out))
    np.testing.assert_allclose(out, [0, 0, 0], atol=1e-15)
    L.orc_calc_position_after_collision(1.0, 1.0, dp(np.array([-1.0, 0, 0])), dp(np.array([-1.0, 0, 0])), dp(out))
    np.testing.assert_allclose(out, [-1, 0, 0], atol=1e-15)
    p = np.zeros(4)
    L.orc_calc_physical_properties(dp(np.array([1.0, 1.0, 1.0, 1.0])), dp(np.array([1.0, 1.0, 1.0, 1.0])), dp(p))
    rho = 0.23873241463784300365332564505877
    assert abs(p[0] - 2.0) <= 1e-15 and abs(p[2] - rho) <= 1e-15 and abs(p[3] - 1.0) <= 1e-15
    assert abs(p[1] - L.orc_calculate_radius(2.0, rho)) <= 1e-15
    # volume is conserved: (4pi/3) R^3 rho = M
    assert abs(4.1887902047863909846168578443727 * p[1] ** 3 * p[2] - 2.0) < 1e-14


def test_red_test_kepler(oracle_mod):
    L = oracle_mod.lib()
    E = oracle_mod.ctypes.c_double()
    deg = 1.745329251994329509e-2
    for M, want in ((27.0, 48.43417991487915), (225.0, 210.47211284530107)):
        assert L.orc_kepler_equation_solver(0.5, M * deg, 1.0e-15, oracle_mod.ctypes.byref(E)) == 0
        assert abs(E.value / deg - want) < 1e-13


def test_orbital_elements_round_trip(oracle_mod):
    L = oracle_mod.lib()
    c = oracle_mod.ctypes
    dp = lambda a: a.ctypes.data_as(c.POINTER(c.c_double))
    mu = files.GAUSS2 * 1.001
    oe = np.array([1.7, 0.3, 0.2, 1.1, 2.3, 0.7])
    r = np.zeros(3); v = np.zeros(3); back = np.zeros(6)
    assert L.orc_calculate_phase(mu, dp(oe), dp(r), dp(v)) == 0
    assert L.orc_calculate_orbital_element(mu, dp(r), dp(v), dp(back)) == 0
    np.testing.assert_allclose(back, oe, rtol=1e-10, atol=1e-12)


# ---- two-body convergence ---------------------------------------------------------------
def kepler_relative(mu, r0, v0, t):
    """Analytic relative orbit (universal elliptic case) after time t."""
    r0 = np.asarray(r0, float); v0 = np.asarray(v0, float)
    R = np.linalg.norm(r0); V2 = v0 @ v0
    a = 1.0 / (2.0 / R - V2 / mu)
    n = math.sqrt(mu / a ** 3)
    h = np.cross(r0, v0)
    evec = np.cross(v0, h) / mu - r0 / R
    e = np.linalg.norm(evec)
    cosE0 = (1 - R / a) / e
    sinE0 = (r0 @ v0) / (e * math.sqrt(mu * a))
    E0 = math.atan2(sinE0, cosE0)
    M = E0 - e * math.sin(E0) + n * t
    E = M
    for _ in range(100):
        E = E - (E - e * math.sin(E) - M) / (1 - e * math.cos(E))
    f = a / R * (math.cos(E - E0) - 1) + 1
    g = t + (math.sin(E - E0) - (E - E0)) / n
    r = f * r0 + g * v0
    Rn = np.linalg.norm(r)
    fd = -math.sqrt(mu * a) / (Rn * R) * math.sin(E - E0)
    gd = 1 + a / Rn * (math.cos(E - E0) - 1)
    return r, fd * r0 + gd * v0


def two_body(m2=1e-3, e=0.2):
    # relative orbit a = 1 AU in internal units (mu = m1 + m2, G = 1)
    mu = 1.0 + m2
    a = 1.0
    rp = a * (1 - e)
    vp = math.sqrt(mu * (1 + e) / rp)
    rrel = np.array([rp, 0.0, 0.0]); vrel = np.array([0.0, vp, 0.0])
    r = np.array([-m2 / mu * rrel, 1.0 / mu * rrel]); v = np.array([-m2 / mu * vrel, 1.0 / mu * vrel])
    return _bodies(r, v, [1.0, m2], types=[0, 1]), mu, rrel, vrel


def pos_error(oracle_mod, integ, dt_day, n_steps, adaptive=False):
    b, mu, rrel, vrel = two_body()
    o = oracle_mod.Oracle(b, integrator=integ, dt_day=dt_day, adaptive=adaptive, tolerance=1.0)
    for _ in range(n_steps):
        o.step()
    s = o.state()
    rr, _ = kepler_relative(mu, rrel, vrel, o.t)
    return np.linalg.norm((s["r"][1] - s["r"][0]) - rr)


@pytest.mark.parametrize("integ,order,dt", [(0, 1, 1.0), (1, 2, 4.0), (2, 4, 8.0), (3, 5, 8.0), (4, 7, 6.0)])
def test_integrator_order(oracle_mod, integ, order, dt):
    T = 8 * 24.0   # days
    e1 = pos_error(oracle_mod, integ, dt, int(T / dt))
    e2 = pos_error(oracle_mod, integ, dt / 2, int(2 * T / dt))
    p = math.log2(e1 / e2)
    assert order - 0.35 < p < order + 0.9, (integ, e1, e2, p)


@pytest.mark.parametrize("integ,expo,dt", [(2, 4, 4.0), (3, 5, 4.0), (4, 8, 16.0)])
def test_embedded_error_order(oracle_mod, integ, expo, dt):
    """One adaptive attempt: the reference's error estimate scales as dt^expo.
    RK4 (int_rungekutta4.cu:231-253): dt * |f(t+dt, y + dt k3) - f(t+dt, y1)|; the
    O(dt) parts of the two stage states cancel, leaving dt * O(dt^3).  Cash-Karp:
    dt * sum (bh - b) f = dt * O(dt^4).  RKF7(8): dt * (f0 + f10 - f11 - f12) = dt * O(dt^7)."""
    errs = []
    for h in (dt, dt / 2):
        b, *_ = two_body()
        o = oracle_mod.Oracle(b, integrator=integ, dt_day=h, adaptive=True, tolerance=1.0)
        o.step()
        errs.append(o.last_err())
    p = math.log2(errs[0] / errs[1])
    assert expo - 0.5 < p < expo + 1.0, (errs, p)


# ---- gravity ------------------------------------------------------------------------------
def brute_force(b, r):
    """pp_disk.cu:404-572 written out with numpy, sinks vectorised, sources in order."""
    n = b.n
    ns = sum(b.counts[:4]); nm = ns + sum(b.counts[4:6])
    a = np.zeros((n, 3))
    sweeps = [(0, ns, 0, nm), (ns, nm, 0, ns), (nm, n, 0, nm)]
    for s0, s1, j0, j1 in sweeps:
        if s1 <= s0:
            continue
        idx = np.arange(s0, s1)
        acc = np.zeros((s1 - s0, 3))
        for j in range(j0, j1):
            if b.ids[j] < 0:
                continue
            dx = r[j, 0] - r[idx, 0]; dy = r[j, 1] - r[idx, 1]; dz = r[j, 2] - r[idx, 2]
            r2 = (dx * dx + dy * dy) + dz * dz
            d = np.sqrt(r2)
            with np.errstate(divide="ignore", invalid="ignore"):
                s = b.mass[j] / (d * r2)
            use = idx != j
            acc[use, 0] += s[use] * dx[use]
            acc[use, 1] += s[use] * dy[use]
            acc[use, 2] += s[use] * dz[use]
        live = b.ids[idx] > 0
        acc[~live] = 0.0
        a[idx] = acc
    return a


@pytest.mark.parametrize("maker", ["small", "mixed", "allsi", "powerlaw"])
def test_oracle_force_brute_force(oracle_mod, maker):
    b = {"small": lambda: generate.disk_small(60, seed=1),
         "mixed": lambda: generate.disk_mixed(n_pl=40, n_tp=30, seed=2),
         "allsi": lambda: generate.disk_allsi(50, seed=3),
         "powerlaw": lambda: generate.disk_powerlaw(300, seed=4, m_split=3e-10)}[maker]()
    b = generate.to_internal(b)
    b.ids[[3, 7]] *= -1
    o = oracle_mod.Oracle(b)
    assert_bitwise(o.accel(0.0, b.r, b.v), brute_force(b, b.r), "accel")


def test_momentum_balance_all_si(oracle_mod):
    b = generate.to_internal(generate.disk_allsi(80, seed=5))
    a = oracle_mod.Oracle(b).accel(0.0, b.r, b.v)
    F = (b.mass[:, None] * a).sum(axis=0)
    scale = np.abs(b.mass[:, None] * a).sum(axis=0)
    assert np.all(np.abs(F) <= 1e-13 * scale)


def test_energy_conserved_rkf78(oracle_mod):
    """Energy and angular momentum over three eccentric orbits; the drift shrinks with
    the tolerance (an error in a stage sum would leave a tolerance-independent drift)."""
    drift = []
    for tol in (1e-10, 1e-13):
        b, *_ = two_body(m2=1e-3, e=0.5)
        o = oracle_mod.Oracle(b, integrator=4, adaptive=True, tolerance=tol, dt_day=1.0)
        E0, L0 = o.energy()
        while o.t < 2 * math.pi * 3:
            o.step()
        E1, L1 = o.energy()
        drift.append(abs((E1 - E0) / E0))
        assert np.linalg.norm(L1 - L0) / np.linalg.norm(L0) < (1e-6 if tol > 1e-12 else 1e-8)
    assert drift[1] < 5e-9 and drift[1] < drift[0] / 10, drift


@pytest.mark.parametrize("frac,expect", [(0.99, True), (1.01, False)])
def test_mutual_hill_criterion_by_hand(oracle_mod, frac, expect):
    """Star + two protoplanets placed frac x f R_H apart, with R_H = ((m_1 + m_2) / (3 M))^(1/3)
    (|r_1| + |r_2|) / 2: one close encounter below the reach, none above; the star, always
    within a few such radii of each body, is never paired."""
    M, m1, m2, f = 1.0, 3e-7, 2e-7, 4.0
    rh0 = ((m1 + m2) / (3.0 * M)) ** (1.0 / 3.0)
    # body 1 at x = 1; body 2 on the x axis at 1 + s, so |r_1| + |r_2| = 2 + s and
    # s = frac f rh0 (2 + s) / 2
    k = frac * f * rh0 / 2.0
    s = 2.0 * k / (1.0 - k)
    r = [[0, 0, 0], [1.0, 0, 0], [1.0 + s, 0, 0]]
    v = [[0, 0, 0], [0, 1.0, 0], [0, 1.0, 0]]
    b = _bodies(r, v, [M, m1, m2], types=[0, 3, 3], radius=[1e-6, 1e-9, 1e-9])
    o = oracle_mod.Oracle(b, encounter_hill_factor=f, radii_enhance_factor=1.0, close_encounter=True)
    o.accel(0.0, b.r, b.v)
    ev = o.raw_events()
    pairs = [(e["idx1"], e["idx2"]) for e in ev]
    assert all(0 not in p for p in pairs)
    assert pairs == ([(2, 1)] if expect else [])
    if expect:
        assert ev[0]["d"] == (1.0 + s) - 1.0
