"""Seeded synthetic disks in the reference body-file layout.

The reference's own generator (red.cuda/src/red.cuda.initial/main.cu:29-85,
551-600, distribution.cpp:11-250) seeds with time(NULL); these disks are
reproducible instead: every draw comes from numpy's MT19937 bit generator with
a fixed seed.  Orbital elements are turned into barycentre-free phase with
mu = k^2 (M_star + m) as red.cuda.initial/main.cu:577-578 does (the Kepler
solver and rotation follow tools.cpp:292-359); the loader then moves the state
to the barycentre.

Workloads (masses in M_sun, lengths in AU, velocities in AU/day):

* ``disk_small``  1 star (1 M_sun, R = 0.00465 AU) + giants (1e-3 at 5.2 AU,
  3e-4 at 9.5 AU, e = i = 0) + planetesimals with m log-uniform in
  [1e-12, 1e-10], density 2 g/cm^3, a uniform in [0.4, 4], e ~ Rayleigh(0.01),
  i ~ Rayleigh(0.005), angles uniform.  Planetesimals are type 5 (non-self-
  interacting) as in the reference's typing.
* ``disk_mixed``  1 star + 2 giants + n_pl planetesimals (m log-uniform in
  [1e-12, 1e-9]) + n_tp massless test particles, elements as disk_small.
* ``disk_powerlaw`` 1 star + n bodies with dN/dm ~ m^-2.5 on [1e-11, 3e-7]
  and surface density ~ a^-1.5 on [0.5, 5] AU; bodies with m >= 1e-10 are
  protoplanets (type 3, self-interacting), the rest planetesimals (type 5).
* ``disk_allsi`` n self-interacting protoplanets around a star (kernel tests).
* ``ensemble``    independent disks of N log-uniform in [64, 4096] bodies, each drawn as
  disk_small with its planetesimals out to a per-disk radius uniform in [3, 8] AU.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .files import GAUSS, BodySet

SOLAR_TO_GRAM = 1.0e3 * 1.98911e30
METER_TO_AU = 1.0 / 1.495978707e11
GRAM_PER_CM3_TO_SOLAR_PER_AU3 = (1.0 / SOLAR_TO_GRAM) / (1.0e-2 * METER_TO_AU) ** 3
FOUR_PI_OVER_THREE = 4.1887902047863909846168578443727


def radius_from_density(m: np.ndarray, rho: float) -> np.ndarray:
    """tools.cpp:238-243"""
    return np.power(1.0 / FOUR_PI_OVER_THREE * m / rho, 1.0 / 3.0)


def kepler_E(ecc: np.ndarray, mean: np.ndarray) -> np.ndarray:
    E = mean + ecc * np.sin(mean) / (1.0 - np.sin(mean + ecc) + np.sin(mean))
    for _ in range(30):
        E = E - (E - ecc * np.sin(E) - mean) / (1.0 - ecc * np.cos(E))
    return E


def phase_from_elements(mu, a, e, inc, peri, node, mean) -> Tuple[np.ndarray, np.ndarray]:
    """Vectorised tools.cpp:317-359."""
    E = np.where(e == 0.0, mean, kepler_E(e, mean))
    nu = 2.0 * np.arctan(np.sqrt((1.0 + e) / (1.0 - e)) * np.tan(E / 2.0))
    p = a * (1.0 - e * e)
    r = p / (1.0 + e * np.cos(nu))
    kszi, eta = r * np.cos(nu), r * np.sin(nu)
    vk = -np.sqrt(mu / p) * np.sin(nu)
    ve = np.sqrt(mu / p) * (e + np.cos(nu))
    cw, sw, cO, sO, ci, si = np.cos(peri), np.sin(peri), np.cos(node), np.sin(node), np.cos(inc), np.sin(inc)
    P = np.stack([cw * cO - sw * sO * ci, cw * sO + sw * cO * ci, sw * si], axis=1)
    Q = np.stack([-sw * cO - cw * sO * ci, -sw * sO + cw * cO * ci, cw * si], axis=1)
    rv = kszi[:, None] * P + eta[:, None] * Q
    vv = vk[:, None] * P + ve[:, None] * Q
    return rv, vv


class _Builder:
    def __init__(self):
        self.rows = []   # (type, name, mass, radius, density, r(3), v(3))

    def add(self, btype, name, mass, radius, density, r, v):
        self.rows.append((btype, name, float(mass), float(radius), float(density), r, v))

    def build(self) -> BodySet:
        rows = sorted(enumerate(self.rows), key=lambda kv: (kv[1][0], kv[0]))   # body-file class order
        rows = [r for _, r in rows]
        n = len(rows)
        counts = [sum(1 for r in rows if r[0] == t) for t in range(7)]
        return BodySet(
            counts=counts,
            ids=np.arange(1, n + 1, dtype=np.int32),
            names=[r[1] for r in rows],
            types=np.array([r[0] for r in rows], dtype=np.int32),
            epoch=np.zeros(n),
            mass=np.array([r[2] for r in rows]),
            radius=np.array([r[3] for r in rows]),
            density=np.array([r[4] for r in rows]),
            cd=np.zeros(n),
            mig_type=np.zeros(n, dtype=np.int32),
            mig_stop_at=np.zeros(n),
            r=np.array([r[5] for r in rows], dtype=np.float64).reshape(n, 3),
            v=np.array([r[6] for r in rows], dtype=np.float64).reshape(n, 3),
        )


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.MT19937(seed))


def _elements(rng, n, a_lo, a_hi, sig_e=0.01, sig_i=0.005, a_sampler=None):
    if a_sampler is None:
        a = rng.uniform(a_lo, a_hi, n)
    else:
        a = a_sampler(rng, n)
    e = np.minimum(sig_e * np.sqrt(-2.0 * np.log(1.0 - rng.uniform(0, 1, n))), 0.9)
    inc = sig_i * np.sqrt(-2.0 * np.log(1.0 - rng.uniform(0, 1, n)))
    peri = rng.uniform(0, 2 * math.pi, n)
    node = rng.uniform(0, 2 * math.pi, n)
    mean = rng.uniform(0, 2 * math.pi, n)
    return a, e, inc, peri, node, mean


def _add_star(B: _Builder, m_star=1.0, r_star=0.00465):
    rho = m_star / (FOUR_PI_OVER_THREE * r_star ** 3)
    B.add(0, "star", m_star, r_star, rho, np.zeros(3), np.zeros(3))


def _add_giants(B: _Builder, rng, m_star, giants=((1.0e-3, 5.2), (3.0e-4, 9.5))):
    rho = 1.33 * GRAM_PER_CM3_TO_SOLAR_PER_AU3
    for k, (m, a) in enumerate(giants):
        mu = GAUSS ** 2 * (m_star + m)
        ang = rng.uniform(0, 2 * math.pi, 3)
        r, v = phase_from_elements(mu, np.array([a]), np.array([0.0]), np.array([0.0]), ang[0:1], ang[1:2], ang[2:3])
        B.add(1, f"giant{k}", m, radius_from_density(np.array([m]), rho)[0], rho, r[0], v[0])


def _add_bodies(B, btype, prefix, m, rho, elems, m_star):
    a, e, inc, peri, node, mean = elems
    mu = GAUSS ** 2 * (m_star + m)
    r, v = phase_from_elements(mu, a, e, inc, peri, node, mean)
    R = radius_from_density(m, rho) if rho > 0 else np.zeros_like(m)
    dens = np.full_like(m, rho)
    for k in range(len(m)):
        B.add(btype, f"{prefix}{k}", m[k], R[k], dens[k], r[k], v[k])


def disk_small(n_pl: int = 253, seed: int = 101, a_range=(0.4, 4.0), giants=True) -> BodySet:
    rng = _rng(seed)
    B = _Builder()
    _add_star(B)
    if giants:
        _add_giants(B, rng, 1.0)
    m = np.exp(rng.uniform(math.log(1e-12), math.log(1e-10), n_pl))
    _add_bodies(B, 5, "pl", m, 2.0 * GRAM_PER_CM3_TO_SOLAR_PER_AU3, _elements(rng, n_pl, *a_range), 1.0)
    return B.build()


def disk_mixed(n_pl: int = 65536, n_tp: int = 65533, seed: int = 303, a_range=(0.4, 4.0)) -> BodySet:
    rng = _rng(seed)
    B = _Builder()
    _add_star(B)
    _add_giants(B, rng, 1.0)
    m = np.exp(rng.uniform(math.log(1e-12), math.log(1e-9), n_pl))
    _add_bodies(B, 5, "pl", m, 2.0 * GRAM_PER_CM3_TO_SOLAR_PER_AU3, _elements(rng, n_pl, *a_range), 1.0)
    _add_bodies(B, 6, "tp", np.zeros(n_tp), 0.0, _elements(rng, n_tp, *a_range), 1.0)
    return B.build()


def disk_powerlaw(n: int = 16383, seed: int = 202, m_split: float = 1.0e-10) -> BodySet:
    rng = _rng(seed)
    B = _Builder()
    _add_star(B)
    m0, m1 = 1.0e-11, 3.0e-7
    u = rng.uniform(0, 1, n)
    m = (m0 ** -1.5 + u * (m1 ** -1.5 - m0 ** -1.5)) ** (-1.0 / 1.5)

    def a_sampler(rng_, k):
        uu = rng_.uniform(0, 1, k)
        return (math.sqrt(0.5) + uu * (math.sqrt(5.0) - math.sqrt(0.5))) ** 2

    el = _elements(rng, n, 0.5, 5.0, a_sampler=a_sampler)
    rho = 2.0 * GRAM_PER_CM3_TO_SOLAR_PER_AU3
    big = m >= m_split
    _add_bodies(B, 3, "pp", m[big], rho, tuple(x[big] for x in el), 1.0)
    _add_bodies(B, 5, "pl", m[~big], rho, tuple(x[~big] for x in el), 1.0)
    return B.build()


def disk_allsi(n: int, seed: int = 7, a_range=(0.4, 4.0), m_range=(1e-12, 1e-9)) -> BodySet:
    rng = _rng(seed)
    B = _Builder()
    _add_star(B)
    m = np.exp(rng.uniform(math.log(m_range[0]), math.log(m_range[1]), n - 1))
    _add_bodies(B, 3, "pp", m, 2.0 * GRAM_PER_CM3_TO_SOLAR_PER_AU3, _elements(rng, n - 1, *a_range), 1.0)
    return B.build()


def ensemble(n_disks: int = 512, seed: int = 404, n_range=(64, 4096), r_out_range=(3.0, 8.0),
             giants: bool = True) -> List[BodySet]:
    """Independent disks: N log-uniform, planetesimal outer radius uniform; each drawn as the
    256-body disk (star, the two giants, planetesimals), N bodies in all."""
    rng = _rng(seed)
    out = []
    for d in range(n_disks):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        r_out = float(rng.uniform(*r_out_range))
        out.append(disk_small(n_pl=n - (3 if giants else 1), seed=seed * 1000 + d, a_range=(0.4, r_out),
                              giants=giants))
    return out


def to_internal(b: BodySet, transform_bc: Optional[bool] = None) -> BodySet:
    """What options::create_pp_disk does before integration (options.cpp:197-235):
    barycentric shift when the star sits at rest at the origin, epochs * k, v / k.
    Uses numpy in the reference's summation order (last body first)."""
    c = b.copy()
    if transform_bc is None:
        transform_bc = bool(np.all(c.r[0] == 0.0) and np.all(c.v[0] == 0.0))
    if transform_bc:
        R0 = np.zeros(3); V0 = np.zeros(3)
        for j in range(c.n - 1, -1, -1):
            R0 += c.mass[j] * c.r[j]
            V0 += c.mass[j] * c.v[j]
        M0 = 0.0
        for j in range(c.n - 1, -1, -1):
            M0 += c.mass[j]
        R0 /= M0; V0 /= M0
        c.r = c.r - R0
        c.v = c.v - V0
    c.epoch = c.epoch * GAUSS
    c.v = c.v / GAUSS
    return c


# ---- command line: body and parameter files for the benchmark workloads -------------------
# Each workload is written in the reference's input layout, so that the drop-in driver runs
# it as `red -iDir <dir> -p parameters.txt -ic disk.txt` (add -mixed for the mixed disk).
# Collisions merge at the physical radii (collision_factor 1), close encounters are reported
# at 3 mutual Hill radii, bodies beyond 100 AU are ejected.
_RULES = "collision_factor = 1\nencounter_hill_factor = 3\nejection = 100\n"
WORKLOADS = {
    # 256 bodies, fixed-step RK4, dt = 1 day, 10000 steps (the loop steps while t <= stop)
    "small": dict(make=lambda seed: disk_small(seed=seed), seed=101,
                  params="integrator = rk4\ndt = 1\nlength = {!r}\noutput_interval = 1\n".format(9999.5 / 365.25)),
    # 16384 bodies, RKF7(8) at tolerance 1e-12, 1000 yr
    "powerlaw": dict(make=lambda seed: disk_powerlaw(seed=seed), seed=202,
                     params="integrator = rkf8\ntolerance = 1e-12\nlength = 1000\noutput_interval = 10\n"),
    # 131072 bodies (65536 planetesimals, 65533 test particles), RK5(4) at 1e-10, 1000 yr
    "mixed": dict(make=lambda seed: disk_mixed(seed=seed), seed=303,
                  params="integrator = rk5\ntolerance = 1e-10\nlength = 1000\noutput_interval = 10\n"),
}


def write_workload(name: str, out_dir: str, seed: Optional[int] = None, n_disks: int = 512) -> List[str]:
    """Writes the body file(s) and parameters.txt of a workload; returns the paths written.
    'ensemble' writes disk_000.txt ... (512 disks, RKF7(8) at 1e-11, 100 yr each)."""
    import os
    from .files import write_body_file
    os.makedirs(out_dir, exist_ok=True)
    written = []
    if name == "ensemble":
        disks = ensemble(n_disks=n_disks, seed=404 if seed is None else seed)
        for k, b in enumerate(disks):
            p = os.path.join(out_dir, f"disk_{k:03d}.txt")
            write_body_file(p, b)
            written.append(p)
        params = "integrator = rkf8\ntolerance = 1e-11\nlength = 100\noutput_interval = 10\n"
    else:
        w = WORKLOADS[name]
        p = os.path.join(out_dir, "disk.txt")
        write_body_file(p, w["make"](w["seed"] if seed is None else seed))
        written.append(p)
        params = w["params"]
    p = os.path.join(out_dir, "parameters.txt")
    with open(p, "w") as f:
        f.write(f"name = {name}\n" + params + _RULES)
    written.append(p)
    return written


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="Writes a benchmark disk in the reference's body-file layout.")
    ap.add_argument("workload", choices=sorted(list(WORKLOADS) + ["ensemble"]))
    ap.add_argument("--out", required=True, help="output directory")
    ap.add_argument("--seed", type=int, default=None, help="seed (default: the workload's own)")
    ap.add_argument("--disks", type=int, default=512, help="number of disks of the ensemble")
    a = ap.parse_args(argv)
    for p in write_workload(a.workload, a.out, a.seed, a.disks):
        print(p)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
