"""Reference file formats: body list, parameter file, result and event records.

Python mirror of the readers and writers the reference uses at its boundary:

* body file: 7-count header ``ns ngp nrp npp nspl npl ntp`` followed by 16
  whitespace-separated fields per body
  ``id name type epoch mass radius density cd mig_type mig_stop_at x y z vx vy vz``
  with v in AU/day (red.cuda/src/red.cuda/pp_disk.cu:1307-1322, 1576-1612);
* parameter file: ``key = value`` lines, ``#`` comments
  (red.cuda/src/red.cuda.utility/file_util.cpp:85-108,
  red.cuda/src/red.cuda/parameter.cpp:32-292);
* ASCII result snapshot, one line per active body
  (red.cuda/src/red.cuda/pp_disk.cu:1685-1725);
* event record line (red.cuda/src/red.cuda/pp_disk.cu:1772-1848).

The C++ driver of this package implements the same formats natively; these
functions exist for tests and for generating inputs.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

GAUSS = 1.720209895e-2
GAUSS2 = 2.959122082855911025e-4
YEAR_TO_DAY = 365.25

BODY_TYPES = ("STAR", "GIANTPLANET", "ROCKYPLANET", "PROTOPLANET",
              "SUPERPLANETESIMAL", "PLANETESIMAL", "TESTPARTICLE", "PADDINGPARTICLE")
EVENT_NAMES = ("NONE", "HIT_CENTRUM", "EJECTION", "CLOSE_ENCOUNTER", "COLLISION")

INTEGRATORS = {
    "e": 0, "euler": 0,
    "rk2": 1, "rungekutta2": 1,
    "rk4": 2, "rungekutta4": 2,
    "rk5": 3, "rungekutta5": 3,
    "rkf8": 4, "rungekuttafehlberg8": 4,
    "rkn": 5, "rungekuttanystrom": 5,
}
INTEGRATOR_SHORT = {0: "E", 1: "RK2", 2: "RK4", 3: "RK5", 4: "RKF8"}


@dataclasses.dataclass
class BodySet:
    counts: List[int]            # ns ngp nrp npp nspl npl ntp
    ids: np.ndarray              # int32
    names: List[str]
    types: np.ndarray            # int32
    epoch: np.ndarray            # float64
    mass: np.ndarray
    radius: np.ndarray
    density: np.ndarray
    cd: np.ndarray
    mig_type: np.ndarray         # int32
    mig_stop_at: np.ndarray
    r: np.ndarray                # (n, 3)
    v: np.ndarray                # (n, 3)

    @property
    def n(self) -> int:
        return int(sum(self.counts))

    def params4(self) -> np.ndarray:
        return np.stack([self.mass, self.radius, self.density, self.cd], axis=1).copy()

    def copy(self) -> "BodySet":
        return BodySet(list(self.counts), self.ids.copy(), list(self.names), self.types.copy(),
                       self.epoch.copy(), self.mass.copy(), self.radius.copy(), self.density.copy(),
                       self.cd.copy(), self.mig_type.copy(), self.mig_stop_at.copy(),
                       self.r.copy(), self.v.copy())


def read_body_file(path: str) -> BodySet:
    """Reads a body list; the name is cut to 30 characters as pp_disk.cu:1585-1589 does."""
    with open(path) as f:
        tok = f.read().split()
    if len(tok) < 7:
        raise ValueError("Cannot open " + path + ".")
    counts = [int(x) for x in tok[:7]]
    n = sum(counts)
    body = tok[7:7 + 16 * n]
    if len(body) < 16 * n:
        raise ValueError(f"{path}: expected {n} body records")
    rec = [body[16 * k:16 * (k + 1)] for k in range(n)]
    f64 = lambda col: np.array([float(x[col]) for x in rec], dtype=np.float64)
    i32 = lambda col: np.array([int(x[col]) for x in rec], dtype=np.int32)
    return BodySet(
        counts=counts,
        ids=i32(0), names=[x[1][:30] for x in rec], types=i32(2), epoch=f64(3),
        mass=f64(4), radius=f64(5), density=f64(6), cd=f64(7),
        mig_type=i32(8), mig_stop_at=f64(9),
        r=np.stack([f64(10), f64(11), f64(12)], axis=1),
        v=np.stack([f64(13), f64(14), f64(15)], axis=1),
    )


def _g(x: float) -> str:
    return repr(float(x))


def write_body_file(path: str, b: BodySet) -> None:
    """Writes a body list that read_body_file reads back bit-exactly (17 significant digits)."""
    lines = [" ".join(str(c) for c in b.counts)]
    for k in range(b.n):
        lines.append(" ".join([
            str(int(b.ids[k])), b.names[k], str(int(b.types[k])), _g(b.epoch[k]),
            _g(b.mass[k]), _g(b.radius[k]), _g(b.density[k]), _g(b.cd[k]),
            str(int(b.mig_type[k])), _g(b.mig_stop_at[k]),
            _g(b.r[k, 0]), _g(b.r[k, 1]), _g(b.r[k, 2]),
            _g(b.v[k, 0]), _g(b.v[k, 1]), _g(b.v[k, 2])]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _is_number(s: str) -> bool:  # tools.cpp:20-34
    return all(c.isdigit() or c in "eE.-+" for c in s)


def _atof(s: str) -> float:
    """C atof: the longest numeric prefix, 0.0 if none."""
    best = 0.0
    for k in range(len(s), 0, -1):
        try:
            best = float(s[:k])
            break
        except ValueError:
            continue
    return best


def default_parameters() -> Dict:
    """parameter.cpp:32-53"""
    return dict(name="", description="", adaptive=False, error_check_for_tp=False,
                close_encounter=False, inner_steps=True, integrator=0, tolerance=1.0e-10,
                start_time=0.0, length=0.0, output_interval=0.0, output_type="text",
                frame_center=1, hit_centrum=0.0, ejection=0.0, radii_enhance_factor=0.0,
                dt=0.001, encounter_hill_factor=0.0)


def parse_parameters(text: str) -> Dict:
    """parameter.cpp:55-292.  Times come back in internal units (days * k).

    Two keys are additions of this package and default to the reference's
    behaviour when absent: ``dt`` (initial step in days, the reference
    hard-codes 0.001 in main.cu:141) and ``encounter_hill_factor`` (close
    encounters at that many mutual Hill radii; 0 = off).
    """
    p = default_parameters()
    lines = []
    for raw in text.splitlines():   # file_util.cpp:85-108
        if len(raw) == 0 or raw[0] == "#":
            continue
        lines.append(raw.split("#", 1)[0])
    for line in lines:
        if line.strip("\r\n") == "":
            continue
        parts = [t for t in line.split("=") if t != ""]
        if len(parts) < 2:
            raise ValueError("Invalid key/value pair: " + line + ".")
        key, value = parts[0].strip(" \t").lower(), parts[1].strip(" \t")
        num_keys = ("tolerance", "start_time", "length", "output_interval", "ejection",
                    "hit_centrum", "radii_enhance_factor", "collision_factor", "dt",
                    "encounter_hill_factor")
        if key in num_keys and not _is_number(value):
            raise ValueError("Invalid number at: " + key)
        if key == "name":
            p["name"] = value
        elif key == "description":
            p["description"] = value
        elif key == "integrator":
            v = value.lower()
            if v not in INTEGRATORS:
                raise ValueError("Invalid integrator type: " + v)
            p["integrator"] = INTEGRATORS[v]
        elif key == "tolerance":
            p["adaptive"] = True
            p["tolerance"] = _atof(value)
        elif key in ("error_check_for_tp", "close_encounter", "inner_steps"):
            if value not in ("true", "false"):
                raise ValueError("Invalid value at: " + key)
            p[key] = value == "true"
        elif key in ("start_time", "length", "output_interval"):
            p[key] = _atof(value) * YEAR_TO_DAY
        elif key == "ejection":
            p["ejection"] = _atof(value)
        elif key == "hit_centrum":
            p["hit_centrum"] = _atof(value)
        elif key in ("radii_enhance_factor", "collision_factor"):
            p["radii_enhance_factor"] = _atof(value)
        elif key == "output_type":
            if value in ("text", "ascii"):
                p["output_type"] = "text"
            elif value == "binary":
                p["output_type"] = "binary"
            else:
                raise ValueError("Invalid value at: " + key)
        elif key == "frame_center":
            if value == "astro":
                p["frame_center"] = 0
            elif value == "bary":
                p["frame_center"] = 1
            else:
                raise ValueError("Invalid value at: " + key)
        elif key == "dt":
            p["dt"] = _atof(value)
        elif key == "encounter_hill_factor":
            p["encounter_hill_factor"] = _atof(value)
        else:
            raise ValueError("Invalid parameter: " + key + ".")
    for k in ("start_time", "length", "output_interval"):   # parameter.cpp:287-292
        p[k] *= GAUSS
    p["stop_time"] = p["start_time"] + p["length"]
    return p


def read_parameters(path: str) -> Dict:
    with open(path) as f:
        return parse_parameters(f.read())


def parse_result_ascii(path: str) -> List[List[str]]:
    with open(path) as f:
        return [ln.split() for ln in f if ln.strip()]


def read_result_snapshot(path: str) -> Tuple[float, BodySet]:
    """Last snapshot of a text result file (the trailing lines sharing the time column) as a
    body list plus its time in days: the reference's implicit restart, since its snapshot has
    the body file's 16 columns with t in the epoch column (pp_disk.cu:1707-1722 against
    1576-1612).  Counts follow the body types, which must be in type order."""
    rows = [r for r in parse_result_ascii(path)]
    if not rows:
        raise ValueError(f"No snapshot in {path}.")
    for r in rows:
        if len(r) != 16:
            raise ValueError(f"Invalid result line in {path}: {' '.join(r)}")
    t = float(rows[-1][3])
    k = len(rows)
    while k > 0 and float(rows[k - 1][3]) == t:
        k -= 1
    rec = rows[k:]
    types = [int(x[2]) for x in rec]
    if any(b < a for a, b in zip(types, types[1:])) or min(types) < 0 or max(types) > 6:
        raise ValueError(f"Snapshot bodies out of type order in {path}.")
    counts = [types.count(c) for c in range(7)]
    f64 = lambda col: np.array([float(x[col]) for x in rec], dtype=np.float64)
    i32 = lambda col: np.array([int(x[col]) for x in rec], dtype=np.int32)
    return t, BodySet(
        counts=counts,
        ids=i32(0), names=[x[1][:30] for x in rec], types=i32(2), epoch=f64(3),
        mass=f64(4), radius=f64(5), density=f64(6), cd=f64(7),
        mig_type=i32(8), mig_stop_at=f64(9),
        r=np.stack([f64(10), f64(11), f64(12)], axis=1),
        v=np.stack([f64(13), f64(14), f64(15)], axis=1),
    )


def _c_e16(x: float) -> str:
    """x as C++ iostream prints it with scientific, precision 16 and width 25 (a NaN keeps its
    sign, as printf's "%e" does)."""
    if x != x:
        return ("-nan" if math.copysign(1.0, x) < 0 else "nan").rjust(25)
    return f"{x: .16e}".rjust(25)


def format_result_line(id_, name, btype, t_day, p4, mig_type, mig_stop, r, v_day) -> str:
    """One line of pp_disk.cu:1707-1722 (C++ iostream scientific, precision 16)."""
    e = _c_e16
    fields = [str(int(id_)).rjust(8), name.rjust(30), str(int(btype)).rjust(2), e(t_day),
              e(p4[0]), e(p4[1]), e(p4[2]), e(p4[3]), str(int(mig_type)).rjust(2), e(mig_stop),
              e(r[0]), e(r[1]), e(r[2]), e(v_day[0]), e(v_day[1]), e(v_day[2])]
    return " ".join(fields)


def format_event_line(e: Dict) -> str:
    """One line of the event file (pp_disk.cu:1786-1837): name, t in days, d, id1, id2, then
    for the survivor before, the merger before and the survivor after: mass, density,
    radius, x y z, vx vy vz in AU/day and the six orbital elements (C++ iostream
    scientific, precision 16)."""
    e16 = _c_e16
    f = [EVENT_NAMES[e["name"]].rjust(16), e16(e["t"] / GAUSS), e16(e["d"]), str(int(e["id1"])).rjust(8),
         str(int(e["id2"])).rjust(8)]
    for p, r, v, oe in (("p1", "r1", "v1", "oe1"), ("p2", "r2", "v2", "oe2"), ("ps", "rs", "vs", "oes")):
        f += [e16(e[p][0]), e16(e[p][2]), e16(e[p][1])]
        f += [e16(x) for x in e[r][:3]]
        f += [e16(x * GAUSS) for x in e[v][:3]]
        f += [e16(x) for x in e[oe][:6]]
    return " ".join(f)


def mutual_hill_radius(m_i: float, m_j: float, m_star: float, ri: float, rj: float) -> float:
    """R_H = ((m_i + m_j) / (3 M_star))^(1/3) * (|r_i| + |r_j|) / 2 (this package's definition)."""
    return math.pow((m_i + m_j) / (3.0 * m_star), 1.0 / 3.0) * (ri + rj) * 0.5
