"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by the benchmark's CPU-baseline leg.  The product never imports it.
The oracle restates the reference's FP64 CPU path; see red_oracle.cpp for the
reference file:line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

EV_NAMES = ("NONE", "HIT_CENTRUM", "EJECTION", "CLOSE_ENCOUNTER", "COLLISION")


class OrcEvent(ctypes.Structure):
    _fields_ = [("name", ctypes.c_int), ("id1", ctypes.c_int), ("idx1", ctypes.c_int),
                ("id2", ctypes.c_int), ("idx2", ctypes.c_int), ("attempt", ctypes.c_int),
                ("stage", ctypes.c_int), ("sink", ctypes.c_int), ("source", ctypes.c_int),
                ("pad", ctypes.c_int), ("t", ctypes.c_double), ("d", ctypes.c_double),
                ("p1", ctypes.c_double * 4), ("p2", ctypes.c_double * 4), ("ps", ctypes.c_double * 4),
                ("r1", ctypes.c_double * 3), ("v1", ctypes.c_double * 3),
                ("r2", ctypes.c_double * 3), ("v2", ctypes.c_double * 3),
                ("rs", ctypes.c_double * 3), ("vs", ctypes.c_double * 3),
                ("oe1", ctypes.c_double * 6), ("oe2", ctypes.c_double * 6), ("oes", ctypes.c_double * 6)]


class OrcParams(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int), ("adaptive", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("error_check_for_tp", ctypes.c_int), ("close_encounter", ctypes.c_int),
                ("inner_steps", ctypes.c_int), ("frame_center", ctypes.c_int),
                ("stage_time_events", ctypes.c_int), ("hc_reactivate", ctypes.c_int),
                ("n_threads", ctypes.c_int),
                ("hit_centrum", ctypes.c_double), ("ejection", ctypes.c_double),
                ("radii_enhance_factor", ctypes.c_double), ("encounter_hill_factor", ctypes.c_double),
                ("dt_day", ctypes.c_double), ("padded", ctypes.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = P
        L.orc_create.argtypes = [ip, ip, ip, ip, dp, dp, dp, dp, dp, ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        for name in ("orc_n_total", "orc_check_ejection_hit_centrum", "orc_check_collision",
                     "orc_check_close_encounter", "orc_n_raw_events"):
            getattr(L, name).restype = I
            getattr(L, name).argtypes = [P]
        L.orc_check_rebuild.restype = I
        L.orc_check_rebuild.argtypes = [P, I]
        L.orc_get_counts.argtypes = [P, ip]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [P]
        for name in ("orc_transform_to_bc", "orc_transform_to_ac", "orc_transform_time_velocity",
                     "orc_clear_events"):
            getattr(L, name).argtypes = [P]
        L.orc_set_t.argtypes = [P, D]
        L.orc_set_dt_try.argtypes = [P, D]
        for name in ("orc_get_t", "orc_dt_try", "orc_step", "orc_last_err"):
            getattr(L, name).restype = D
            getattr(L, name).argtypes = [P]
        L.orc_get_state.argtypes = [P, dp, dp, dp, ip, ip]
        L.orc_set_state.argtypes = [P, dp, dp]
        L.orc_get_raw_events.restype = I
        L.orc_get_raw_events.argtypes = [P, ctypes.POINTER(OrcEvent), I]
        L.orc_get_sp_events.restype = I
        L.orc_get_sp_events.argtypes = [P, ctypes.POINTER(OrcEvent), I]
        L.orc_accel.argtypes = [P, D, I, dp, dp, dp]
        L.orc_counters.argtypes = [P, ip]
        L.orc_energy.argtypes = [P, dp, dp]
        L.orc_calculate_radius.restype = D
        L.orc_calculate_radius.argtypes = [D, D]
        L.orc_calc_physical_properties.argtypes = [dp, dp, dp]
        L.orc_calc_position_after_collision.argtypes = [D, D, dp, dp, dp]
        L.orc_kepler_equation_solver.restype = I
        L.orc_kepler_equation_solver.argtypes = [D, D, D, dp]
        L.orc_calculate_phase.restype = I
        L.orc_calculate_phase.argtypes = [D, dp, dp, dp]
        L.orc_calculate_orbital_element.restype = I
        L.orc_calculate_orbital_element.argtypes = [D, dp, dp, dp]
        L.orc_create_sp_events.restype = I
        L.orc_create_sp_events.argtypes = [ctypes.POINTER(OrcEvent), I, ctypes.POINTER(OrcEvent)]
        L.orc_max_error.restype = D
        L.orc_max_error.argtypes = [dp, dp, I, D, D]
        L.orc_step_factor.restype = D
        L.orc_step_factor.argtypes = [D, D]
        L.orc_handle_collision_pair.restype = I
        L.orc_handle_collision_pair.argtypes = [P, ctypes.POINTER(OrcEvent), I]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def event_dict(e: OrcEvent) -> dict:
    d = {f: getattr(e, f) for f, _ in OrcEvent._fields_ if f != "pad"}
    for k, v in list(d.items()):
        if not isinstance(v, (int, float)):
            d[k] = tuple(v)
    return d


class Oracle:
    """One simulated disk integrated by the CPU oracle (internal units: t*k, v/k)."""

    def __init__(self, bodies, integrator=0, adaptive=False, tolerance=1e-10, dt_day=0.001,
                 error_check_for_tp=False, close_encounter=False, inner_steps=True, frame_center=1,
                 hit_centrum=0.0, ejection=0.0, radii_enhance_factor=0.0, encounter_hill_factor=0.0,
                 stage_time_events=True, hc_reactivate=False, n_threads=1, padded=False, r=None, v=None):
        L = lib()
        self.L = L
        prm = OrcParams(integrator, int(adaptive), tolerance, int(error_check_for_tp), int(close_encounter),
                        int(inner_steps), frame_center, int(stage_time_events), int(hc_reactivate), n_threads,
                        hit_centrum, ejection, radii_enhance_factor, encounter_hill_factor, dt_day, int(padded))
        counts = np.ascontiguousarray(bodies.counts, dtype=np.int32)
        ids = np.ascontiguousarray(bodies.ids, dtype=np.int32)
        types = np.ascontiguousarray(bodies.types, dtype=np.int32)
        mig = np.ascontiguousarray(bodies.mig_type, dtype=np.int32)
        migs = np.ascontiguousarray(bodies.mig_stop_at, dtype=np.float64)
        ep = np.ascontiguousarray(bodies.epoch, dtype=np.float64)
        p4 = np.ascontiguousarray(bodies.params4(), dtype=np.float64)
        rr = np.ascontiguousarray(bodies.r if r is None else r, dtype=np.float64)
        vv = np.ascontiguousarray(bodies.v if v is None else v, dtype=np.float64)
        self.h = L.orc_create(_ip(counts), _ip(ids), _ip(types), _ip(mig), _dp(migs), _dp(ep), _dp(p4),
                              _dp(rr), _dp(vv), ctypes.byref(prm))
        if not self.h:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    @property
    def n(self):
        return self.L.orc_n_total(self.h)

    def counts(self):
        c = np.zeros(7, dtype=np.int32)
        self.L.orc_get_counts(self.h, _ip(c))
        return [int(x) for x in c]

    def transform_to_bc(self):
        self.L.orc_transform_to_bc(self.h)

    def transform_to_ac(self):
        self.L.orc_transform_to_ac(self.h)

    def transform_time_velocity(self):
        self.L.orc_transform_time_velocity(self.h)

    @property
    def t(self):
        return self.L.orc_get_t(self.h)

    @t.setter
    def t(self, value):
        self.L.orc_set_t(self.h, value)

    @property
    def dt_try(self):
        return self.L.orc_dt_try(self.h)

    @dt_try.setter
    def dt_try(self, value):
        self.L.orc_set_dt_try(self.h, value)

    def last_err(self):
        return self.L.orc_last_err(self.h)

    def step(self) -> float:
        dt = self.L.orc_step(self.h)
        if dt != dt:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())
        return dt

    def state(self):
        n = self.n
        r = np.zeros((n, 3)); v = np.zeros((n, 3)); p = np.zeros((n, 4))
        ids = np.zeros(n, dtype=np.int32); types = np.zeros(n, dtype=np.int32)
        self.L.orc_get_state(self.h, _dp(r), _dp(v), _dp(p), _ip(ids), _ip(types))
        return dict(r=r, v=v, p=p, id=ids, type=types)

    def set_state(self, r, v):
        r = np.ascontiguousarray(r, dtype=np.float64); v = np.ascontiguousarray(v, dtype=np.float64)
        self.L.orc_set_state(self.h, _dp(r), _dp(v))

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())
        return bool(rc)

    def check_for_ejection_hit_centrum(self) -> bool:
        return self._check(self.L.orc_check_ejection_hit_centrum(self.h))

    def check_for_collision(self) -> bool:
        return self._check(self.L.orc_check_collision(self.h))

    def check_for_close_encounter(self) -> bool:
        return self._check(self.L.orc_check_close_encounter(self.h))

    def check_for_rebuild_vectors(self, n: int) -> bool:
        return self._check(self.L.orc_check_rebuild(self.h, n))

    def clear_event_counter(self):
        self.L.orc_clear_events(self.h)

    def n_event(self):
        return self.L.orc_n_raw_events(self.h)

    def raw_events(self):
        n = self.L.orc_n_raw_events(self.h)
        buf = (OrcEvent * max(n, 1))()
        self.L.orc_get_raw_events(self.h, buf, n)
        return [event_dict(buf[k]) for k in range(n)]

    def sp_events(self):
        buf = (OrcEvent * 1)()
        n = self.L.orc_get_sp_events(self.h, buf, 0)
        buf = (OrcEvent * max(n, 1))()
        self.L.orc_get_sp_events(self.h, buf, n)
        return [event_dict(buf[k]) for k in range(n)]

    def accel(self, t, r, v, stage=0):
        r = np.ascontiguousarray(r, dtype=np.float64); v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.zeros_like(r)
        self.L.orc_accel(self.h, t, stage, _dp(r), _dp(v), _dp(a))
        return a

    def handle_collision_pair(self, event: OrcEvent, kernel_negated: bool = False) -> OrcEvent:
        """pp_disk.cu:1151-1202 on a caller-built record; returns the completed record."""
        self._check(self.L.orc_handle_collision_pair(self.h, ctypes.byref(event), int(kernel_negated)))
        return event

    def counters(self):
        c = np.zeros(8, dtype=np.int32)
        self.L.orc_counters(self.h, _ip(c))
        keys = ("tried", "failed", "passed", "n_collision", "n_ejection", "n_hit_centrum",
                "n_close_encounter", "n_event_last_clear")
        return dict(zip(keys, (int(x) for x in c)))

    def energy(self):
        E = np.zeros(1); L = np.zeros(3)
        self.L.orc_energy(self.h, _dp(E), _dp(L))
        return float(E[0]), L


def create_sp_events(records):
    """pp_disk.cu:1111-1149 on a list of OrcEvent records; returns the kept records."""
    L = lib()
    n = len(records)
    src = (OrcEvent * max(n, 1))(*records)
    out = (OrcEvent * max(n, 1))()
    k = L.orc_create_sp_events(src, n, out)
    return [out[i] for i in range(k)]


def max_error(err_r, err_v, n_var, dt, lam):
    """integrator.cu:211-256 (CPU branch) on given error vectors."""
    er = np.ascontiguousarray(err_r, dtype=np.float64)
    ev = np.ascontiguousarray(err_v, dtype=np.float64)
    return lib().orc_max_error(_dp(er), _dp(ev), int(n_var), float(dt), float(lam))


def step_factor(x, y):
    """The controllers' factor x^y as the oracle computes it (powq, rounded to a double)."""
    return lib().orc_step_factor(float(x), float(y))
