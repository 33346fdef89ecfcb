"""ctypes binding of librednative.so (the C ABI in include/rednative.h).

``Disk`` mirrors the surface the reference's driver uses on pp_disk and the
integrator (red.cuda/src/red.cuda/pp_disk.h:50-108, integrator.h:42): ``step``,
``check_for_ejection_hit_centrum``, ``check_for_collision``,
``check_for_close_encounter``, ``check_for_rebuild_vectors``,
``clear_event_counter``, ``t``, plus state and event accessors.  Failures raise
``RedError`` with the library's message, the way the reference throws
``nbody_exception`` / ``std::string`` (red.cuda/src/red.cuda/nbody_exception.h).

There is no fallback: if the HIP library is not built this module raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "librednative.so")

RED_OK = 0
ERR_NAMES = {1: "RED_ERR_ARG", 2: "RED_ERR_HIP", 3: "RED_ERR_EVENT_OVERFLOW", 4: "RED_ERR_NAN",
             5: "RED_ERR_UNSUPPORTED", 6: "RED_ERR_NO_STAR", 7: "RED_ERR_IO", 8: "RED_ERR_STEP_LIMIT"}
PREC_FP64, PREC_MIXED = 0, 1
ERR_UNSUPPORTED = 5
# stop reasons of red_run_fixed
FUSED_STOP_STEPS, FUSED_STOP_TIME, FUSED_STOP_OUTPUT, FUSED_STOP_EJECTION, FUSED_STOP_EVENTS = 1, 2, 3, 4, 5


class RedError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class RedEvent(ctypes.Structure):
    _fields_ = [("name", ctypes.c_int), ("id1", ctypes.c_int), ("idx1", ctypes.c_int),
                ("id2", ctypes.c_int), ("idx2", ctypes.c_int), ("attempt", ctypes.c_int),
                ("stage", ctypes.c_int), ("sink", ctypes.c_int), ("source", ctypes.c_int),
                ("pad", ctypes.c_int), ("t", ctypes.c_double), ("d", ctypes.c_double),
                ("p1", ctypes.c_double * 4), ("p2", ctypes.c_double * 4), ("ps", ctypes.c_double * 4),
                ("r1", ctypes.c_double * 3), ("v1", ctypes.c_double * 3),
                ("r2", ctypes.c_double * 3), ("v2", ctypes.c_double * 3),
                ("rs", ctypes.c_double * 3), ("vs", ctypes.c_double * 3),
                ("oe1", ctypes.c_double * 6), ("oe2", ctypes.c_double * 6), ("oes", ctypes.c_double * 6)]


class RedParams(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int), ("adaptive", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("error_check_for_tp", ctypes.c_int), ("close_encounter", ctypes.c_int),
                ("inner_steps", ctypes.c_int), ("frame_center", ctypes.c_int),
                ("hit_centrum", ctypes.c_double), ("ejection", ctypes.c_double),
                ("radii_enhance_factor", ctypes.c_double), ("encounter_hill_factor", ctypes.c_double),
                ("dt_day", ctypes.c_double), ("precision", ctypes.c_int), ("event_capacity", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("kahan_pairs", ctypes.c_int), ("device", ctypes.c_int),
                ("block", ctypes.c_int), ("sinks_per_thread", ctypes.c_int), ("strategy", ctypes.c_int),
                ("fp64_split", ctypes.c_int), ("source_parts", ctypes.c_int), ("host_threads", ctypes.c_int),
                ("padded_storage", ctypes.c_int)]


class RedBodies(ctypes.Structure):
    _fields_ = [("counts", ctypes.c_int * 7), ("id", ctypes.c_void_p), ("type", ctypes.c_void_p),
                ("mig_type", ctypes.c_void_p), ("mig_stop_at", ctypes.c_void_p), ("epoch", ctypes.c_void_p),
                ("mass", ctypes.c_void_p), ("radius", ctypes.c_void_p), ("density", ctypes.c_void_p),
                ("cd", ctypes.c_void_p), ("r", ctypes.c_void_p), ("v", ctypes.c_void_p)]


class RedStepInfo(ctypes.Structure):
    _fields_ = [("dt_did", ctypes.c_double), ("dt_next", ctypes.c_double), ("err", ctypes.c_double),
                ("tries", ctypes.c_int), ("n_events", ctypes.c_int)]


# Every entry point the header declares, with its ctypes signature.
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "red_ctx_create": (_I, [ctypes.POINTER(RedBodies), ctypes.POINTER(RedParams), ctypes.POINTER(_P)]),
    "red_ctx_destroy": (_I, [_P]),
    "red_last_error": (ctypes.c_char_p, [_P]),
    "red_get_counts": (_I, [_P, _IP]),
    "red_get_time": (_I, [_P, _DP]),
    "red_set_time": (_I, [_P, _D]),
    "red_get_dt": (_I, [_P, _DP]),
    "red_set_dt": (_I, [_P, _D]),
    "red_get_state": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "red_set_state": (_I, [_P, _P, _P]),
    "red_accel": (_I, [_P, _D, _I, _P, _P, _P]),
    "red_state_device": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "red_sync": (_I, [_P]),
    "red_step": (_I, [_P, ctypes.POINTER(RedStepInfo)]),
    "red_run_fixed": (_I, [_P, _I, _D, _D, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int)]),
    "red_check_ejection_hit_centrum": (_I, [_P, _IP]),
    "red_check_collision": (_I, [_P, _IP]),
    "red_check_close_encounter": (_I, [_P, _IP]),
    "red_check_rebuild": (_I, [_P, _I, _IP]),
    "red_clear_events": (_I, [_P]),
    "red_raw_events": (_I, [_P, ctypes.POINTER(RedEvent), _I, _IP, _IP]),
    "red_sp_events": (_I, [_P, ctypes.POINTER(RedEvent), _I, _IP]),
    "red_last_rebuild_map": (_I, [_P, _IP, _I, _IP]),
    "red_counters": (_I, [_P, _IP]),
    "red_energy": (_I, [_P, _DP, _DP]),
    "red_energy_noalloc": (_I, [_P, _P]),
    "red_host_transform_to_bc": (_I, [_I, _P, _P, _P]),
    "red_host_transform_to_ac": (_I, [_I, _P, _P, _P]),
    "red_host_transform_velocity": (_I, [_I, _P]),
    "red_set_profiling": (_I, [_P, _I]),
    "red_get_profile": (_I, [_P, _DP, _IP, _I]),
    "red_device_alloc": (_I, [_P, ctypes.c_ulonglong, ctypes.POINTER(_P)]),
    "red_device_free": (_I, [_P, _P]),
    "red_copy_to_device": (_I, [_P, _P, _P, ctypes.c_ulonglong]),
    "red_copy_to_host": (_I, [_P, _P, _P, ctypes.c_ulonglong]),
    "red_ensemble_create": (_I, [_I, _P, ctypes.POINTER(RedParams), ctypes.POINTER(_P)]),
    "red_ensemble_destroy": (_I, [_P]),
    "red_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "red_ensemble_set_time": (_I, [_P, _D]),
    "red_ensemble_advance": (_I, [_P, _D, _I, _P]),
    "red_ensemble_get_state": (_I, [_P, _I, _P, _P]),
    "red_ensemble_get_bodies": (_I, [_P, _I, _IP, _P, _P, _P, _P, _P, _P]),
    "red_ensemble_disk_counters": (_I, [_P, _I, _IP]),
    "red_ensemble_disk_events": (_I, [_P, _I, ctypes.POINTER(RedEvent), _I, _IP]),
    "red_ensemble_disk_info": (_I, [_P, _I, _IP, _DP, _DP]),
    "red_ensemble_autotune": (_I, [_P, _I, _IP, _IP, _DP]),
    "red_ensemble_set_batch": (_I, [_P, _I]),
    "red_step_pow": (_I, [_P, _P, _P, _I, _I, _I]),
    "red_valu_rates": (_I, [_I, _P, _I]),
    "red_mlp_param_count": (_I, []),
    "red_mlp_create": (_I, [_I, _P, ctypes.POINTER(_P)]),
    "red_mlp_destroy": (_I, [_P]),
    "red_mlp_last_error": (ctypes.c_char_p, [_P]),
    "red_mlp_set_normalizer": (_I, [_P, _P, _P]),
    "red_mlp_get_params": (_I, [_P, _P, _P, _P]),
    "red_mlp_set_loss": (_I, [_P, _I, ctypes.c_float]),
    "red_mlp_predict": (_I, [_P, _P, _I, _P]),
    "red_mlp_gradient": (_I, [_P, _P, _P, _I, _P, _P]),
    "red_mlp_train": (_I, [_P, _P, _P, _I, _I, _I, ctypes.c_float, _P]),
    "red_mlp_adam_state": (_I, [_P, _P, _P, ctypes.POINTER(ctypes.c_long)]),
    "red_history_enable": (_I, [_P, _I]),
    "red_history_get": (_I, [_P, _P, _I, _IP]),
    "red_set_proposer": (_I, [_P, _P, _D, _D]),
    "red_proposer_stats": (_I, [_P, _IP]),
    "red_version": (ctypes.c_char_p, []),
}
HIST_FIELDS = 16
MLP_FEATURES = 12


class RedEnsembleInfo(ctypes.Structure):
    _fields_ = [("supersteps", ctypes.c_int), ("active_disks", ctypes.c_int), ("seconds", ctypes.c_double),
                ("event_seconds", ctypes.c_double), ("rebuild_seconds", ctypes.c_double), ("batches", ctypes.c_int),
                ("full_batches", ctypes.c_int)]

_lib = None


def load_library(path: Optional[str] = None):
    """Loads librednative.so; raises if it is not built (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run make -C {PKG}")
    L = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def event_dict(e: RedEvent) -> dict:
    d = {f: getattr(e, f) for f, _ in RedEvent._fields_ if f != "pad"}
    for k, v in list(d.items()):
        if not isinstance(v, (int, float)):
            d[k] = tuple(v)
    return d


def host_transform_to_ac(btype: np.ndarray, r: np.ndarray, v: np.ndarray) -> None:
    """Astro-centric shift in place (tools.cpp:209-236); body 0 must be the star."""
    L = load_library()
    t = np.ascontiguousarray(btype, np.int32)
    rc = L.red_host_transform_to_ac(len(t), _ptr(t), _ptr(r), _ptr(v))
    if rc:
        raise RedError(rc, "red_host_transform_to_ac")


def host_transform_to_bc(mass: np.ndarray, r: np.ndarray, v: np.ndarray) -> None:
    L = load_library()
    rc = L.red_host_transform_to_bc(len(mass), _ptr(mass), _ptr(r), _ptr(v))
    if rc:
        raise RedError(rc, "red_host_transform_to_bc")


def host_transform_velocity(v: np.ndarray) -> None:
    L = load_library()
    rc = L.red_host_transform_velocity(len(v), _ptr(v))
    if rc:
        raise RedError(rc, "red_host_transform_velocity")


class Disk:
    """One disk integrated on the GPU through the C ABI (internal units)."""

    def __init__(self, bodies, integrator=0, adaptive=False, tolerance=1e-10, dt_day=0.001,
                 error_check_for_tp=False, close_encounter=False, inner_steps=True, frame_center=1,
                 hit_centrum=0.0, ejection=0.0, radii_enhance_factor=0.0, encounter_hill_factor=0.0,
                 precision=PREC_FP64, event_capacity=0, use_graph=False, kahan_pairs=False, device=0,
                 block=0, sinks_per_thread=0, strategy=0, fp64_split=False, source_parts=0, padded_storage=False,
                 r=None, v=None):
        self.L = load_library()
        self.ctx = ctypes.c_void_p()
        prm = RedParams()
        prm.integrator = integrator
        prm.adaptive = int(adaptive)
        prm.tolerance = tolerance
        prm.error_check_for_tp = int(error_check_for_tp)
        prm.close_encounter = int(close_encounter)
        prm.inner_steps = int(inner_steps)
        prm.frame_center = frame_center
        prm.hit_centrum = hit_centrum
        prm.ejection = ejection
        prm.radii_enhance_factor = radii_enhance_factor
        prm.encounter_hill_factor = encounter_hill_factor
        prm.dt_day = dt_day
        prm.precision = precision
        prm.event_capacity = event_capacity
        prm.use_graph = int(use_graph)
        prm.kahan_pairs = int(kahan_pairs)
        prm.device = device
        prm.block = block
        prm.sinks_per_thread = sinks_per_thread
        prm.strategy = strategy
        prm.fp64_split = int(fp64_split)
        prm.source_parts = source_parts
        prm.padded_storage = int(padded_storage)
        self._keep = dict(
            id=np.ascontiguousarray(bodies.ids, dtype=np.int32),
            type=np.ascontiguousarray(bodies.types, dtype=np.int32),
            mig_type=np.ascontiguousarray(bodies.mig_type, dtype=np.int32),
            mig_stop_at=np.ascontiguousarray(bodies.mig_stop_at, dtype=np.float64),
            epoch=np.ascontiguousarray(bodies.epoch, dtype=np.float64),
            mass=np.ascontiguousarray(bodies.mass, dtype=np.float64),
            radius=np.ascontiguousarray(bodies.radius, dtype=np.float64),
            density=np.ascontiguousarray(bodies.density, dtype=np.float64),
            cd=np.ascontiguousarray(bodies.cd, dtype=np.float64),
            r=np.ascontiguousarray(bodies.r if r is None else r, dtype=np.float64),
            v=np.ascontiguousarray(bodies.v if v is None else v, dtype=np.float64),
        )
        b = RedBodies()
        for k in range(7):
            b.counts[k] = int(bodies.counts[k])
        for k, a in self._keep.items():
            setattr(b, k, _ptr(a))
        rc = self.L.red_ctx_create(ctypes.byref(b), ctypes.byref(prm), ctypes.byref(self.ctx))
        if rc:
            msg = self.L.red_last_error(self.ctx).decode() if self.ctx else ""
            self.close()
            raise RedError(rc, msg or "red_ctx_create failed")
        self._keep = None
        self.params = prm

    def close(self):
        if getattr(self, "ctx", None) and self.ctx:
            self.L.red_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise RedError(rc, self.L.red_last_error(self.ctx).decode())

    # ---- counts / time ---------------------------------------------------------
    def counts(self):
        c = np.zeros(7, dtype=np.int32)
        self._chk(self.L.red_get_counts(self.ctx, c.ctypes.data_as(_IP)))
        return [int(x) for x in c]

    @property
    def n(self):
        return int(sum(self.counts()))

    @property
    def t(self):
        x = ctypes.c_double()
        self._chk(self.L.red_get_time(self.ctx, ctypes.byref(x)))
        return x.value

    @t.setter
    def t(self, value):
        self._chk(self.L.red_set_time(self.ctx, float(value)))

    @property
    def dt_try(self):
        x = ctypes.c_double()
        self._chk(self.L.red_get_dt(self.ctx, ctypes.byref(x)))
        return x.value

    @dt_try.setter
    def dt_try(self, value):
        self._chk(self.L.red_set_dt(self.ctx, float(value)))

    # ---- state --------------------------------------------------------------------
    def state(self):
        n = self.n
        r = np.zeros((n, 3)); v = np.zeros((n, 3))
        m = np.zeros(n); R = np.zeros(n); rho = np.zeros(n); cd = np.zeros(n)
        ids = np.zeros(n, dtype=np.int32); types = np.zeros(n, dtype=np.int32)
        self._chk(self.L.red_get_state(self.ctx, _ptr(r), _ptr(v), _ptr(m), _ptr(R), _ptr(rho), _ptr(cd),
                                       _ptr(ids), _ptr(types)))
        return dict(r=r, v=v, p=np.stack([m, R, rho, cd], axis=1), id=ids, type=types)

    def set_state(self, r, v):
        r = np.ascontiguousarray(r, dtype=np.float64); v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk(self.L.red_set_state(self.ctx, _ptr(r), _ptr(v)))

    def state_device(self):
        rp, vp = ctypes.c_void_p(), ctypes.c_void_p()
        self._chk(self.L.red_state_device(self.ctx, ctypes.byref(rp), ctypes.byref(vp)))
        return rp.value, vp.value

    def sync(self):
        self._chk(self.L.red_sync(self.ctx))

    def accel_device(self, t, stage, r_dev, v_dev, a_dev):
        """red_accel on caller-owned device buffers (raw addresses)."""
        self._chk(self.L.red_accel(self.ctx, float(t), int(stage), ctypes.c_void_p(r_dev),
                                   ctypes.c_void_p(v_dev), ctypes.c_void_p(a_dev)))

    def device_alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._chk(self.L.red_device_alloc(self.ctx, nbytes, ctypes.byref(p)))
        return p.value

    def device_free(self, dev: int):
        self._chk(self.L.red_device_free(self.ctx, ctypes.c_void_p(dev)))

    def to_device(self, dev: int, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self._chk(self.L.red_copy_to_device(self.ctx, ctypes.c_void_p(dev), _ptr(host), host.nbytes))

    def to_host(self, host: np.ndarray, dev: int):
        self._chk(self.L.red_copy_to_host(self.ctx, _ptr(host), ctypes.c_void_p(dev), host.nbytes))

    def accel(self, t, r, v, stage=0):
        """Force evaluation through red_accel on device buffers of the library's allocator."""
        n = len(r)
        rr = np.zeros((n, 4)); vv = np.zeros((n, 4))
        rr[:, :3] = r; vv[:, :3] = v
        rr[:, 3] = self.state()["p"][:, 0]
        nb = n * 32
        bufs = [self.device_alloc(nb) for _ in range(3)]
        try:
            self.to_device(bufs[0], rr)
            self.to_device(bufs[1], vv)
            self.to_device(bufs[2], np.full((n, 4), np.nan))
            self.accel_device(t, stage, bufs[0], bufs[1], bufs[2])
            self.sync()
            out = np.empty((n, 4))
            self.to_host(out, bufs[2])
        finally:
            for b in bufs:
                self.device_free(b)
        return out[:, :3].copy()

    # ---- integration and events -----------------------------------------------------
    def step(self) -> float:
        info = RedStepInfo()
        self._chk(self.L.red_step(self.ctx, ctypes.byref(info)))
        self.last_step = info
        return info.dt_did

    def run_fixed(self, max_steps: int, t_stop: float, output_interval: float, ps: float = 0.0):
        """Fixed-step main-loop iterations on the device (red_run_fixed): runs until one
        needs the host and returns (steps done, accumulated output step ps, stop reason)."""
        p = ctypes.c_double(ps); k = ctypes.c_int(0); why = ctypes.c_int(0)
        self._chk(self.L.red_run_fixed(self.ctx, int(max_steps), float(t_stop), float(output_interval),
                                       ctypes.byref(p), ctypes.byref(k), ctypes.byref(why)))
        return k.value, p.value, why.value

    def _flag(self, fn, *args):
        h = ctypes.c_int(0)
        self._chk(fn(self.ctx, *args, ctypes.byref(h)))
        return bool(h.value)

    def check_for_ejection_hit_centrum(self) -> bool:
        return self._flag(self.L.red_check_ejection_hit_centrum)

    def check_for_collision(self) -> bool:
        return self._flag(self.L.red_check_collision)

    def check_for_close_encounter(self) -> bool:
        return self._flag(self.L.red_check_close_encounter)

    def check_for_rebuild_vectors(self, n: int) -> bool:
        return self._flag(self.L.red_check_rebuild, int(n))

    def clear_event_counter(self):
        self._chk(self.L.red_clear_events(self.ctx))

    def raw_events(self):
        n = ctypes.c_int(0); ov = ctypes.c_int(0)
        self._chk(self.L.red_raw_events(self.ctx, None, 0, ctypes.byref(n), ctypes.byref(ov)))
        if ov.value:
            raise RedError(3, f"{n.value} pair events exceed the capacity")
        buf = (RedEvent * max(n.value, 1))()
        self._chk(self.L.red_raw_events(self.ctx, buf, n.value, ctypes.byref(n), ctypes.byref(ov)))
        return [event_dict(buf[k]) for k in range(n.value)]

    def n_event(self):
        n = ctypes.c_int(0); ov = ctypes.c_int(0)
        self._chk(self.L.red_raw_events(self.ctx, None, 0, ctypes.byref(n), ctypes.byref(ov)))
        return n.value

    def sp_events(self):
        n = ctypes.c_int(0)
        self._chk(self.L.red_sp_events(self.ctx, None, 0, ctypes.byref(n)))
        buf = (RedEvent * max(n.value, 1))()
        self._chk(self.L.red_sp_events(self.ctx, buf, n.value, ctypes.byref(n)))
        return [event_dict(buf[k]) for k in range(n.value)]

    def rebuild_map(self):
        n = ctypes.c_int(0)
        self._chk(self.L.red_last_rebuild_map(self.ctx, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        self._chk(self.L.red_last_rebuild_map(self.ctx, out.ctypes.data_as(_IP), n.value, ctypes.byref(n)))
        return out[:n.value]

    def counters(self):
        c = np.zeros(8, dtype=np.int32)
        self._chk(self.L.red_counters(self.ctx, c.ctypes.data_as(_IP)))
        keys = ("tried", "failed", "passed", "n_collision", "n_ejection", "n_hit_centrum",
                "n_close_encounter", "n_event_last_clear")
        return dict(zip(keys, (int(x) for x in c)))

    PROF_CLASSES = ("si", "nsi", "ni", "stage", "err", "pack", "learn", "eject")

    def set_profiling(self, on: bool):
        self._chk(self.L.red_set_profiling(self.ctx, int(on)))

    def profile(self, reset=True):
        ms = (ctypes.c_double * 8)(); cnt = (ctypes.c_int * 8)()
        self._chk(self.L.red_get_profile(self.ctx, ms, cnt, int(reset)))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.PROF_CLASSES)}

    def energy(self):
        E = ctypes.c_double(); L = (ctypes.c_double * 3)()
        self._chk(self.L.red_energy(self.ctx, ctypes.byref(E), L))
        return E.value, np.array(list(L))

    # ---- learned step-size proposer -------------------------------------------------
    def history_enable(self, capacity: int):
        self._chk(self.L.red_history_enable(self.ctx, int(capacity)))

    def history(self, cap: int = 1 << 20) -> np.ndarray:
        """Step records (n, 16) float32, oldest first; the buffer is emptied."""
        out = []
        while True:
            buf = np.zeros((min(cap, 65536), HIST_FIELDS), dtype=np.float32)
            n = ctypes.c_int(0)
            self._chk(self.L.red_history_get(self.ctx, _ptr(buf), buf.shape[0], ctypes.byref(n)))
            out.append(buf[: n.value])
            cap -= n.value
            if n.value < buf.shape[0] or cap <= 0:
                break
        return np.concatenate(out) if out else np.zeros((0, HIST_FIELDS), np.float32)

    def set_proposer(self, net: Optional["Mlp"], safety: float = 0.9, max_log_change: float = 3.0):
        self._proposer = net          # the network must outlive its use
        self._chk(self.L.red_set_proposer(self.ctx, net.h if net is not None else None, float(safety),
                                          float(max_log_change)))

    def proposer_stats(self):
        o = np.zeros(2, dtype=np.int32)
        self._chk(self.L.red_proposer_stats(self.ctx, o.ctypes.data_as(_IP)))
        return dict(proposals=int(o[0]), first_accepted=int(o[1]))


def _params(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001, error_check_for_tp=False,
            use_graph=True, device=0, block=0, frame_center=1, precision=PREC_FP64, radii_enhance_factor=0.0,
            encounter_hill_factor=0.0, close_encounter=False, ejection=0.0, hit_centrum=0.0, inner_steps=True,
            event_capacity=0, host_threads=0, padded_storage=False):
    """Ensemble parameters; every option is passed on, so that the library can refuse the ones
    an ensemble does not support instead of their being dropped here."""
    prm = RedParams()
    prm.integrator = integrator
    prm.adaptive = int(adaptive)
    prm.tolerance = tolerance
    prm.error_check_for_tp = int(error_check_for_tp)
    prm.inner_steps = int(inner_steps)
    prm.event_capacity = event_capacity
    prm.frame_center = frame_center
    prm.precision = precision
    prm.radii_enhance_factor = radii_enhance_factor
    prm.encounter_hill_factor = encounter_hill_factor
    prm.close_encounter = int(close_encounter)
    prm.ejection = ejection
    prm.hit_centrum = hit_centrum
    prm.dt_day = dt_day
    prm.use_graph = int(use_graph)
    prm.device = device
    prm.block = block
    prm.host_threads = host_threads
    prm.padded_storage = int(padded_storage)
    return prm


class Ensemble:
    """Independent disks integrated together through red_ensemble_* (internal units)."""

    def __init__(self, disks, batch=None, **kw):
        self.L = load_library()
        self.h = ctypes.c_void_p()
        self.sizes = [d.n for d in disks]
        keep = []
        arr = (RedBodies * len(disks))()
        for k, bodies in enumerate(disks):
            a = dict(id=np.ascontiguousarray(bodies.ids, dtype=np.int32),
                     type=np.ascontiguousarray(bodies.types, dtype=np.int32),
                     mass=np.ascontiguousarray(bodies.mass, dtype=np.float64),
                     radius=np.ascontiguousarray(bodies.radius, dtype=np.float64),
                     density=np.ascontiguousarray(bodies.density, dtype=np.float64),
                     cd=np.ascontiguousarray(bodies.cd, dtype=np.float64),
                     r=np.ascontiguousarray(bodies.r, dtype=np.float64),
                     v=np.ascontiguousarray(bodies.v, dtype=np.float64))
            keep.append(a)
            for c in range(7):
                arr[k].counts[c] = int(bodies.counts[c])
            for f, x in a.items():
                setattr(arr[k], f, _ptr(x))
        prm = _params(**kw)
        rc = self.L.red_ensemble_create(len(disks), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(prm),
                                        ctypes.byref(self.h))
        if rc:
            msg = self.L.red_ensemble_last_error(self.h).decode() if self.h else ""
            self.close()
            raise RedError(rc, msg or "red_ensemble_create failed")
        if batch is not None:
            self.set_batch(batch)

    def close(self):
        if getattr(self, "h", None) and self.h:
            self.L.red_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise RedError(rc, self.L.red_ensemble_last_error(self.h).decode())

    def set_time(self, t):
        self._chk(self.L.red_ensemble_set_time(self.h, float(t)))

    def advance(self, t_stop, max_attempts=1 << 30):
        info = RedEnsembleInfo()
        self._chk(self.L.red_ensemble_advance(self.h, float(t_stop), int(max_attempts), ctypes.byref(info)))
        return dict(supersteps=info.supersteps, active_disks=info.active_disks, seconds=info.seconds,
                    event_seconds=info.event_seconds, rebuild_seconds=info.rebuild_seconds, batches=info.batches,
                    full_batches=info.full_batches)

    def counts(self, disk):
        c = np.zeros(7, dtype=np.int32)
        self._chk(self.L.red_ensemble_get_bodies(self.h, disk, c.ctypes.data_as(_IP), None, None, None, None, None,
                                                 None))
        return [int(x) for x in c]

    def state(self, disk):
        n = sum(self.counts(disk))
        r = np.zeros((n, 3)); v = np.zeros((n, 3))
        self._chk(self.L.red_ensemble_get_state(self.h, disk, _ptr(r), _ptr(v)))
        return r, v

    def bodies(self, disk):
        """Parameters of a disk's current bodies: p = (mass, radius, density, cd), id, type."""
        n = sum(self.counts(disk))
        m = np.zeros(n); R = np.zeros(n); rho = np.zeros(n); cd = np.zeros(n)
        ids = np.zeros(n, dtype=np.int32); types = np.zeros(n, dtype=np.int32)
        self._chk(self.L.red_ensemble_get_bodies(self.h, disk, None, _ptr(m), _ptr(R), _ptr(rho), _ptr(cd),
                                                 _ptr(ids), _ptr(types)))
        return dict(p=np.stack([m, R, rho, cd], axis=1), id=ids, type=types)

    def counters(self, disk):
        c = np.zeros(8, dtype=np.int32)
        self._chk(self.L.red_ensemble_disk_counters(self.h, disk, c.ctypes.data_as(_IP)))
        keys = ("tried", "failed", "passed", "n_collision", "n_ejection", "n_hit_centrum",
                "n_close_encounter", "n_event_last_clear")
        return dict(zip(keys, (int(x) for x in c)))

    def drain_events(self, disk, buf):
        """Moves the disk's handled records (oldest first) into `buf`, a RedEvent array, as
        many as fit; returns (copied, left before the call)."""
        n = ctypes.c_int()
        self._chk(self.L.red_ensemble_disk_events(self.h, disk, buf, len(buf), ctypes.byref(n)))
        return min(n.value, len(buf)), n.value

    def events(self, disk):
        """Records of the event passes the disk handled since the last call, in order."""
        n = ctypes.c_int()
        self._chk(self.L.red_ensemble_disk_events(self.h, disk, None, 0, ctypes.byref(n)))
        buf = (RedEvent * max(n.value, 1))()
        self._chk(self.L.red_ensemble_disk_events(self.h, disk, buf, n.value, ctypes.byref(n)))
        return [event_dict(buf[k]) for k in range(n.value)]

    def disk_info(self, disk):
        c = np.zeros(3, dtype=np.int32); t = ctypes.c_double(); dt = ctypes.c_double()
        self._chk(self.L.red_ensemble_disk_info(self.h, disk, c.ctypes.data_as(_IP), ctypes.byref(t), ctypes.byref(dt)))
        return dict(tried=int(c[0]), failed=int(c[1]), passed=int(c[2]), t=t.value, dt_try=dt.value)

    def autotune(self, reps=5):
        b = ctypes.c_int(); j = ctypes.c_int(); ms = (ctypes.c_double * 9)()
        self._chk(self.L.red_ensemble_autotune(self.h, reps, ctypes.byref(b), ctypes.byref(j), ms))
        return b.value, j.value, np.array(list(ms)).reshape(3, 3)

    def set_batch(self, supersteps):
        self._chk(self.L.red_ensemble_set_batch(self.h, int(supersteps)))


def step_pow(x, y, on_device=False, device=0):
    """The step controllers' factor x^y rounded to nearest (red_step_pow), elementwise."""
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    rc = L.red_step_pow(_ptr(x), _ptr(y), _ptr(out), int(x.size), 1 if on_device else 0, int(device))
    if rc:
        raise RedError(rc, "red_step_pow failed")
    return out


VALU_RATE_NAMES = ("fma32", "mul32", "add64", "cvt", "rsq32", "fp32_fp64_waves")


def valu_rates(device=0):
    """Vector-ALU issue rates (red_valu_rates): lane-instructions per second per kind, and
    pairs per second of the mixed and the FP64 sweeps' pair sequences with sources in cache."""
    L = load_library()
    out = np.zeros(len(VALU_RATE_NAMES) + 2)
    rc = L.red_valu_rates(int(device), _ptr(out), int(out.size))
    if rc:
        raise RedError(rc, "red_valu_rates failed")
    d = {k: float(v) for k, v in zip(VALU_RATE_NAMES, out)}
    d["mixed_pairs_per_s"] = float(out[-2])
    d["fp64_pairs_per_s"] = float(out[-1])
    return d


class Mlp:
    """The proposer network (red_mlp_*): 12 -> 64 -> 64 -> 1 GELU, BF16 weights, FP32 sums."""

    NF = MLP_FEATURES

    def __init__(self, params: Optional[np.ndarray] = None, device: int = 0, norm: Optional[np.ndarray] = None):
        self.L = load_library()
        self.h = ctypes.c_void_p()
        self.n_param = int(self.L.red_mlp_param_count())
        p = None
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.shape != (self.n_param,):
                raise ValueError(f"expected {self.n_param} parameters, got {p.shape}")
        rc = self.L.red_mlp_create(int(device), _ptr(p), ctypes.byref(self.h))
        if rc:
            msg = self.L.red_mlp_last_error(self.h).decode() if self.h else ""
            self.close()
            raise RedError(rc, msg or "red_mlp_create failed")
        if norm is not None:
            self.set_normalizer(norm[: self.NF], norm[self.NF:])

    def close(self):
        if getattr(self, "h", None) and self.h:
            self.L.red_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise RedError(rc, self.L.red_mlp_last_error(self.h).decode())

    def set_normalizer(self, mean, inv_std):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(inv_std, dtype=np.float32)
        self._chk(self.L.red_mlp_set_normalizer(self.h, _ptr(m), _ptr(s)))

    LOSS_MSE, LOSS_PINBALL = 0, 1

    def set_loss(self, kind: int, tau: float = 0.5):
        self._chk(self.L.red_mlp_set_loss(self.h, int(kind), ctypes.c_float(tau)))

    def params(self):
        """(FP32 master parameters, BF16 image as uint16, normaliser[24])."""
        w = np.zeros(self.n_param, np.float32); b = np.zeros(self.n_param, np.uint16)
        nrm = np.zeros(2 * self.NF, np.float32)
        self._chk(self.L.red_mlp_get_params(self.h, _ptr(w), _ptr(b), _ptr(nrm)))
        return w, b, nrm

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.NF)
        y = np.zeros(X.shape[0], np.float32)
        self._chk(self.L.red_mlp_predict(self.h, _ptr(X), X.shape[0], _ptr(y)))
        return y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.NF)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.n_param, np.float32); loss = np.zeros(1, np.float32)
        self._chk(self.L.red_mlp_gradient(self.h, _ptr(X), _ptr(T), X.shape[0], _ptr(g), _ptr(loss)))
        return float(loss[0]), g

    def train(self, X, T, epochs=1, batch=4096, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.NF)
        T = np.ascontiguousarray(T, dtype=np.float32)
        losses = np.zeros(epochs, np.float32)
        self._chk(self.L.red_mlp_train(self.h, _ptr(X), _ptr(T), X.shape[0], int(epochs), int(batch),
                                       ctypes.c_float(lr), _ptr(losses)))
        return losses

    def adam_state(self):
        m = np.zeros(self.n_param, np.float32); v = np.zeros(self.n_param, np.float32)
        st = ctypes.c_long(0)
        self._chk(self.L.red_mlp_adam_state(self.h, _ptr(m), _ptr(v), ctypes.byref(st)))
        return m, v, st.value


def mlp_param_count() -> int:
    return int(load_library().red_mlp_param_count())
