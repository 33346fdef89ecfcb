"""The reference driver loop, restated over any engine with the pp_disk/integrator surface.

Follows red.cuda/src/red.cuda/main.cu:159-281: initial snapshot, then while
``t <= stop_time``: snapshot when the accumulated step ``ps`` reaches the output
interval, ejection/hit-centrum check, one integrator step, then close-encounter
or collision handling and the rebuild check (threshold 8 events), and a final
snapshot if anything was integrated since the last one.  Wall-clock info lines
and log files are left to the caller.

An engine provides: ``t``, ``step()``, ``check_for_ejection_hit_centrum()``,
``check_for_collision()``, ``check_for_close_encounter()``,
``check_for_rebuild_vectors(n)``, ``sp_events()`` and ``clear_event_counter()``.
"""
from __future__ import annotations

from typing import Callable, Dict, Optional


# red_run_fixed's stop reasons and the error code of a configuration it does not run
_STOP_OUTPUT, _STOP_EJECTION, _STOP_EVENTS = 3, 4, 5
_ERR_UNSUPPORTED = 5


def run_main_loop(eng, prm: Dict, on_snapshot: Callable, on_events: Optional[Callable] = None,
                  max_steps: Optional[int] = None, fused: bool = True) -> Dict:
    on_events = on_events or (lambda evs: None)
    if not prm.get("inner_steps", True):
        raise NotImplementedError("inner_steps = false (outer close-encounter scan) is not supported")
    radii = prm.get("radii_enhance_factor", 0.0)
    hill = prm.get("encounter_hill_factor", 0.0)
    ps = 0.0
    dt = 0.0
    steps = 0
    rebuilds = 0
    # An engine with run_fixed (the GPU library, fixed-step integrators) runs the iterations
    # that need nothing from the host on the device and hands back the rest.
    run_fixed = getattr(eng, "run_fixed", None) if fused else None
    limit = (1 << 31) - 1 if max_steps is None else max_steps

    def after_step():
        nonlocal rebuilds
        if prm.get("close_encounter", False):
            if (0.0 < radii or 0.0 < hill) and eng.check_for_close_encounter():
                on_events(eng.sp_events())
                eng.clear_event_counter()
        else:
            if (0.0 < radii or 0.0 < hill) and eng.check_for_collision():
                on_events(eng.sp_events())
                eng.clear_event_counter()
            if eng.check_for_rebuild_vectors(8):
                rebuilds += 1

    on_snapshot(eng)
    while eng.t <= prm["stop_time"]:
        if steps >= limit:
            break
        if prm["output_interval"] <= abs(ps):
            ps = 0.0
            on_snapshot(eng)
        if run_fixed is not None:
            try:
                k, ps, why = run_fixed(limit - steps, prm["stop_time"], prm["output_interval"], ps)
            except RuntimeError as e:
                if getattr(e, "code", None) != _ERR_UNSUPPORTED:   # else: step one at a time
                    raise
                run_fixed = None
            else:
                steps += k
                if k > 0:
                    dt = eng.dt_try   # fixed steps: every step is dt_try
                if why == _STOP_EVENTS:   # a step recorded pair events: its checks are left to do
                    after_step()
                    continue
                if why != _STOP_EJECTION and not (why == _STOP_OUTPUT and k == 0):
                    continue        # time, output or step limit: the loop head decides
                # an ejection or hit-centrum is due, or an output with no step done
                # (output_interval <= 0): this iteration runs below
        if eng.check_for_ejection_hit_centrum():
            on_events(eng.sp_events())
            eng.clear_event_counter()
        dt = eng.step()
        steps += 1
        ps += abs(dt)
        after_step()
    if 0.0 < ps:
        on_snapshot(eng)
    return dict(steps=steps, rebuilds=rebuilds, last_dt=dt)
