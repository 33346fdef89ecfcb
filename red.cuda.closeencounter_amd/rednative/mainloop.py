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


def run_main_loop(eng, prm: Dict, on_snapshot: Callable, on_events: Optional[Callable] = None,
                  max_steps: Optional[int] = None) -> Dict:
    on_events = on_events or (lambda evs: None)
    if not prm.get("inner_steps", True):
        raise NotImplementedError("inner_steps = false (outer close-encounter scan) is not supported")
    radii = prm.get("radii_enhance_factor", 0.0)
    hill = prm.get("encounter_hill_factor", 0.0)
    ps = 0.0
    dt = 0.0
    steps = 0
    rebuilds = 0
    on_snapshot(eng)
    while eng.t <= prm["stop_time"]:
        if max_steps is not None and steps >= max_steps:
            break
        if prm["output_interval"] <= abs(ps):
            ps = 0.0
            on_snapshot(eng)
        if eng.check_for_ejection_hit_centrum():
            on_events(eng.sp_events())
            eng.clear_event_counter()
        dt = eng.step()
        steps += 1
        ps += abs(dt)
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
    if 0.0 < ps:
        on_snapshot(eng)
    return dict(steps=steps, rebuilds=rebuilds, last_dt=dt)
