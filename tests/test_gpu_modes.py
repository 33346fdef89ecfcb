"""Less common options of the driver loop on the GPU against the CPU oracle, in lock step:
the error norm over test particles too (`error_check_for_tp`, integrator.cu:211-256 with
int_rungekutta8.cu:766-774), RK2 and Euler with every event kind, and close-encounter mode
under RKF7(8) step control with the Hill test (events of rejected attempts discarded,
int_rungekutta8.cu:730-733).  States, step sizes, counters and event records must be
identical after every check."""
import numpy as np
import pytest

from conftest import assert_bitwise
from test_gpu_events import events_equal, lockstep, planted_disk

pytestmark = pytest.mark.gpu

EVENTS = dict(hit_centrum=0.3, ejection=6.0, radii_enhance_factor=2500.0)
CASES = {
    "tp_error_rkf78": dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5, error_check_for_tp=True, **EVENTS),
    "rk2": dict(integrator=1, dt_day=0.25, **EVENTS),
    "euler": dict(integrator=0, dt_day=0.1, **EVENTS),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_loop_options_bit_exact(oracle_mod, native_mod, name):
    print(f"case {name}", flush=True)
    kw = CASES[name]
    b = planted_disk(seed=46)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, **kw)
    n_events, cnt = lockstep(o, d, 30)
    print("events", n_events, cnt, flush=True)
    assert cnt["n_collision"] > 0 and cnt["n_ejection"] > 0 and cnt["n_hit_centrum"] > 0


def test_close_encounter_mode_rkf78_hill(oracle_mod, native_mod):
    kw = dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5, close_encounter=True,
              radii_enhance_factor=1.0, encounter_hill_factor=10.0)
    print("case ce rkf78 hill", flush=True)
    b = planted_disk(seed=46)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, event_capacity=20000, **kw)
    seen = 0
    for k in range(30):
        dto, dtg = o.step(), d.step()
        assert dto == dtg, (k, dto, dtg)
        ho, hg = o.check_for_close_encounter(), d.check_for_close_encounter()
        assert ho == hg, k
        if ho:
            eo, eg = o.sp_events(), d.sp_events()
            events_equal(eo, eg, f"ce step {k}")
            seen += len(eg)
            o.clear_event_counter(); d.clear_event_counter()
        so, sg = o.state(), d.state()
        assert_bitwise(sg["r"], so["r"], f"r step {k}")
        assert_bitwise(sg["v"], so["v"], f"v step {k}")
    assert o.counters() == d.counters()
    assert seen > 0 and o.counters()["failed"] > 0     # a rejected attempt's events were dropped
