"""The oracle reproduces the reference's stored TwoBody run bit for bit.

Fixture: red.cuda/TestRun/InputTest/Debug/TwoBody/{TwoBody2.txt, parameters2.txt,
D_cpu_ns_fs_E_result.txt} (Euler, CPU, 2 yr, dt = 0.001 day, output every 0.1 yr).
The stored file predates the current printers: its epoch column holds t and its velocity
columns v in internal units (k = 0.01720209895), so t and v compare directly with the
oracle's internal state.  This pins the barycentric transform order, the unit scaling,
the FP64 force without fused multiply-adds, the Euler update and the driver cadence.
"""
import os

import numpy as np

from conftest import GOLDEN
from rednative import files, mainloop


def run_twobody(engine_factory, fused=True):
    b = files.read_body_file(os.path.join(GOLDEN, "TwoBody2.txt"))
    prm = files.read_parameters(os.path.join(GOLDEN, "parameters2.txt"))
    eng, bc = engine_factory(b, prm)
    eng.t = prm["start_time"]
    snaps = []

    def snap(e):
        s = e.state()
        snaps.append((e.t, s["r"].copy(), s["v"].copy()))

    info = mainloop.run_main_loop(eng, prm, snap, fused=fused)
    return snaps, info


def compare_golden(snaps):
    gold = files.parse_result_ascii(os.path.join(GOLDEN, "D_cpu_ns_fs_E_result.txt"))
    assert len(gold) == 42 and len(snaps) == 21
    n_vals = 0
    for k, (t, r, v) in enumerate(snaps):
        for i in range(2):
            row = gold[2 * k + i]
            assert float(row[3]) == t, (k, i, row[3], t)
            for c in range(3):
                assert float(row[10 + c]) == r[i, c], (k, i, c)
                assert float(row[13 + c]) == v[i, c], (k, i, c)
            n_vals += 7
    return n_vals


def test_oracle_twobody_euler_bit_exact(oracle_mod):
    def factory(b, prm):
        o = oracle_mod.Oracle(b, integrator=prm["integrator"], dt_day=prm["dt"],
                              stage_time_events=False, hc_reactivate=True)
        s = o.state()
        assert np.all(s["r"][0] == 0) and np.all(s["v"][0] == 0)
        o.transform_to_bc()
        o.transform_time_velocity()
        return o, True

    snaps, info = run_twobody(factory)
    assert info["steps"] == 730500
    assert compare_golden(snaps) == 294
