"""Integrator steps on the GPU (red_step) against the CPU oracle, FP64 mode.

Every scheme (Euler, RK2, RK4 fixed and adaptive, Cash-Karp RK5(4), RKF7(8) fixed and
adaptive) runs the same steps on both; positions, velocities, step sizes and the
tried/failed/passed counters must be bit-identical, with and without HIP-graph replay.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu

CASES = [
    ("euler", 0, False, 0.0, 0.5),
    ("rk2", 1, False, 0.0, 0.5),
    ("rk4", 2, False, 0.0, 1.0),
    ("rk4a", 2, True, 1e-9, 1.0),
    ("rk5a", 3, True, 1e-10, 1.0),
    ("rkf78", 4, False, 0.0, 2.0),
    ("rkf78a", 4, True, 1e-12, 2.0),
]


@pytest.mark.parametrize("name,integ,adaptive,tol,dt_day", CASES)
@pytest.mark.parametrize("graph", [False, True])
def test_steps_bit_exact(oracle_mod, native_mod, name, integ, adaptive, tol, dt_day, graph):
    print(f"case {name} graph={graph}", flush=True)
    b = generate.to_internal(generate.disk_small(120, seed=31))
    kw = dict(integrator=integ, adaptive=adaptive, tolerance=tol, dt_day=dt_day)
    o = oracle_mod.Oracle(b, **kw)
    d = native_mod.Disk(b, use_graph=graph, **kw)
    for k in range(12):
        dt_o = o.step()
        dt_g = d.step()
        assert dt_o == dt_g, (k, dt_o, dt_g)
        assert o.dt_try == d.dt_try, (k, o.dt_try, d.dt_try)
    so, sg = o.state(), d.state()
    assert_bitwise(sg["r"], so["r"], "r")
    assert_bitwise(sg["v"], so["v"], "v")
    assert o.t == d.t
    co, cg = o.counters(), d.counters()
    for key in ("tried", "failed", "passed"):
        assert co[key] == cg[key], (key, co, cg)


def test_mixed_rk5_tracks_fp64(native_mod):
    """Mixed-precision RK5(4) stays close to the FP64 trajectory over a few steps."""
    b = generate.to_internal(generate.disk_mixed(n_pl=2000, n_tp=2000, seed=12))
    kw = dict(integrator=3, adaptive=True, tolerance=1e-10, dt_day=1.0)
    d64 = native_mod.Disk(b, **kw)
    dmx = native_mod.Disk(b, precision=native_mod.PREC_MIXED, **kw)
    for _ in range(10):
        d64.step()
        dmx.step()
    s64, smx = d64.state(), dmx.state()
    dr = np.abs(smx["r"] - s64["r"]).max()
    dv = np.abs(smx["v"] - s64["v"]).max()
    print("max |dr|", dr, "max |dv|", dv)
    assert dr < 1e-10 and dv < 1e-10
