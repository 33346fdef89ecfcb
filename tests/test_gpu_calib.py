"""The vector-ALU calibration (red_valu_rates) that the force sweeps' rooflines are read
against: every rate is positive and finite, single-kind rates stay within one instruction per
lane and clock of the GPU's 1024 SIMDs (2.4 GHz), the half-rate reciprocal square root runs
slower than an FMA, and the two pair sequences give plausible pair rates."""
import math

import pytest

pytestmark = pytest.mark.gpu

LANE_CLOCKS = 1024 * 16 * 2.4e9      # SIMDs x lanes x clock: one instruction per lane and clock


def test_valu_rates_plausible(native_mod):
    print("case red_valu_rates", flush=True)
    r = native_mod.valu_rates()
    print(r, flush=True)
    for k, v in r.items():
        assert math.isfinite(v) and v > 0.0, k
    for k in ("fma32", "mul32", "add64", "cvt"):
        assert 0.5 * LANE_CLOCKS < r[k] <= 1.05 * LANE_CLOCKS, (k, r[k])
    assert r["rsq32"] < 0.75 * r["fma32"]
    assert r["fp32_fp64_waves"] > r["fma32"]
    assert 1e12 < r["mixed_pairs_per_s"] < 1e13
    assert 1e11 < r["fp64_pairs_per_s"] < r["mixed_pairs_per_s"]


def test_valu_rates_rejects_short_output(native_mod):
    import ctypes

    import numpy as np
    L = native_mod.load_library()
    out = np.zeros(3)
    rc = L.red_valu_rates(0, out.ctypes.data_as(ctypes.c_void_p), 3)
    assert rc != 0
