"""Edge shapes on the GPU against the CPU oracle: tiny and ragged body counts (sinks and
sources that fill no block, tile or chunk evenly), disks with empty body classes, a star
alone, and an empty body list.  FP64 results must equal the oracle's bit for bit; the
mixed path stays within its 2e-6 per-body bound."""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return (np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)).max()


@pytest.mark.parametrize("n", [2, 3, 17, 64, 65, 257, 1025, 4099])
def test_ragged_allsi_fp64(oracle_mod, native_mod, n):
    print(f"case allsi n={n}", flush=True)
    b = generate.to_internal(generate.disk_allsi(n, seed=n))
    o = oracle_mod.Oracle(b, radii_enhance_factor=1.0)
    d = native_mod.Disk(b, radii_enhance_factor=1.0)
    assert_bitwise(d.accel(0.25, b.r, b.v), o.accel(0.25, b.r, b.v), "accel")


@pytest.mark.parametrize("n_pl,n_tp", [(0, 7), (5, 0), (1, 1), (63, 65), (2049, 3), (3, 4100)])
def test_body_classes_fp64_and_mixed(oracle_mod, native_mod, n_pl, n_tp):
    print(f"case mixed classes pl={n_pl} tp={n_tp}", flush=True)
    b = generate.to_internal(generate.disk_mixed(n_pl=n_pl, n_tp=n_tp, seed=n_pl + 7 * n_tp))
    o = oracle_mod.Oracle(b, radii_enhance_factor=1.0)
    a_o = o.accel(0.5, b.r, b.v)
    d = native_mod.Disk(b, radii_enhance_factor=1.0)
    assert_bitwise(d.accel(0.5, b.r, b.v), a_o, "fp64")
    d.close()
    m = native_mod.Disk(b, precision=native_mod.PREC_MIXED, radii_enhance_factor=1.0)
    assert _rel(m.accel(0.5, b.r, b.v), a_o) < 2e-6


def test_star_alone(oracle_mod, native_mod):
    b = generate.to_internal(generate.disk_allsi(1, seed=1))
    assert b.n == 1
    d = native_mod.Disk(b, integrator=2, dt_day=1.0)
    assert np.all(d.accel(0.0, b.r, b.v) == 0.0)
    for _ in range(3):
        d.step()
    s = d.state()
    assert_bitwise(s["r"], b.r, "r")
    assert_bitwise(s["v"], b.v, "v")


def test_empty_body_list(native_mod):
    b = generate.to_internal(generate.disk_allsi(2, seed=1))
    e = b.copy()
    e.counts = [0] * 7
    for k in ("ids", "types", "mig_type"):
        setattr(e, k, getattr(b, k)[:0])
    for k in ("epoch", "mass", "radius", "density", "cd", "mig_stop_at"):
        setattr(e, k, getattr(b, k)[:0])
    e.names = []
    e.r = b.r[:0]
    e.v = b.v[:0]
    d = native_mod.Disk(e, integrator=2, dt_day=1.0)
    assert d.n == 0
    t0 = d.t
    d.step()
    assert d.t > t0
    assert d.state()["r"].shape == (0, 3)
