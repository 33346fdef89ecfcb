"""The step controllers' factor (tol/err)^(1/(p+1)) (int_rungekutta8.cu:777,
int_rungekutta4.cu:253): the oracle takes it correctly rounded through libquadmath's powq,
the product through its own double-double step_pow, which the host controller and the
ensembles' device controller share.  Both are pinned here against Python's decimal module
at 60 digits, rounded to the nearest double, on inputs spread over the controller's range
and on the hard cases next to x = 1 (x = 1 - 2^-53 with y = 1/2 lies 2^-110 below a midpoint);
the product's host value must equal the oracle's on many more inputs.  The platform's pow
(glibc here) is reported for comparison only: it is not correctly rounded everywhere."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

EXPONENTS = (1.0 / 4.0, 1.0 / 5.0, 1.0 / 8.0, 1.0 / 2.0)


def _rounded(x, y):
    getcontext().prec = 60
    return float((Decimal(x).ln() * Decimal(y)).exp())


def _samples(n, seed):
    g = np.random.default_rng(seed)
    x = 10.0 ** g.uniform(-40.0, 40.0, n)
    near = g.integers(-64, 65, n // 8)
    x[: n // 8] = 1.0 + near * 2.0 ** -52           # a few ulps around 1
    y = np.array(EXPONENTS)[g.integers(0, 4, n)]
    return x, y


def test_oracle_factor_is_correctly_rounded(oracle_mod):
    x, y = _samples(4000, 1)
    bad = [(a, b) for a, b in zip(x, y) if oracle_mod.step_factor(a, b) != _rounded(a, b)]
    assert not bad, bad[:5]


def test_product_factor_is_correctly_rounded(native_mod):
    x, y = _samples(4000, 2)
    got = native_mod.step_pow(x, y)
    want = np.array([_rounded(a, b) for a, b in zip(x, y)])
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:5]


def test_hard_cases_next_to_one(native_mod, oracle_mod):
    # x = 1 - 2^-53, y = 1/2: sqrt lies 2^-110 below the midpoint 1 - 2^-54, so rounds down
    x = 1.0 - 2.0 ** -53
    assert native_mod.step_pow(np.array([x]), 0.5)[0] == x
    assert oracle_mod.step_factor(x, 0.5) == x
    xs = 1.0 + np.arange(-40, 41) * 2.0 ** -52
    for y in EXPONENTS:
        got = native_mod.step_pow(xs, y)
        want = np.array([_rounded(a, y) for a in xs])
        assert np.array_equal(got, want), (y, np.flatnonzero(got != want))


def test_special_values(native_mod):
    x = np.array([0.0, math.inf, 1.0, -1.0, math.nan, 5e-324, 1.7976931348623157e308])
    got = native_mod.step_pow(x, 0.125)
    assert got[0] == 0.0 and got[1] == math.inf and got[2] == 1.0
    assert math.isnan(got[3]) and math.isnan(got[4])
    assert got[5] == _rounded(5e-324, 0.125) and got[6] == _rounded(1.7976931348623157e308, 0.125)


def test_product_host_equals_oracle_at_scale(native_mod, oracle_mod):
    x, y = _samples(200000, 3)
    got = native_mod.step_pow(x, y)
    want = np.array([oracle_mod.step_factor(a, b) for a, b in zip(x, y)])
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:5]
    glibc = np.array([math.pow(a, b) for a, b in zip(x, y)])
    print(f"platform pow differs from the correctly rounded factor in {int((glibc != want).sum())} of {x.size}")


@pytest.mark.gpu
def test_device_factor_equals_host(native_mod):
    x, y = _samples(1 << 20, 4)
    host = native_mod.step_pow(x, y)
    dev = native_mod.step_pow(x, y, on_device=True)
    assert np.array_equal(host, dev), np.flatnonzero(host != dev)[:5]
