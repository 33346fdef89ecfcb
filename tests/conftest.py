"""Shared test setup: import paths, the `gpu` marker, small helpers.

CPU tests (-m "not gpu") check the oracle against the reference's golden vectors and
known answers, the host-side file formats and that the C ABI library loads and exports
every declared symbol.  GPU tests (-m gpu) drive the HIP path through the C ABI and
compare it with the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "red.cuda.closeencounter_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def native_mod():
    from rednative import native
    native.load_library()
    return native


def assert_bitwise(a, b, what=""):
    a = np.asarray(a); b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    if a.dtype.kind == "f":
        assert a.dtype == b.dtype, f"{what}: dtype {a.dtype} != {b.dtype}"
        u = np.dtype(f"u{a.dtype.itemsize}")
        ok = (a.view(u) == b.view(u)) | ((a == 0) & (b == 0))
    else:
        ok = a == b
    if not np.all(ok):
        idx = np.argwhere(~ok)[:5]
        raise AssertionError(f"{what}: {int((~ok).sum())} values differ, first at {idx.tolist()}: "
                             f"{a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")
