"""red_energy_noalloc: the energy diagnostic with no allocation, through the C ABI.

Its partials live in a static device array shared by every call on the GPU, and its sums
are formed on the device in red_energy's order, so its {E, Lx, Ly, Lz} must equal
red_energy's bit for bit: here on two calls in a row, the second on a smaller disk (so any
partial left over from the first call would show), then the first disk again, and a disk
whose partials exceed the array is refused.
"""
import ctypes

import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def _noalloc(native_mod, d):
    out = d.device_alloc(4 * 8)
    try:
        d.to_device(out, np.full(4, np.nan))
        rc = d.L.red_energy_noalloc(d.ctx, ctypes.c_void_p(out))
        if rc:
            raise native_mod.RedError(rc, d.L.red_last_error(d.ctx).decode())
        d.sync()
        res = np.zeros(4)
        d.to_host(res, out)
    finally:
        d.device_free(out)
    return res


def _expected(d):
    E, L = d.energy()
    return np.array([E, L[0], L[1], L[2]])


def test_energy_noalloc_matches_red_energy(native_mod):
    big = generate.to_internal(generate.disk_mixed(n_pl=20000, n_tp=500, seed=31))
    small = generate.to_internal(generate.disk_allsi(700, seed=32))
    print(f"case energy_noalloc n={big.n} then n={small.n}", flush=True)
    d1 = native_mod.Disk(big, integrator=2, dt_day=0.5)
    d2 = native_mod.Disk(small, integrator=2, dt_day=0.5)
    for d in (d1, d2):
        d.step()
    r1 = _noalloc(native_mod, d1)
    r2 = _noalloc(native_mod, d2)
    assert_bitwise(r1, _expected(d1), "first call")
    assert_bitwise(r2, _expected(d2), "second call, smaller disk")
    assert not np.array_equal(r1, r2)
    assert_bitwise(_noalloc(native_mod, d1), _expected(d1), "first disk again")
    d1.close()
    d2.close()


def test_energy_noalloc_refuses_oversized_disk(native_mod):
    # 30000 self-interacting bodies: 30000 x 8 potential partials > 2^17
    b = generate.to_internal(generate.disk_allsi(30000, seed=33))
    print(f"case energy_noalloc oversized n={b.n}", flush=True)
    d = native_mod.Disk(b)
    out = d.device_alloc(32)
    rc = d.L.red_energy_noalloc(d.ctx, ctypes.c_void_p(out))
    assert rc == 1 and b"static buffer" in d.L.red_last_error(d.ctx)
    d.device_free(out)
    d.close()
