"""Contexts on several host threads run concurrently (one context per thread, as the C ABI
states): captured graphs, allocations and set-up copies of one context must not disturb
another's.  Four threads step four disks with HIP graphs on while a fifth thread keeps
creating and destroying contexts; every disk must end bit-identical to the same disk
stepped alone."""
import threading

import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu

KW = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01, use_graph=True)


def run(native_mod, b, steps):
    d = native_mod.Disk(b, **KW)
    for _ in range(steps):
        d.step()
    r = d.state()["r"]
    d.close()
    return r


def test_concurrent_contexts_match_sequential(native_mod):
    disks = [generate.to_internal(generate.disk_small(seed=5000 + k)) for k in range(4)]
    alone = [run(native_mod, b, 150) for b in disks]
    out = [None] * len(disks)
    errors = []
    stop = threading.Event()

    def worker(k):
        try:
            out[k] = run(native_mod, disks[k], 150)
        except Exception as e:          # reported below
            errors.append(e)

    def churn():
        b = generate.to_internal(generate.disk_small(n_pl=40, seed=6000))
        try:
            while not stop.is_set():
                d = native_mod.Disk(b, **KW)
                d.step()
                d.close()
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(len(disks))]
    ch = threading.Thread(target=churn)
    ch.start()
    for t in th:
        t.start()
    for t in th:
        t.join()
    stop.set()
    ch.join()
    assert not errors, errors
    for k in range(len(disks)):
        assert_bitwise(out[k], alone[k], f"disk {k}")
