"""The one-work-item-per-sink FP64 sweep parks an inactive source at (1e30, 0, 0) with mass 0
instead of skipping it (the reference skips sources with id < 0, pp_disk.cu:242).  Checked
here on the CPU with IEEE float64 arithmetic in the sweep's operation order: the parked
source's term is a signed zero, and adding it leaves every sum bit-identical to the sum
without that source -- for sums that start at +0, pass through zero by cancellation, and for
sinks anywhere, including at the origin and far out."""
import numpy as np


def term(q, p):
    dx = q[0] - p[0]
    dy = q[1] - p[1]
    dz = q[2] - p[2]
    r2 = (dx * dx + dy * dy) + dz * dz
    d = np.sqrt(r2)
    sc = q[3] / (d * r2)
    return sc * dx, sc * dy, sc * dz


def sweep(sources, p, skip):
    ax = ay = az = np.float64(0.0)
    for k, q in enumerate(sources):
        if skip[k]:
            continue
        tx, ty, tz = term(q, p)
        ax += tx
        ay += ty
        az += tz
    return ax, ay, az


def bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


def test_parked_source_term_is_signed_zero():
    parked = np.array([1e30, 0.0, 0.0, 0.0])
    for p in ([0.0, 0.0, 0.0], [1.5, -2.0, 0.1], [-3e3, 7.0, -1e-3], [1e20, 0.0, 0.0]):
        t = term(parked, np.array(p))
        assert all(v == 0.0 for v in t)


def test_parked_sources_leave_sums_bit_identical():
    rng = np.random.default_rng(12)
    parked = np.array([1e30, 0.0, 0.0, 0.0])
    for case in range(200):
        n = int(rng.integers(1, 40))
        src = np.zeros((n, 4))
        src[:, :3] = rng.normal(0.0, 3.0, (n, 3))
        src[:, 3] = np.exp(rng.uniform(np.log(1e-12), np.log(1e-3), n))
        if case % 5 == 0:
            # a mirrored pair: the sums cancel to exactly +0 in between
            src[1:2] = src[0:1] * np.array([-1.0, -1.0, -1.0, 1.0])
        p = np.zeros(3) if case % 7 == 0 else rng.normal(0.0, 3.0, 3)
        inactive = rng.random(n) < 0.3
        packed = src.copy()
        packed[inactive] = parked
        ref = sweep(src, p, inactive)                    # the reference: inactive skipped
        got = sweep(packed, p, np.zeros(n, dtype=bool))  # parked sources summed
        assert np.array_equal(bits(ref), bits(got)), (case, ref, got)
