"""The oracle's energy (a diagnostic the reference does not have) against an exactly
rounded sum of the same terms (math.fsum over every kinetic and pair term), on a disk
where the potential is dominated by the star and millions of tiny pair terms follow."""
import math

import numpy as np

from rednative import generate


def exact_energy(b):
    r, v, m = b.r, b.v, b.mass
    counts = b.counts
    nsi = sum(counts[:4]); nm = nsi + sum(counts[4:6])
    terms = list(0.5 * m[:nm] * np.sum(v[:nm] ** 2, axis=1))
    for i in range(nm):
        jend = nm if i < nsi else nsi
        if jend <= i + 1:
            continue
        d = np.linalg.norm(r[i + 1:jend] - r[i], axis=1)
        terms.extend(-(m[i] * m[i + 1:jend] / d))
    return math.fsum(terms)


def test_oracle_energy_is_the_exact_sum(oracle_mod):
    for b in (generate.to_internal(generate.disk_allsi(1501, seed=11)),
              generate.to_internal(generate.disk_powerlaw(n=3000, seed=17))):
        E, _ = oracle_mod.Oracle(b).energy()
        Ex = exact_energy(b)
        assert abs(E - Ex) <= 1e-14 * abs(Ex), (E, Ex)
