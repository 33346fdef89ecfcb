"""Energy and angular momentum of the GPU path (red_energy) against an exactly rounded sum
of the same terms and against the CPU oracle on the same state.  E = sum 1/2 m v^2 -
sum_{interacting i<j} m_i m_j / r_ij (NSI-NSI pairs excluded, as in the force model),
L = sum m r x v.  The GPU adds the pair terms in chunk partials, the oracle with
compensated sums: both agree with the exact sum to 1e-13 relative."""
import numpy as np
import pytest

from rednative import generate
from test_oracle_energy import exact_energy

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["typed", "allsi", "mixed", "powerlaw"])
def test_energy_matches_oracle(oracle_mod, native_mod, name):
    b = {"typed": lambda: generate.disk_small(253, seed=101),
         "allsi": lambda: generate.disk_allsi(3001, seed=11),
         "mixed": lambda: generate.disk_mixed(n_pl=5000, n_tp=3000, seed=5),
         "powerlaw": lambda: generate.disk_powerlaw(n=9000, seed=17)}[name]()
    b = generate.to_internal(b)
    print(f"case energy {name} n={b.n}", flush=True)
    o = oracle_mod.Oracle(b)
    d = native_mod.Disk(b)
    Eo, Lo = o.energy()
    Eg, Lg = d.energy()
    Ex = exact_energy(b)
    print(f"E exact {Ex!r} oracle {Eo!r} gpu {Eg!r}", flush=True)
    assert abs(Eg - Ex) <= 1e-13 * abs(Ex)
    assert abs(Eo - Ex) <= 1e-13 * abs(Ex)
    np.testing.assert_allclose(Lg, Lo, rtol=1e-12, atol=1e-14 * np.abs(Lo).max())
