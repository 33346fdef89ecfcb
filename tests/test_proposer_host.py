"""Host-side pieces of the step-size proposer that need no GPU: the shipped network file,
the standardisation floors and the initial parameter layout."""
import json
import os

import numpy as np

import mlp_oracle as mo
from conftest import PKG

DATA = os.path.join(PKG, "rednative", "data", "proposer_rkf78.npz")


def test_shipped_network_file():
    z = np.load(DATA, allow_pickle=False)
    p, nrm = z["params"], z["norm"]
    assert p.shape == (mo.NPARAM,) and p.dtype == np.float32 and np.all(np.isfinite(p))
    assert nrm.shape == (2 * mo.NF,) and np.all(np.isfinite(nrm)) and np.all(nrm[mo.NF:] > 0)
    meta = json.loads(str(z["meta"]))
    assert sum(meta["records"].values()) >= 1_000_000
    assert meta["quantile"] == 0.01


def test_normaliser_floors_and_init_layout():
    from rednative import proposer
    X = np.zeros((100, proposer.NF), np.float32)
    X[:, 0] = np.linspace(-2, -1, 100)          # spread 0.29 > floor
    mean, istd = proposer.normalizer(X)
    assert np.allclose(1 / istd[1:], proposer.STD_FLOOR[1:])      # constant features: floor
    assert abs(1 / istd[0] - X[:, 0].std()) < 1e-6
    p = proposer.init_params(3)
    assert p.shape == (mo.NPARAM,)
    a, b = mo.OFF["b2"][0], mo.NPARAM
    assert np.all(p[a:b] == 0.0)               # b2, W3, b3 start at zero: the output is 0
