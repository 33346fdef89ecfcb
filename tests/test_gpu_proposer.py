"""The learned step-size proposer in the integration loop.

Gain: the network shipped with the package (trained by rednative.train_proposer on step
records written by the library, 1,121,692 records; see rednative/data/README) proposes
every first trial step on disks it never saw -- the 256-body planetary disk of seed 101
and the 16,384-body power-law disk of seed 202 -- while the embedded RKF7(8) error test
still accepts or rejects every attempt.  Against the reference controller
(dt <- 0.9 dt (tol/err)^(1/8)) it must need fewer force-evaluation attempts per simulated
year, reject no larger a share of attempts, and keep |dE/E| below 1e-9.  A controller with
a fixed, hand-tuned margin (0.98, no network) is printed beside it as an ablation.

Pipeline: records -> GPU training with the pinball loss -> the trained network predicts the
1 % quantile of the step change on its own records.
"""
import numpy as np
import pytest

from rednative import generate, native, proposer

pytestmark = pytest.mark.gpu

RKF78 = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01, use_graph=True)


@pytest.mark.parametrize("case", ["c1_seed101", "c2_seed202"])
def test_trained_proposer_beats_reference_controller(native_mod, case):
    if case == "c1_seed101":
        b = generate.to_internal(generate.disk_small(seed=101)); years = 20.0; need = 0.95
    else:
        b = generate.to_internal(generate.disk_powerlaw(n=16383, seed=202)); years = 6.0; need = 0.99
    print(f"case {case} years={years}", flush=True)
    net = proposer.load(proposer.DEFAULT_MODEL)
    zero = native_mod.Mlp(np.zeros(native_mod.mlp_param_count(), np.float32))
    ref = proposer.run(b, years, net=None, **RKF78)
    fixed = proposer.run(b, years, net=zero, safety=0.98, **RKF78)
    lrn = proposer.run(b, years, net=net, safety=proposer.SAFETY, **RKF78)
    for name, st in (("reference", ref), ("fixed 0.98", fixed), ("learned", lrn)):
        print(f"  {name:11s} attempts/yr {st.attempts / st.sim_years:8.1f}  rejected {st.rejected_fraction:.4f}  "
              f"sim-yr/hr {st.sim_yr_per_hr:9.0f}  dE/E {st.rel_energy:.2e}  dL/L {st.rel_angmom:.2e}", flush=True)
    assert lrn.attempts / lrn.sim_years <= need * ref.attempts / ref.sim_years
    assert lrn.rejected_fraction <= ref.rejected_fraction + 1e-3
    assert lrn.rel_energy < 1e-9 and ref.rel_energy < 1e-9


def test_training_pipeline_learns_the_quantile(native_mod):
    disks = [generate.to_internal(generate.disk_small(seed=3000 + k)) for k in range(2)]
    R = proposer.collect_many(disks, 2500, threads=2, explore=(0.8, 1.2), **RKF78)
    assert len(R) > 4000 and np.all(np.isfinite(R))
    net, losses = proposer.train(R, epochs=15, batch=1024, lr=2e-3, quantile=0.01)
    print("pinball loss", losses[[0, -1]], flush=True)
    assert losses[-1] < losses[0]
    pred = net.predict(R[:, :proposer.NF])
    below = float(np.mean(R[:, proposer.LABEL] < pred))
    print(f"share of labels below the predicted 1% quantile: {below:.4f}", flush=True)
    assert 0.001 <= below <= 0.05
