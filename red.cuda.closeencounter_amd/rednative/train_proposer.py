"""Records step histories and trains the step-size proposer on the GPU.

    python -m rednative.train_proposer --out proposer.npz [--c1-disks 64 --c1-steps 15000
        --c3-steps 3000 --c2-disks 8 --c2-steps 12000 --epochs 30]

Records come from the library's own step log (red_history_*):
  * disks drawn as the 256-body planetary disk (1 star, 2 giants, 253 planetesimals, seeds
    1000..), integrated with RKF7(8) at tolerance 1e-12 -- the method and tolerance of the
    16384-body disk the proposer is meant for;
  * the 131072-body mixed disk (star, giants, 65536 planetesimals, 65533 test particles),
    mixed precision, Cash-Karp RK5(4) at tolerance 1e-10;
  * optionally, 4096-body disks of the 16384-body disk's family (protoplanets and
    planetesimals, power-law masses, seeds 2020..; never the evaluation seed 202), because
    only these contain planetesimal-protoplanet encounters.
Every trial step is drawn around the tolerance-matching step (exploration), so that the
records cover accepted and rejected attempts; the label does not depend on the trial step.
The network is trained by red_mlp_train (HIP forward, backward and Adam kernels) with the
pinball loss of level --quantile (default 0.01): it predicts a low quantile of the change of
the tolerance-matching step.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import generate, native, proposer

RKF78_TOL = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01, use_graph=True)
RK5_MIXED = dict(integrator=3, adaptive=True, tolerance=1e-10, dt_day=0.1, use_graph=True,
                 precision=native.PREC_MIXED)
EXPLORE = (0.8, 1.2)


def _log(msg):
    print(msg, flush=True)


def gather(c1_disks: int, c1_steps: int, c3_steps: int, c2_disks: int, c2_steps: int, threads: int = 8):
    parts = {}
    t0 = time.time()
    if c1_disks > 0:
        disks = [generate.to_internal(generate.disk_small(seed=1000 + k)) for k in range(c1_disks)]
        parts["c1"] = proposer.collect_many(
            disks, c1_steps, threads=threads, explore=EXPLORE,
            progress=lambda k, H: _log(f"  c1 disk {k}: {len(H)} records ({time.time() - t0:.0f}s)"), **RKF78_TOL)
    if c3_steps > 0:
        b = generate.to_internal(generate.disk_mixed())
        parts["c3"] = proposer.collect(b, c3_steps, explore=EXPLORE, **RK5_MIXED)
        _log(f"  c3: {len(parts['c3'])} records ({time.time() - t0:.0f}s)")
    if c2_disks > 0:
        disks = [generate.to_internal(generate.disk_powerlaw(n=4095, seed=2020 + k)) for k in range(c2_disks)]
        parts["c2family"] = proposer.collect_many(
            disks, c2_steps, threads=threads, explore=EXPLORE,
            progress=lambda k, H: _log(f"  c2-family disk {k}: {len(H)} records ({time.time() - t0:.0f}s)"),
            **RKF78_TOL)
    return parts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--c1-disks", type=int, default=64)
    ap.add_argument("--c1-steps", type=int, default=15000)
    ap.add_argument("--c3-steps", type=int, default=3000)
    ap.add_argument("--c2-disks", type=int, default=0)
    ap.add_argument("--c2-steps", type=int, default=12000)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--quantile", type=float, default=proposer.QUANTILE)
    ap.add_argument("--records", default=None, help="also save the records (.npz)")
    a = ap.parse_args(argv)
    _log("recording step histories")
    parts = gather(a.c1_disks, a.c1_steps, a.c3_steps, a.c2_disks, a.c2_steps, a.threads)
    R = np.concatenate(list(parts.values()))
    counts = {k: int(len(v)) for k, v in parts.items()}
    _log(f"records: {counts} total {len(R)}")
    if a.records:
        np.savez(a.records, **parts)
    t = time.time()
    net, losses = proposer.train(R, epochs=a.epochs, batch=a.batch, lr=a.lr, quantile=a.quantile)
    y = R[:, proposer.LABEL]
    _log(f"trained {a.epochs} epochs in {time.time() - t:.1f}s: loss {losses[0]:.3e} -> {losses[-1]:.3e} "
         f"(label variance {np.var(y):.3e})")
    w, _, nrm = net.params()
    np.savez(a.out, params=w, norm=nrm, meta=np.array(json.dumps(dict(records=counts, epochs=a.epochs,
                                                                       quantile=a.quantile,
                                                                       final_loss=float(losses[-1]),
                                                                       label_var=float(np.var(y))))))
    _log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
