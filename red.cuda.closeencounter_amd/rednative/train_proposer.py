"""Records step histories and trains the step-size proposer on the GPU.

    python -m rednative.train_proposer --out proposer.npz [--small-disks 64 --small-steps 15000
        --mixed-steps 3000 --powerlaw-disks 8 --powerlaw-steps 12000 --epochs 30]

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


def gather(small_disks: int, small_steps: int, mixed_steps: int, powerlaw_disks: int, powerlaw_steps: int, threads: int = 8):
    parts = {}
    t0 = time.time()
    if small_disks > 0:
        disks = [generate.to_internal(generate.disk_small(seed=1000 + k)) for k in range(small_disks)]
        parts["small"] = proposer.collect_many(
            disks, small_steps, threads=threads, explore=EXPLORE,
            progress=lambda k, H: _log(f"  small disk {k}: {len(H)} records ({time.time() - t0:.0f}s)"), **RKF78_TOL)
    if mixed_steps > 0:
        b = generate.to_internal(generate.disk_mixed())
        parts["mixed"] = proposer.collect(b, mixed_steps, explore=EXPLORE, **RK5_MIXED)
        _log(f"  mixed disk: {len(parts['mixed'])} records ({time.time() - t0:.0f}s)")
    if powerlaw_disks > 0:
        disks = [generate.to_internal(generate.disk_powerlaw(n=4095, seed=2020 + k)) for k in range(powerlaw_disks)]
        parts["powerlaw"] = proposer.collect_many(
            disks, powerlaw_steps, threads=threads, explore=EXPLORE,
            progress=lambda k, H: _log(f"  power-law disk {k}: {len(H)} records ({time.time() - t0:.0f}s)"),
            **RKF78_TOL)
    return parts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--small-disks", type=int, default=64)
    ap.add_argument("--small-steps", type=int, default=15000)
    ap.add_argument("--mixed-steps", type=int, default=3000)
    ap.add_argument("--powerlaw-disks", type=int, default=0)
    ap.add_argument("--powerlaw-steps", type=int, default=12000)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--quantile", type=float, default=proposer.QUANTILE)
    ap.add_argument("--records", default=None, help="also save the records (.npz)")
    a = ap.parse_args(argv)
    _log("recording step histories")
    parts = gather(a.small_disks, a.small_steps, a.mixed_steps, a.powerlaw_disks, a.powerlaw_steps, a.threads)
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
