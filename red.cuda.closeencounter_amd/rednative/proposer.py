"""Learned step-size proposer: record step histories, train the network on the GPU, propose.

Not in the reference, whose step control is the elementary controller
dt <- 0.9 dt (tol/err)^(1/order) (red.cuda/src/red.cuda/int_rungekutta8.cu:776-777).

Target.  For an attempt with step dt and error err, eta = dt (tol/err)^(1/order) is the
step that would have met the tolerance exactly.  The reference controller proposes
0.9 eta_n for step n+1, i.e. it assumes eta does not change and keeps a fixed margin
(the error then sits near 0.9^order = 0.43 of the tolerance for RKF7(8)).  The network
predicts a low quantile (1 %) of y = ln(eta_(n+1) / eta_n) from the 12 features of a step
record (recent eta ratios, the last error ratio and rejections, pair events, and state
features: closest approach in mutual Hill radii, the largest acceleration and the largest
non-Keplerian acceleration ratio, with their changes) -- trained with the pinball loss --
and the library proposes dt = safety * eta_n * exp(y): close to the tolerance where the
history and the state say the step will not shrink, smaller ahead of an encounter.  The
embedded error estimate still accepts or rejects every attempt and a rejected attempt is
retried with the reference controller's step, so the network only decides where a step
starts.  Records are written by the library itself (red_history_*), one per step.

The network (12 -> 64 -> 64 -> 1, GELU, BF16 weights, FP32 accumulation) is trained by
the library's own HIP forward/backward/Adam kernels (red_mlp_train).
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass
from typing import Iterable, List, Optional

import numpy as np

from . import native

NF = native.MLP_FEATURES
# network trained by train_proposer (RKF7(8) step records, see data/README)
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "proposer_rkf78.npz")
LABEL, OUTPUT, TRIED, FIRST_OK = 12, 13, 14, 15
DAY = 0.01720209895        # internal time units per day
YEAR_DAYS = 365.25


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights; the output layer starts at zero (y = 0: the reference step
    scaled by safety / 0.9)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / NF), (64, NF))
    W2 = rng.normal(0.0, np.sqrt(2.0 / 64), (64, 64))
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), np.zeros(64), [0.0]]).astype(np.float32)


# Smallest spread assumed per feature when standardising: a feature that barely varies in
# the training records (no rejections, no pair events) must not be blown up by 1/std when it
# does vary later.  Units follow the features: decades, natural-log ratios, counts.
STD_FLOOR = np.array([0.1, 0.1, 0.005, 0.005, 1.0, 1.0, 0.005, 0.1, 0.01, 0.1, 0.1, 0.01], np.float32)


def normalizer(X: np.ndarray):
    """Per-feature mean and 1/std over the training records, std floored by STD_FLOOR."""
    X = np.asarray(X, np.float64)
    mean = X.mean(axis=0)
    std = np.maximum(X.std(axis=0), STD_FLOOR)
    return mean.astype(np.float32), (1.0 / std).astype(np.float32)


def collect(bodies, n_steps: int, t_stop: Optional[float] = None, explore: Optional[tuple] = None,
            seed: int = 0, device: int = 0, **kw) -> np.ndarray:
    """Integrates one disk and returns its step records.

    explore=(lo, hi): the first trial step of every step is safety * eta_n with safety drawn
    uniformly from [lo, hi] (a zero network makes the proposal y = 0), so the records cover
    errors near and above the tolerance and rejected attempts.  The label ln(eta_(n+1)/eta_n)
    does not depend on the trial step while the error scales as dt^order.  Without explore
    the reference controller runs unchanged."""
    d = native.Disk(bodies, device=device, **kw)
    d.history_enable(n_steps + 8)
    zero = native.Mlp(np.zeros(native.mlp_param_count(), np.float32), device=device) if explore else None
    rng = np.random.default_rng(seed)
    k = 0
    try:
        while k < n_steps and (t_stop is None or d.t <= t_stop):
            if zero is not None:
                d.set_proposer(zero, safety=float(rng.uniform(*explore)), max_log_change=1.0)
            d.step()
            k += 1
    except native.RedError as e:
        # a disk that fails (e.g. the reference controller's infinite step after an error
        # estimate of 0) ends its history; the records up to the failure stay valid
        print(f"collect: disk stopped after {k} steps: {e}", flush=True)
    H = d.history()
    d.close()
    if zero is not None:
        zero.close()
    return H


def collect_many(disks: List, n_steps: int, threads: int = 8, progress=None, **kw) -> np.ndarray:
    """collect() over many disks, one context per host thread (the library releases the
    interpreter lock in every call, so the contexts' streams run side by side)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(k):
        H = collect(disks[k], n_steps, seed=k, **kw)
        if progress:
            progress(k, H)
        return H

    with ThreadPoolExecutor(max_workers=threads) as ex:
        out = list(ex.map(one, range(len(disks))))
    return np.concatenate(out) if out else np.zeros((0, native.HIST_FIELDS), np.float32)


# The network is trained for a low quantile of the step change, so that a proposal
# dt = SAFETY * eta_n * exp(y) is rejected rarely; SAFETY is a small fixed margin.
QUANTILE = 0.01
SAFETY = 0.995


def train(records: np.ndarray, epochs: int = 20, batch: int = 4096, lr: float = 2e-3, seed: int = 0,
          device: int = 0, net: Optional[native.Mlp] = None, quantile: Optional[float] = QUANTILE):
    """Trains a network on the records' features and labels with the pinball loss of level
    `quantile` (None: mean squared error); returns (net, epoch losses)."""
    R = np.asarray(records, np.float32)
    R = R[np.all(np.isfinite(R), axis=1)]
    rng = np.random.default_rng(seed)
    R = R[rng.permutation(len(R))]
    X, T = R[:, :NF], R[:, LABEL]
    if net is None:
        net = native.Mlp(init_params(seed), device=device)
        mean, istd = normalizer(X)
        net.set_normalizer(mean, istd)
    if quantile is None:
        net.set_loss(native.Mlp.LOSS_MSE)
    else:
        net.set_loss(native.Mlp.LOSS_PINBALL, quantile)
    losses = net.train(X, T, epochs=epochs, batch=batch, lr=lr)
    return net, losses


def save(path: str, net: native.Mlp) -> None:
    w, _, nrm = net.params()
    np.savez(path, params=w, norm=nrm)


def load(path: str, device: int = 0) -> native.Mlp:
    z = np.load(path, allow_pickle=False)
    return native.Mlp(z["params"], device=device, norm=z["norm"])


@dataclass
class RunStats:
    steps: int
    attempts: int
    rejected: int
    sim_years: float
    seconds: float
    rel_energy: float
    rel_angmom: float

    @property
    def rejected_fraction(self) -> float:
        return self.rejected / max(self.attempts, 1)

    @property
    def sim_yr_per_hr(self) -> float:
        return self.sim_years / max(self.seconds, 1e-12) * 3600.0


def run(bodies, t_years: float, net: Optional[native.Mlp] = None, safety: float = SAFETY,
        max_log_change: float = 1.0, max_steps: int = 10 ** 9, warm_steps: int = 0, **kw) -> RunStats:
    """Integrates for t_years (from the disk's start) with or without the proposer and
    reports steps, attempts, rejections, wall time and the energy / angular-momentum drift."""
    d = native.Disk(bodies, **kw)
    for _ in range(warm_steps):
        d.step()
    if net is not None:
        d.set_proposer(net, safety=safety, max_log_change=max_log_change)
    E0, L0 = d.energy()
    c0 = d.counters()
    t0 = d.t
    t_stop = t0 + t_years * YEAR_DAYS * DAY
    d.sync()
    w0 = time.perf_counter()
    k = 0
    while d.t <= t_stop and k < max_steps:
        d.step()
        k += 1
    d.sync()
    sec = time.perf_counter() - w0
    E1, L1 = d.energy()
    c1 = d.counters()
    sim = (d.t - t0) / DAY / YEAR_DAYS
    d.close()
    return RunStats(steps=c1["passed"] - c0["passed"], attempts=c1["tried"] - c0["tried"],
                    rejected=c1["failed"] - c0["failed"], sim_years=sim, seconds=sec,
                    rel_energy=abs((E1 - E0) / E0),
                    rel_angmom=float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0)))
