"""Ensembles of independent disks (red_ensemble_*): every member must follow exactly the
steps it takes alone.  Each disk of a small ensemble of uneven sizes is compared bit for
bit with the single-disk path (red_step in the reference loop `while t <= t_stop`) and,
for two disks, with the CPU oracle; the autotuned launch shape and graph replay must not
change any result."""
import numpy as np
import pytest

from conftest import assert_bitwise
from rednative import generate

pytestmark = pytest.mark.gpu


def small_ensemble(n=10, seed=404):
    return [generate.to_internal(d) for d in generate.ensemble(n_disks=n, seed=seed, n_range=(16, 700))]


def run_alone(native_mod, b, t_stop, **kw):
    d = native_mod.Disk(b, **kw)
    while d.t <= t_stop:
        d.step()
    return d


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_members_match_single_disk(native_mod, oracle_mod, graph):
    disks = small_ensemble()
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    print(f"case ensemble n_disks={len(disks)} sizes={[d.n for d in disks]} graph={graph}", flush=True)
    E = native_mod.Ensemble(disks, use_graph=graph, **kw)
    t_stop = 4.0 * 0.01720209895          # four days in internal units
    info = E.advance(t_stop)
    assert info["active_disks"] == 0 and info["supersteps"] > 0
    for k, b in enumerate(disks):
        alone = run_alone(native_mod, b, t_stop, **kw)
        r, v = E.state(k)
        s = alone.state()
        assert_bitwise(r, s["r"], f"disk {k} r")
        assert_bitwise(v, s["v"], f"disk {k} v")
        di = E.disk_info(k)
        c = alone.counters()
        assert (di["tried"], di["failed"], di["passed"]) == (c["tried"], c["failed"], c["passed"])
        assert di["t"] == alone.t and di["dt_try"] == alone.dt_try
    for k in (0, 3):
        o = oracle_mod.Oracle(disks[k], **kw)
        while o.t <= t_stop:
            o.step()
        r, _ = E.state(k)
        assert_bitwise(r, o.state()["r"], f"disk {k} vs oracle")


def test_ensemble_autotune_keeps_results(native_mod):
    disks = small_ensemble(n=8, seed=405)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    E1 = native_mod.Ensemble(disks, **kw)
    E2 = native_mod.Ensemble(disks, **kw)
    block, jtile, table = E2.autotune(reps=3)
    print("autotune", block, jtile, table.tolist())
    assert block in (64, 128, 256) and jtile in (64, 128, 256)
    assert np.all(table > 0)
    t_stop = 2.0 * 0.01720209895
    E1.advance(t_stop); E2.advance(t_stop)
    for k in range(len(disks)):
        assert_bitwise(E1.state(k)[0], E2.state(k)[0], f"disk {k}")


@pytest.mark.parametrize("integ,adaptive,tol,dt", [(0, False, 0.0, 0.1), (1, False, 0.0, 0.2), (2, True, 1e-10, 0.5),
                                                    (3, True, 1e-10, 0.5), (4, False, 0.0, 0.5), (3, False, 0.0, 0.5)])
def test_ensemble_every_scheme_matches_single_disk(native_mod, integ, adaptive, tol, dt):
    """Every integrator and step control: the force launches form the stage states, y_(n+1)
    and the error stages in their epilogues (retried attempts from the stored f0), and every
    member equals its disk stepped alone, bit for bit."""
    disks = small_ensemble(n=5, seed=408)
    kw = dict(integrator=integ, adaptive=adaptive, tolerance=tol, dt_day=dt)
    print(f"case ensemble scheme {kw}", flush=True)
    E = native_mod.Ensemble(disks, **kw)
    t_stop = 3.0 * 0.01720209895
    info = E.advance(t_stop)
    assert info["active_disks"] == 0
    for k, b in enumerate(disks):
        alone = run_alone(native_mod, b, t_stop, **kw)
        r, v = E.state(k)
        s = alone.state()
        assert_bitwise(r, s["r"], f"disk {k} r")
        assert_bitwise(v, s["v"], f"disk {k} v")
        c = alone.counters()
        di = E.disk_info(k)
        assert (di["tried"], di["failed"], di["passed"]) == (c["tried"], c["failed"], c["passed"])


def test_ensemble_autotune_between_advances(native_mod):
    """Autotuning after the first advance (a graph already captured for the old launch shape)
    leaves the results as they are."""
    disks = small_ensemble(n=6, seed=409)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    E1 = native_mod.Ensemble(disks, block=64, **kw)
    E2 = native_mod.Ensemble(disks, block=64, **kw)
    t1, t2 = 1.0 * 0.01720209895, 2.0 * 0.01720209895
    E1.advance(t1); E2.advance(t1)
    E2.autotune(reps=2)
    E1.advance(t2); E2.advance(t2)
    for k in range(len(disks)):
        assert_bitwise(E1.state(k)[0], E2.state(k)[0], f"disk {k}")


def test_ensemble_fixed_step_rk4(native_mod):
    disks = small_ensemble(n=6, seed=406)
    kw = dict(integrator=2, adaptive=False, dt_day=0.5)
    E = native_mod.Ensemble(disks, **kw)
    t_stop = 3.0 * 0.01720209895
    E.advance(t_stop)
    for k, b in enumerate(disks):
        alone = run_alone(native_mod, b, t_stop, **kw)
        assert_bitwise(E.state(k)[0], alone.state()["r"], f"disk {k}")


def test_ensemble_advance_in_parts(native_mod):
    """A run to t1 and then to t2 steps every member as one run to t2 does (the loop
    condition only decides whether a step is taken), also when the first call is cut short
    by max_attempts; every member still equals the single-disk path."""
    disks = small_ensemble(n=6, seed=407)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001)
    print("case ensemble advance in parts", flush=True)
    t1, t2 = 1.5 * 0.01720209895, 4.0 * 0.01720209895
    whole = native_mod.Ensemble(disks, **kw)
    whole.advance(t2)
    parts = native_mod.Ensemble(disks, **kw)
    info = parts.advance(t1)
    assert info["supersteps"] > 0 and info["active_disks"] == 0
    cut = parts.advance(t2, max_attempts=1)
    assert cut["supersteps"] == 1 and cut["active_disks"] > 0
    rest = parts.advance(t2)
    assert rest["supersteps"] > 0 and rest["active_disks"] == 0
    for k, b in enumerate(disks):
        assert_bitwise(parts.state(k)[0], whole.state(k)[0], f"disk {k} r")
        assert_bitwise(parts.state(k)[1], whole.state(k)[1], f"disk {k} v")
        assert parts.disk_info(k) == whole.disk_info(k)
        alone = run_alone(native_mod, b, t2, **kw)
        assert_bitwise(parts.state(k)[0], alone.state()["r"], f"disk {k} alone")


def _alone_main_loop(native_mod, b, t_stop, loop_kw, **kw):
    """One disk through the reference's driver loop (ejection check, step, collision or
    close-encounter pass, rebuild check; main.cu:176-255) up to t_stop: the state and the
    handled event records."""
    from rednative import mainloop
    d = native_mod.Disk(b, **kw)
    events = []
    prm = dict(stop_time=t_stop, output_interval=1e300, inner_steps=True, **loop_kw)
    mainloop.run_main_loop(d, prm, lambda e: None, lambda evs: events.extend(evs))
    return d, events


def _member_equal(E, k, d, events, what):
    from test_gpu_events import events_equal
    r, v = E.state(k)
    s = d.state()
    assert E.counts(k) == d.counts(), (what, k)
    assert_bitwise(r, s["r"], f"{what} disk {k} r")
    assert_bitwise(v, s["v"], f"{what} disk {k} v")
    bo = E.bodies(k)
    assert_bitwise(bo["p"], s["p"], f"{what} disk {k} p")
    assert np.array_equal(bo["id"], s["id"]), (what, k)
    assert E.counters(k) == d.counters(), (what, k, E.counters(k), d.counters())
    di = E.disk_info(k)
    assert di["t"] == d.t and di["dt_try"] == d.dt_try, (what, k)
    events_equal(events, E.events(k), f"{what} disk {k} events")


EVENT_CASES = {
    # integrator, step control, the pair and scan rules
    "rk4_merge": dict(integrator=2, adaptive=False, dt_day=0.5, hit_centrum=0.3, ejection=6.0,
                      radii_enhance_factor=2500.0),
    "rkf78_merge": dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5, hit_centrum=0.3, ejection=6.0,
                        radii_enhance_factor=2500.0),
    "rkf78_padded_hill": dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.5, ejection=6.0,
                              radii_enhance_factor=1.0, encounter_hill_factor=3.0, padded_storage=True),
    "ck_hill": dict(integrator=3, adaptive=True, tolerance=1e-10, dt_day=0.5, radii_enhance_factor=1.0,
                    encounter_hill_factor=10.0),
    "rk4_ce_mode": dict(integrator=2, adaptive=False, dt_day=0.5, close_encounter=True, radii_enhance_factor=500.0),
}


@pytest.mark.parametrize("case", sorted(EVENT_CASES))
def test_ensemble_events_match_single_disk(native_mod, case):
    """Members with collisions merged, close encounters reported (Hill test and close-encounter
    mode), ejections, hit-centrum merges and compactions: every member's state, bodies,
    counters and event records equal those of the same disk run alone through the driver
    loop, bit for bit."""
    from test_gpu_events import planted_disk
    kw = EVENT_CASES[case]
    gen = small_ensemble(n=3, seed=407)
    for b in gen:
        b.radius[0] = 1e-7   # as the planted disks' star: 2500 stellar radii must not swallow the disk
    disks = [planted_disk(seed=41), planted_disk(seed=44)] + gen + [planted_disk(seed=45)]
    print(f"case ensemble events {case} sizes={[d.n for d in disks]}", flush=True)
    t_stop = 30.0 * 0.01720209895
    loop_kw = {k: kw[k] for k in ("radii_enhance_factor", "encounter_hill_factor", "close_encounter") if k in kw}
    E = native_mod.Ensemble(disks, event_capacity=200000, **kw)
    info = E.advance(t_stop)
    assert info["active_disks"] == 0
    total = dict(n_collision=0, n_close_encounter=0, n_ejection=0, n_hit_centrum=0)
    for k, b in enumerate(disks):
        d, events = _alone_main_loop(native_mod, b, t_stop, loop_kw, event_capacity=200000, **kw)
        _member_equal(E, k, d, events, case)
        for key in total:
            total[key] += d.counters()[key]
        d.close()
    print("events", total, flush=True)
    if case.endswith("merge"):
        assert total["n_collision"] > 0 and total["n_ejection"] > 0 and total["n_hit_centrum"] > 0
    else:
        assert total["n_close_encounter"] > 0


def test_ensemble_config4_slice(native_mod, oracle_mod):
    """A 64-disk slice of the 512-disk ensemble workload (seed 404: the 32 largest disks, up to
    4096 bodies, and every 15th of the rest), each disk a star, the two giants and
    planetesimals out to its own radius, with the 256-body disk's rules: collisions at
    1 x (R_i + R_j), close encounters at 3 mutual Hill radii, ejection at 100 AU; RKF7(8) at
    tolerance 1e-11 for two days.  Every member equals its disk run alone through the driver
    loop bit for bit; the smallest and the largest equal the oracle."""
    all_disks = generate.ensemble(n_disks=512, seed=404)
    order = sorted(range(512), key=lambda k: -all_disks[k].n)
    pick = sorted(order[:32] + order[32::15][:32])
    disks = [generate.to_internal(all_disks[k]) for k in pick]
    assert len(disks) == 64 and max(d.n for d in disks) >= 4000 and all(d.counts[1] == 2 for d in disks)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001, ejection=100.0, radii_enhance_factor=1.0,
              encounter_hill_factor=3.0)
    loop_kw = dict(radii_enhance_factor=1.0, encounter_hill_factor=3.0)
    t_stop = 2.0 * 0.01720209895
    print(f"case ensemble c4 slice n_disks={len(disks)} bodies={sum(d.n for d in disks)}", flush=True)
    E = native_mod.Ensemble(disks, **kw)
    info = E.advance(t_stop)
    assert info["active_disks"] == 0
    ce = 0
    for k, b in enumerate(disks):
        d, events = _alone_main_loop(native_mod, b, t_stop, loop_kw, **kw)
        _member_equal(E, k, d, events, "c4")
        ce += d.counters()["n_close_encounter"]
        d.close()
    print("close encounters", ce, flush=True)
    from rednative import mainloop
    for k in (min(range(64), key=lambda q: disks[q].n), max(range(64), key=lambda q: disks[q].n)):
        o = oracle_mod.Oracle(disks[k], **kw)
        prm = dict(stop_time=t_stop, output_interval=1e300, inner_steps=True, **loop_kw)
        mainloop.run_main_loop(o, prm, lambda e: None)
        r, v = E.state(k)
        so = o.state()
        assert_bitwise(r, so["r"], f"disk {k} vs oracle r")
        assert_bitwise(v, so["v"], f"disk {k} vs oracle v")


@pytest.mark.parametrize("case", ["rkf78_merge", "ck_hill"])
def test_ensemble_batch_sizes_agree(native_mod, case):
    """The device runs each disk's controller for a batch of supersteps and holds a disk for
    the host where its step needs one (merges, ejection scans, compactions that move bodies):
    batches of 1, 3 and 16 supersteps give every member the same state, bodies, counters and
    event records."""
    from test_gpu_events import planted_disk
    kw = EVENT_CASES[case]
    gen = small_ensemble(n=3, seed=411)
    for b in gen:
        b.radius[0] = 1e-7
    disks = [planted_disk(seed=41), planted_disk(seed=44)] + gen
    t_stop = 20.0 * 0.01720209895
    runs = []
    for batch in (1, 3, 16):
        print(f"case ensemble batch {case} batch={batch}", flush=True)
        E = native_mod.Ensemble(disks, event_capacity=200000, batch=batch, **kw)
        info = E.advance(t_stop)
        assert info["active_disks"] == 0
        runs.append([(E.state(k), E.bodies(k), E.counters(k), E.disk_info(k), E.events(k)) for k in range(len(disks))])
        E.close()
    from test_gpu_events import events_equal
    for other in runs[1:]:
        for k in range(len(disks)):
            (r0, v0), b0, c0, i0, e0 = runs[0][k]
            (r1, v1), b1, c1, i1, e1 = other[k]
            assert_bitwise(r0, r1, f"disk {k} r")
            assert_bitwise(v0, v1, f"disk {k} v")
            assert np.array_equal(b0["id"], b1["id"]) and c0 == c1 and i0 == i1, k
            events_equal(e0, e1, f"disk {k} events")


def test_ensemble_small_record_capacity(native_mod):
    """With a record capacity small enough that a batch's kept records pass half of it, the
    batch's later supersteps idle and the next batch resumes them: members still equal those
    of a run with one superstep per batch and ample capacity (and the batch size is checked)."""
    from test_gpu_events import planted_disk
    kw = EVENT_CASES["ck_hill"]
    disks = [planted_disk(seed=41), planted_disk(seed=44)] + small_ensemble(n=2, seed=412)
    for b in disks[2:]:
        b.radius[0] = 1e-7
    # 25-95 kept records per superstep (a few hundred raw ones per attempt): a batch of 16
    # supersteps passes half of a 400-record capacity after a few of them
    t_stop = 20.0 * 0.01720209895
    print("case ensemble small record capacity", flush=True)
    ref = native_mod.Ensemble(disks, event_capacity=200000, batch=1, **kw)
    ref.advance(t_stop)
    small = native_mod.Ensemble(disks, event_capacity=400, batch=16, **kw)
    for bad in (0, 65):
        with pytest.raises(native_mod.RedError):
            small.set_batch(bad)
    info = small.advance(t_stop)
    assert info["active_disks"] == 0
    print("batches", info["batches"], "full", info["full_batches"], "supersteps", info["supersteps"], flush=True)
    assert info["full_batches"] > 0
    from test_gpu_events import events_equal
    n_records = 0
    for k in range(len(disks)):
        assert_bitwise(small.state(k)[0], ref.state(k)[0], f"disk {k} r")
        assert small.counters(k) == ref.counters(k) and small.disk_info(k) == ref.disk_info(k), k
        e_small, e_ref = small.events(k), ref.events(k)
        events_equal(e_small, e_ref, f"disk {k} events")
        n_records += len(e_ref)
    print("records", n_records, flush=True)
