#!/usr/bin/env python3
"""Benchmark of the integration hot path on one MI355X.

Workload (one step = one adaptive Cash-Karp RK5(4) step, tolerance 1e-10, with its
rejected attempts): 1 star + 2 giant planets + 65536 planetesimals (m log-uniform in
[1e-12, 1e-9] M_sun, non-self-interacting as the reference types planetesimals) +
65533 massless test particles, a uniform in [0.4, 4] AU; N = 131072.  Collisions at
1 x (R_i + R_j) with merging, ejection at 100 AU.  Synthetic, seeded (303).

Reported: pair-interactions/s of the mixed-precision path (FP64 differences, FP32
r^-3, compensated FP64 sums), simulated years per hour, the relative energy change
over the timed steps, the same figures for the plain FP64 path on the same state,
the roofline position of the dominant kernel (the test-particle force sweep) from
HIP events around its launches, beside the ALU ceiling of its pair sequence measured on the
same GPU (red_valu_rates), and the CPU oracle's pair rate on the host.  Further
sections (each also alone with --only): the learned step-size proposer on the 16384-body
disk, the 512-disk ensemble, the 256-body disk's 10^4 fixed RK4 steps through the fused
device loop and step by step (small_disk), and, only with --only drift, the energy drift
over a long run.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "red.cuda.closeencounter_amd"))

GAUSS = 1.720209895e-2
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (packed), vendor figure
PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64
FLOP_PER_PAIR = 20           # 3 sub, 6 r^2, 1 rsqrt, 2 r^-3, 1 mass, 6 accumulate (+1 select)


def pair_counts(counts):
    ns, ngp, nrp, npp, nspl, npl, ntp = counts
    n_si = ns + ngp + nrp + npp
    n_nsi = nspl + npl
    n_m = n_si + n_nsi
    si = n_si * n_m - n_si          # self pairs excluded
    nsi = n_nsi * n_si
    ni = ntp * n_m
    return si + nsi + ni, dict(si=si, nsi=nsi, ni=ni)


def _driver_steps(d, steps):
    """`steps` iterations of the driver loop (ejection check, step, collision check)."""
    tries = 0
    sim_t = 0.0
    for _ in range(steps):
        sim_t += d.step()
        tries += d.last_step.tries
        d.check_for_ejection_hit_centrum()
        d.check_for_collision()
        d.clear_event_counter()
    return sim_t, tries


def run_path(native, bodies, precision, steps, warmup, kahan=False, tol=1e-10, use_graph=True):
    """The timed steps replay each attempt from its captured HIP graph with no timing events
    (events around every launch cost 1-2 % of a step, and graphs cannot carry them); the
    per-launch-class durations come from a second pass over the same number of steps, on a
    disk stepped without graphs with HIP events on the context streams around every launch."""
    kw = dict(integrator=3, adaptive=True, tolerance=tol, dt_day=1.0, precision=precision,
              radii_enhance_factor=1.0, ejection=100.0, kahan_pairs=kahan)
    d = native.Disk(bodies, use_graph=use_graph, **kw)
    _driver_steps(d, warmup)
    E0, _ = d.energy()
    d.sync()
    t0 = time.perf_counter()
    sim_t, tries = _driver_steps(d, steps)
    d.sync()
    t1 = time.perf_counter()
    E1, _ = d.energy()
    res = dict(seconds=t1 - t0, tries=tries, sim_t=sim_t, E0=E0, E1=E1, counts=d.counts(), t_end=d.t)
    d.close()
    p = native.Disk(bodies, use_graph=False, **kw)
    _driver_steps(p, 1)
    p.sync()
    p.set_profiling(True)
    p.profile(reset=True)
    _driver_steps(p, steps)
    p.sync()
    res["prof"] = p.profile(reset=True)
    p.close()
    return res


def reference_workload_section(native, reps=20):
    """The workload of the reference's only published timings (nvprof traces of
    calc_grav_accel on a GeForce GT 630, red.cuda/nvprofile.results.xlsx, from the harness
    red.code_snippet.cuda.test/main.cu:1049): 1 star + 5119 protoplanets + 5120 test
    particles, FP64, 5120 sinks x 5120 sources per sweep.  Here: the FP64 path with the
    reference's sequential sums, the SI and NI sweeps timed with HIP events."""
    import math
    import numpy as np
    from rednative import generate
    rng = generate._rng(1049)
    B = generate._Builder()
    generate._add_star(B)
    rho = 2.0 * generate.GRAM_PER_CM3_TO_SOLAR_PER_AU3
    m = np.exp(rng.uniform(math.log(1e-12), math.log(1e-9), 5119))
    generate._add_bodies(B, 3, "pp", m, rho, generate._elements(rng, 5119, 0.4, 4.0), 1.0)
    generate._add_bodies(B, 6, "tp", np.zeros(5120), 0.0, generate._elements(rng, 5120, 0.4, 4.0), 1.0)
    b = generate.to_internal(B.build())
    d = native.Disk(b, radii_enhance_factor=1.0)
    n = b.n
    rr = np.zeros((n, 4)); vv = np.zeros((n, 4))
    rr[:, :3] = b.r; rr[:, 3] = b.mass; vv[:, :3] = b.v
    bufs = [d.device_alloc(n * 32) for _ in range(3)]
    d.to_device(bufs[0], rr)
    d.to_device(bufs[1], vv)
    d.accel_device(0.0, 0, bufs[0], bufs[1], bufs[2])
    d.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        d.accel_device(0.0, 0, bufs[0], bufs[1], bufs[2])
    d.sync()
    eval_us = (time.perf_counter() - t0) / reps * 1e6
    d.set_profiling(True)
    d.profile(reset=True)
    for _ in range(reps):
        d.accel_device(0.0, 0, bufs[0], bufs[1], bufs[2])
    d.sync()
    prof = d.profile(reset=True)
    for p in bufs:
        d.device_free(p)
    d.close()
    si_us = prof["si"][0] / max(prof["si"][1], 1) * 1e3
    ni_us = prof["ni"][0] / max(prof["ni"][1], 1) * 1e3
    pairs = 5120 * 5120
    return {"workload": "1 star + 5119 protoplanets + 5120 test particles, FP64 sequential sums, collision test on",
            "eval_us": eval_us, "reference_published_eval_ms": 135.45 + 166.45,
            "speedup_vs_published_eval": (135.45 + 166.45) * 1e3 / eval_us,
            "si_sweep_us": si_us, "ni_sweep_us": ni_us,
            "pairs_per_s_si": (pairs - 5120) / (si_us * 1e-6), "pairs_per_s_ni": pairs / (ni_us * 1e-6),
            "reference_published_ms": {"si": [134.9, 136.0], "ni": [165.7, 167.2], "padded": [72.6, 73.2],
                                       "hardware": "GeForce GT 630"},
            "speedup_vs_published_si": 135.45e3 / si_us, "speedup_vs_published_ni": 166.45e3 / ni_us}


def accuracy_section(native, bodies, steps=10, dt_day=1.0):
    """Mixed against FP64 on the same disk: the relative acceleration error per body of one
    force evaluation, and the position / velocity difference after `steps` fixed Cash-Karp
    steps of dt_day from the same state (fixed steps, so both paths take the same times)."""
    import numpy as np
    kw = dict(radii_enhance_factor=1.0, ejection=100.0)
    d64 = native.Disk(bodies, integrator=3, adaptive=False, dt_day=dt_day, **kw)
    dmx = native.Disk(bodies, integrator=3, adaptive=False, dt_day=dt_day, precision=native.PREC_MIXED, **kw)
    a64 = d64.accel(0.0, bodies.r, bodies.v)
    amx = dmx.accel(0.0, bodies.r, bodies.v)
    rel = np.linalg.norm(amx - a64, axis=1) / np.maximum(np.linalg.norm(a64, axis=1), 1e-300)
    for _ in range(steps):
        d64.step()
        dmx.step()
    s64, smx = d64.state(), dmx.state()
    dr = np.linalg.norm(smx["r"] - s64["r"], axis=1) / np.maximum(np.linalg.norm(s64["r"], axis=1), 1e-300)
    dv = np.linalg.norm(smx["v"] - s64["v"], axis=1) / np.maximum(np.linalg.norm(s64["v"], axis=1), 1e-300)
    d64.close(); dmx.close()
    return {"accel_rel_err_max": float(rel.max()), "accel_rel_err_median": float(np.median(rel)),
            "steps": steps, "dt_day": dt_day,
            "pos_rel_err_max": float(dr.max()), "pos_rel_err_median": float(np.median(dr)),
            "vel_rel_err_max": float(dv.max()), "vel_rel_err_median": float(np.median(dv))}


def host_info():
    """The host the CPU legs run on: lscpu's model and topology, the CPUs this process may
    use, and the OpenMP thread count the job is given."""
    import subprocess
    info = {"os_cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads():
    """Threads of the CPU leg: every CPU this process may use, capped by OMP_NUM_THREADS when
    the job sets it (the share of the host that comes with one GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(bodies_full, target_s=4.0, sim_steps=3):
    """The CPU oracle (FP64, the reference's operation order) on the same disk, timed on the
    host: pair rate of force evaluations on all the job's threads and on one thread (each on
    a sample of the test particles sized to about target_s), and simulated years per hour
    from sim_steps full adaptive Cash-Karp steps on all threads after one untimed step (the
    main loop's work per step: ejection check, step, collision check), extrapolated to one
    thread by the pair rates."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from rednative import files

    threads = cpu_threads()
    m = sum(bodies_full.counts[:6])

    def sample(ntp):
        b = bodies_full
        keep = np.r_[0:m, m:m + ntp]
        c = list(b.counts[:6]) + [ntp]
        return files.BodySet(c, b.ids[keep], [b.names[k] for k in keep], b.types[keep], b.epoch[keep],
                             b.mass[keep], b.radius[keep], b.density[keep], b.cd[keep], b.mig_type[keep],
                             b.mig_stop_at[keep], b.r[keep], b.v[keep])

    def rate(ntp, nthr):
        s = sample(ntp)
        o = pyoracle.Oracle(s, n_threads=nthr, radii_enhance_factor=1.0)
        t0 = time.perf_counter()
        o.accel(0.0, s.r, s.v)
        dt = time.perf_counter() - t0
        return pair_counts(s.counts)[0], dt

    def measured(nthr):
        p, dt = rate(256, nthr)
        ntp = int(max(256, min(bodies_full.counts[6], 256 * target_s / max(dt, 1e-6))))
        p, dt = rate(ntp, nthr)
        return p / dt, f"{ntp} of its {bodies_full.counts[6]} test particles: {p} pairs in {dt:.2f} s"

    rate_all, sample_all = measured(threads)
    rate_one, sample_one = measured(1)
    # full steps of the whole disk on all threads
    o = pyoracle.Oracle(bodies_full, integrator=3, adaptive=True, tolerance=1e-10, dt_day=1.0,
                        radii_enhance_factor=1.0, ejection=100.0, n_threads=threads)
    def one_step():
        o.check_for_ejection_hit_centrum()
        o.step()
        o.check_for_collision()
        o.clear_event_counter()

    one_step()                    # the first step's trial and rejections, untimed as on the GPU
    t_start, tried0 = o.t, o.counters()["tried"]
    t0 = time.perf_counter()
    for _ in range(sim_steps):
        one_step()
    wall = time.perf_counter() - t0
    yrs = (o.t - t_start) / GAUSS / 365.25
    attempts = o.counters()["tried"] - tried0
    sim_all = yrs / (wall / 3600.0)
    return dict(value=rate_all, unit="pair-interactions/s", cores=threads, kind="port",
                sample=f"one FP64 force evaluation (oracle, OpenMP over sinks, {threads} threads) of the same disk "
                       f"with {sample_all}",
                one_thread={"pairs_per_s": rate_one, "sample": f"the same on 1 thread with {sample_one}"},
                sim_yr_per_hr_extrapolated={
                    "threads": threads, "value": sim_all, "one_thread_value": sim_all * rate_one / rate_all,
                    "basis": f"{sim_steps} full adaptive Cash-Karp RK5(4) steps (tol 1e-10, {attempts} attempts) "
                             f"of the whole disk after one untimed step: {yrs * 365.25:.3f} simulated days in "
                             f"{wall:.1f} s on {threads} threads; one thread scaled by the pair rates"},
                host=host_info())


def proposer_section(native, years, use_graph=True):
    """Learned step-size proposer on the 16384-body disk (RKF7(8), tol 1e-12, FP64):
    reference controller against the shipped network, same span of simulated time."""
    import numpy as np
    from rednative import generate, proposer
    b = generate.to_internal(generate.disk_powerlaw(n=16383, seed=202))
    kw = dict(integrator=4, adaptive=True, tolerance=1e-12, dt_day=0.01, use_graph=use_graph)
    net = proposer.load(proposer.DEFAULT_MODEL)
    out = {"workload": "1 star + 16383 protoplanets/planetesimals (m >= 1e-10 M_sun self-interacting), "
                       "power-law masses, RKF7(8) tol 1e-12, FP64 sequential sums",
           "sim_years": years}
    for name, nn, s in (("reference_controller", None, 0.9), ("learned", net, proposer.SAFETY)):
        st = proposer.run(b, years, net=nn, safety=s, **kw)
        out[name] = {"attempts_per_sim_yr": st.attempts / st.sim_years, "rejected_fraction": st.rejected_fraction,
                     "sim_yr_per_hr": st.sim_yr_per_hr, "dE_over_E": st.rel_energy, "dL_over_L": st.rel_angmom,
                     "hours_for_1000_yr": 1000.0 / st.sim_yr_per_hr}
    ref, lrn = out["reference_controller"], out["learned"]
    out["attempts_ratio"] = lrn["attempts_per_sim_yr"] / ref["attempts_per_sim_yr"]
    out["sim_yr_per_hr_ratio"] = lrn["sim_yr_per_hr"] / ref["sim_yr_per_hr"]
    return out


def _checkpoint_bodies(d):
    """The disk's current state as a body set (internal units), for a checkpoint."""
    import numpy as np
    from rednative import files
    s = d.state()
    n = len(s["id"])
    return files.BodySet(d.counts(), s["id"].astype(np.int32), [f"b{k}" for k in range(n)], s["type"].astype(np.int32),
                         np.zeros(n), s["p"][:, 0].copy(), s["p"][:, 1].copy(), s["p"][:, 2].copy(), s["p"][:, 3].copy(),
                         np.zeros(n, dtype=np.int32), np.zeros(n), s["r"].copy(), s["v"].copy())


def drift_section(native, years, disk="powerlaw", report_years=None, tol=None, checkpoint_out=None, resume=None,
                  wall_limit_s=None):
    """Energy and angular-momentum drift over `years` simulated years, driven like the
    reference's main loop: ejection check, step, collision check (merges at 1 x (R_i + R_j)),
    compaction after 8 events.  disk = "powerlaw": the 16384-body disk, FP64 with the
    reference's sequential sums, RKF7(8) at tol 1e-12, close encounters at 3 mutual Hill
    radii reported; disk = "mixed": the 131072-body disk in mixed precision, Cash-Karp
    RK5(4) at tol 1e-10 (no Hill test: with RK5 any pair event accepts the attempt).  A merge
    changes the total energy by itself; the energy is measured on both sides of every
    collision pass, so those jumps are kept apart from the integration drift.  tol replaces
    the disk's step tolerance.  checkpoint_out: write the state, time, trial step and the
    run's totals there at the end (npz), stopping early once the run has taken wall_limit_s
    seconds; resume: continue such a checkpoint up to `years` after the original start (the
    state, time and trial step are the integrator's whole state, so the continued run steps
    as an uninterrupted one would, apart from the count of events since the last compaction,
    which restarts at 0)."""
    import numpy as np
    from rednative import files, generate
    ck = None
    if resume:
        ck = dict(np.load(resume, allow_pickle=False))
        disk = str(ck["disk"])
        tol = float(ck["tol"])
        n = int(ck["r"].shape[0])
        resumed = files.BodySet([int(x) for x in ck["counts"]], ck["id"].astype(np.int32),
                                [f"b{k}" for k in range(n)], ck["type"].astype(np.int32), np.zeros(n), ck["mass"],
                                ck["radius"], ck["density"], ck["cd"], np.zeros(n, dtype=np.int32), np.zeros(n),
                                ck["r"], ck["v"])
    if disk in ("mixed", "mixed-disk-fp64"):
        prec = native.PREC_MIXED if disk == "mixed" else native.PREC_FP64
        b = resumed if ck is not None else generate.to_internal(generate.disk_mixed(n_pl=65536, n_tp=65533, seed=303))
        tol = 1e-10 if tol is None else tol
        d = native.Disk(b, integrator=3, adaptive=True, tolerance=tol, dt_day=1.0, precision=prec,
                        radii_enhance_factor=1.0, ejection=100.0, event_capacity=65536)
        label = ("1 star + 2 giants + 65536 planetesimals + 65533 test particles (seed 303), "
                 + ("mixed precision" if disk == "mixed" else "FP64 sequential sums")
                 + f", Cash-Karp RK5(4) tol {tol:g}, collisions 1x(R_i+R_j) merged, ejection 100 AU")
    else:
        b = resumed if ck is not None else generate.to_internal(generate.disk_powerlaw(n=16383, seed=202))
        tol = 1e-12 if tol is None else tol
        d = native.Disk(b, integrator=4, adaptive=True, tolerance=tol, dt_day=0.01, use_graph=True,
                        radii_enhance_factor=1.0, encounter_hill_factor=3.0, ejection=100.0, event_capacity=65536)
        label = (f"1 star + 16383 protoplanets/planetesimals (power-law masses, seed 202), RKF7(8) tol {tol:g}, "
                 "FP64 sequential sums, collisions 1x(R_i+R_j) merged, close encounters at 3 mutual Hill radii, "
                 "ejection 100 AU")
    if report_years is None:
        report_years = years / 10.0
    E0, L0 = d.energy()
    t0 = d.t
    jumps = 0.0
    checkpoints = []
    wall_before = 0.0
    done_before = dict(passed=0, tried=0, n_collision=0, n_close_encounter=0, n_ejection=0)
    if ck is not None:
        d.t = float(ck["t"])
        d.dt_try = float(ck["dt_try"])
        t0 = float(ck["t0"])
        E0 = float(ck["E0"])
        L0 = ck["L0"]
        jumps = float(ck["jumps"])
        checkpoints = [list(x) for x in ck["checkpoints"]]
        wall_before = float(ck["wall"])
        done_before = {k: int(ck["done_" + k]) for k in done_before}
    t_stop = t0 + years * 365.25 * GAUSS
    next_report = report_years * (1 + int(((d.t - t0) / GAUSS / 365.25) / report_years + 1e-9))
    w0 = time.perf_counter()
    while d.t <= t_stop:
        if d.check_for_ejection_hit_centrum():
            d.clear_event_counter()
        d.step()
        if d.n_event() > 0:
            Ea, _ = d.energy()
            if d.check_for_collision():
                d.clear_event_counter()
            Eb, _ = d.energy()
            jumps += Eb - Ea
        d.check_for_rebuild_vectors(8)
        yrs = (d.t - t0) / GAUSS / 365.25
        if yrs >= next_report:
            E, L = d.energy()
            drift = abs((E - E0 - jumps) / E0)
            checkpoints.append([round(yrs, 3), drift])
            print(f"drift: {yrs:.1f} yr, integration dE/E {drift:.3e}, merges dE/E {jumps / E0:.3e}, "
                  f"{time.perf_counter() - w0 + wall_before:.0f} s", file=sys.stderr, flush=True)
            next_report += report_years
        if checkpoint_out and wall_limit_s and time.perf_counter() - w0 > wall_limit_s:
            break
    d.sync()
    sec = time.perf_counter() - w0 + wall_before
    E1, L1 = d.energy()
    c = {k: v + done_before.get(k, 0) for k, v in d.counters().items()}
    sim = (d.t - t0) / GAUSS / 365.25
    if checkpoint_out:
        st = _checkpoint_bodies(d)
        np.savez(checkpoint_out, disk=disk, tol=tol, counts=np.array(st.counts), id=st.ids, type=st.types,
                 mass=st.mass, radius=st.radius, density=st.density, cd=st.cd, r=st.r, v=st.v, t=d.t,
                 dt_try=d.dt_try, t0=t0, E0=E0, L0=np.asarray(L0), jumps=jumps,
                 checkpoints=np.array(checkpoints, dtype=np.float64).reshape(-1, 2), wall=sec,
                 **{"done_" + k: c[k] for k in ("passed", "tried", "n_collision", "n_close_encounter", "n_ejection")})
    out = {"workload": label,
           "sim_years": sim, "seconds": sec, "sim_yr_per_hr": sim / (sec / 3600.0),
           "dE_over_E_integration": abs((E1 - E0 - jumps) / E0), "dE_over_E_merges": jumps / E0,
           "dE_over_E_total": abs((E1 - E0) / E0),
           "dL_over_L": float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0)),
           "steps": c["passed"], "attempts": c["tried"], "collisions": c["n_collision"],
           "close_encounters": c["n_close_encounter"], "ejections": c["n_ejection"], "bodies_left": d.n,
           "checkpoints": checkpoints, "complete": bool(d.t > t_stop)}
    if len(checkpoints) >= 2:
        # linear drift rate from the checkpoints (least squares through the origin)
        yy = np.array(checkpoints, dtype=np.float64)
        out["dE_over_E_per_year"] = float(np.dot(yy[:, 0], yy[:, 1]) / np.dot(yy[:, 0], yy[:, 0]))
        out["dE_over_E_1000yr_extrapolated"] = out["dE_over_E_per_year"] * 1000.0
    d.close()
    return out


def small_disk_cpu_baseline(bodies, kw, prm):
    """CPU leg of the small-disk section: the oracle (one thread, like the reference's CPU
    path) through the same main loop; returns the wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from rednative import mainloop
    o = pyoracle.Oracle(bodies, **kw)
    t0 = time.perf_counter()
    mainloop.run_main_loop(o, prm, lambda e: None)
    return time.perf_counter() - t0


def small_disk_section(native, with_cpu=True, use_graph=True, fused_only=False):
    """The 256-body workload at its full length: star + 2 giants + 253 planetesimals (seed
    101), RK4 at dt = 1 day, 10^4 steps of the reference's main loop (ejection at 100 AU,
    collisions at 1 x (R_i + R_j), a snapshot read back every 365.25 days).  The GPU runs it
    with the fused fixed-step loop (red_run_fixed) and one step at a time (HIP graphs); the
    final states must be identical.  use_graph=False: the step-by-step leg without HIP graphs
    (for kernel traces); fused_only: the fused leg alone (for counter runs, which sample
    every launch)."""
    import numpy as np
    from rednative import generate, mainloop
    b = generate.to_internal(generate.disk_small(seed=101))
    kw = dict(integrator=2, adaptive=False, dt_day=1.0, ejection=100.0, radii_enhance_factor=1.0)
    prm = dict(stop_time=9999.5 * GAUSS, output_interval=365.25 * GAUSS, inner_steps=True, radii_enhance_factor=1.0)
    out = {"workload": "1 star + 2 giants + 253 planetesimals (seed 101), RK4 dt = 1 day, 10^4 steps, "
                       "ejection 100 AU, collisions 1x(R_i+R_j), snapshot every 365.25 days"}
    states = {}
    for mode in ("fused",) if fused_only else ("fused", "stepwise"):
        d = native.Disk(b, use_graph=use_graph, **kw)
        mainloop.run_main_loop(d, dict(prm, stop_time=20.5 * GAUSS), lambda e: None, fused=(mode == "fused"))
        d = native.Disk(b, use_graph=use_graph, **kw)
        d.sync()
        snaps = []
        t0 = time.perf_counter()
        info = mainloop.run_main_loop(d, prm, lambda e: snaps.append(e.t), fused=(mode == "fused"))
        d.sync()
        sec = time.perf_counter() - t0
        s = d.state()
        states[mode] = (s["r"].copy(), s["v"].copy())
        out[mode] = {"seconds": sec, "us_per_step": sec / info["steps"] * 1e6, "steps": info["steps"],
                     "snapshots": len(snaps), "sim_yr_per_hr": (d.t / GAUSS / 365.25) / (sec / 3600.0)}
        d.close()
    if fused_only:
        return out
    out["identical"] = bool(all(np.array_equal(states["fused"][k].view(np.uint64), states["stepwise"][k].view(np.uint64))
                                for k in range(2)))
    if with_cpu:
        sec = small_disk_cpu_baseline(b, kw, prm)
        out["cpu_baseline"] = {"seconds": sec, "us_per_step": sec / 1e4 * 1e6, "cores": 1, "kind": "port"}
        out["fused_vs_cpu"] = sec / out["fused"]["seconds"]
    out["fused_vs_stepwise"] = out["stepwise"]["seconds"] / out["fused"]["seconds"]
    return out


def ensemble_section(native, days, sample, use_graph=True):
    """512 disks (N log-uniform in [64, 4096], seeded; each a star, the two giants and
    planetesimals out to its own radius) integrated together with the 256-body disk's rules
    (collisions merged at 1 x (R_i + R_j), close encounters at 3 mutual Hill radii reported,
    ejection at 100 AU), against the same disks stepped one context at a time through the
    same driver loop (a sample of them, scaled): disk-steps per second.  use_graph=False:
    both without HIP graphs (for kernel traces)."""
    from rednative import generate, mainloop
    disks = [generate.to_internal(d) for d in generate.ensemble(n_disks=512, seed=404)]
    rules = dict(radii_enhance_factor=1.0, encounter_hill_factor=3.0)
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001, ejection=100.0, **rules)
    t_stop = days * GAUSS
    batch = 8   # supersteps per graph launch
    E = native.Ensemble(disks, use_graph=use_graph, batch=batch, **kw)
    block, jtile, _ = E.autotune(reps=3)
    t0 = time.perf_counter()
    info = E.advance(t_stop)
    t_ens = time.perf_counter() - t0
    cnt = [E.counters(k) for k in range(len(disks))]
    steps_ens = sum(c["passed"] for c in cnt)
    ev = {k: sum(c[k] for c in cnt) for k in ("n_collision", "n_close_encounter", "n_ejection", "n_hit_centrum")}
    E.close()
    prm = dict(stop_time=t_stop, output_interval=1e300, inner_steps=True, **rules)
    t1 = time.perf_counter()
    steps_one = 0
    # a sample spread evenly over the ensemble (disk sizes are drawn in random order)
    picks = sorted({round(i * (len(disks) - 1) / max(sample - 1, 1)) for i in range(sample)})
    for k in picks:
        d = native.Disk(disks[k], use_graph=use_graph, **kw)
        mainloop.run_main_loop(d, prm, lambda e: None)
        d.sync()
        steps_one += d.counters()["passed"]
        d.close()
    t_one = time.perf_counter() - t1
    rate_ens = steps_ens / t_ens
    rate_one = steps_one / t_one
    return {"workload": f"512 disks, N log-uniform in [64, 4096] (seed 404), each star + 2 giants + planetesimals, "
                        f"RKF7(8) tol 1e-11, collisions 1x(R_i+R_j) merged, close encounters at 3 mutual Hill radii, "
                        f"ejection 100 AU, {days} days each",
            "ensemble_disk_steps_per_s": rate_ens, "per_disk_disk_steps_per_s": rate_one,
            "ensemble_bodies": sum(d.n for d in disks),
            "per_disk_sample": len(picks), "per_disk_sample_bodies": sum(disks[k].n for k in picks),
            "speedup": rate_ens / rate_one, "supersteps": info["supersteps"],
            "ms_per_superstep": 1e3 * t_ens / max(info["supersteps"], 1), "seconds": t_ens, "events": ev,
            "host_event_seconds": info["event_seconds"], "host_rebuild_seconds": info["rebuild_seconds"],
            "batches": info["batches"], "supersteps_per_batch": batch,
            "autotuned_block": block, "autotuned_jtile": jtile}


PEAK_HBM_GBS = 8000.0         # MI355X HBM3E, vendor figure
# Non-zero coefficients of the Cash-Karp stage rows 1..5 and of y_(n+1) (5th-order weights),
# and of the error estimate (the differences of the two weight sets).
CK_STAGE_TERMS = (1, 2, 3, 4, 5, 4)
CK_ERR_TERMS = 5


def hbm_rooflines(r, bodies, pmc_path):
    """HBM roofline of the stage combinations, the error estimate and the ejection scan of a
    timed run: algorithmic bytes (x, y, z of each vector read or written, 24 B per body and
    vector: a stage row reads y and its k derivative records and writes the stage state, for
    positions and velocities alike; the error estimate reads its derivative records of the
    massive bodies; the scan reads x, y, z, id and type, 32 B per body) over the launches'
    average duration.  These launches last a few microseconds, so the HIP events recorded
    around each of them on the context stream add their own gaps (about 3 us): the duration
    comes from the committed kernel trace of the same kernels when there is one
    (duration_source), and the live HIP-event figure is reported beside it.  traffic: HBM
    bytes per launch from the committed counter summary (FETCH_SIZE and WRITE_SIZE scaled by
    a copy of known size)."""
    n = bodies.n
    nm = sum(bodies.counts[:6])
    stage_attempt = sum(48 * n * (k + 2) for k in CK_STAGE_TERMS)
    per_launch = {"k_stage": stage_attempt / len(CK_STAGE_TERMS), "k_err": 48.0 * nm * CK_ERR_TERMS,
                  "k_ejection": 32.0 * n}
    pmc = {}
    try:
        with open(pmc_path) as f:
            pmc = json.load(f)
    except (OSError, ValueError):
        pass
    out = {}
    for key, cls in (("k_stage", "stage"), ("k_err", "err"), ("k_ejection", "eject")):
        ms, cnt = r["prof"][cls]
        if not cnt or ms <= 0.0:
            continue
        sec_ev = ms * 1e-3 / cnt
        k = pmc.get(key, {})
        sec = k["duration_us_mean"] * 1e-6 if k.get("duration_us_mean") else sec_ev
        achieved = per_launch[key] / sec / 1e9
        out[key] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": achieved / PEAK_HBM_GBS, "traffic": k.get("hbm_bytes_per_launch"),
                    "algorithmic_bytes_per_launch": per_launch[key], "launch_us": sec * 1e6,
                    "duration_source": "kernel trace (profiles/pmc_stage_err_eject_n131072.md)"
                    if k.get("duration_us_mean") else "HIP events",
                    "launch_us_hip_events": sec_ev * 1e6, "achieved_hip_events": per_launch[key] / sec_ev / 1e9,
                    "launches": cnt}
    return out


def load_drift(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def bound_tolerance_section(native, bodies, steps_mixed, steps_fp64, drift_path, tol=1e-11, use_graph=True):
    """Both paths on the same disk at the tolerance whose energy drift meets dE/E <= 1e-9
    over 1000 yr (Cash-Karp's truncation drift is the same in both precisions; at 1e-10 it
    extrapolates to 3.4e-9): simulated years per hour, their ratio, and the mixed path's
    measured drift rate from the committed long run, extrapolated to 1000 yr."""
    GAUSS_YR = GAUSS * 365.25
    res = {"tolerance": tol}
    for name, prec, steps in (("mixed", native.PREC_MIXED, steps_mixed), ("fp64", native.PREC_FP64, steps_fp64)):
        r = run_path(native, bodies, prec, steps, 2, tol=tol, use_graph=use_graph)
        yrs = r["sim_t"] / GAUSS_YR
        res[name] = {"sim_yr_per_hr": yrs / (r["seconds"] / 3600.0), "ms_per_step": 1e3 * r["seconds"] / steps,
                     "steps": steps, "attempts": r["tries"], "sim_days": r["sim_t"] / GAUSS,
                     "dE_over_E_timed": abs((r["E1"] - r["E0"]) / r["E0"])}
    res["speedup_sim_yr_per_hr_vs_fp64"] = res["mixed"]["sim_yr_per_hr"] / res["fp64"]["sim_yr_per_hr"]
    dr = load_drift(drift_path)
    if dr:
        res["drift"] = {k: dr.get(k) for k in ("workload", "sim_years", "dE_over_E_integration",
                                               "dE_over_E_per_year", "dE_over_E_1000yr_extrapolated",
                                               "dL_over_L", "sim_yr_per_hr", "source")}
        res["drift"]["label"] = "measured over sim_years, extrapolated to 1000 yr"
    return res


def ensemble_span_section(native, years, chunk_days=100.0, use_graph=True):
    """The 512-disk workload of ensemble_section over its full span: advanced in chunks, each
    followed by draining every disk's event records into a reusable buffer (as a caller
    writing the event file would), so host memory stays bounded.  Reports the throughput of
    the advance calls alone and with the drains."""
    from rednative import generate
    disks = [generate.to_internal(d) for d in generate.ensemble(n_disks=512, seed=404)]
    kw = dict(integrator=4, adaptive=True, tolerance=1e-11, dt_day=0.001, ejection=100.0,
              radii_enhance_factor=1.0, encounter_hill_factor=3.0)
    E = native.Ensemble(disks, use_graph=use_graph, **kw)
    block, jtile, _ = E.autotune(reps=3)
    buf = (native.RedEvent * (1 << 16))()
    t_adv = t_drain = 0.0
    supersteps = 0
    records = 0
    t_stop_days = 365.25 * years
    day = 0.0
    while day < t_stop_days:
        day = min(t_stop_days, day + chunk_days)
        t0 = time.perf_counter()
        info = E.advance(day * GAUSS)
        t1 = time.perf_counter()
        for k in range(len(disks)):
            while True:
                got, left = E.drain_events(k, buf)
                records += got
                if got >= left:
                    break
        t2 = time.perf_counter()
        t_adv += t1 - t0
        t_drain += t2 - t1
        supersteps += info["supersteps"]
        print(f"ensemble span: {day / 365.25:.2f} yr, {supersteps} supersteps, {records} records, "
              f"{t_adv:.1f} s advancing, {t_drain:.1f} s draining", file=sys.stderr, flush=True)
    cnt = [E.counters(k) for k in range(len(disks))]
    steps = sum(c["passed"] for c in cnt)
    ev = {k: sum(c[k] for c in cnt) for k in ("n_collision", "n_close_encounter", "n_ejection", "n_hit_centrum")}
    E.close()
    return {"workload": f"512 disks, N log-uniform in [64, 4096] (seed 404), each star + 2 giants + planetesimals, "
                        f"RKF7(8) tol 1e-11, collisions 1x(R_i+R_j) merged, close encounters at 3 mutual Hill "
                        f"radii, ejection 100 AU, {years} yr each, event records drained every {chunk_days} days",
            "disk_steps": steps, "supersteps": supersteps, "seconds_advance": t_adv, "seconds_drain": t_drain,
            "ensemble_disk_steps_per_s": steps / t_adv, "ensemble_disk_steps_per_s_with_drain": steps / (t_adv + t_drain),
            "ms_per_superstep": 1e3 * t_adv / max(supersteps, 1), "records_drained": records, "events": ev,
            "autotuned_block": block, "autotuned_jtile": jtile}


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (the runtime's copy kernel): the reference
    point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0        # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a); hip.hipFree(b)
    print(json.dumps({"copy_bytes": n}))


def load_traffic(path):
    """HBM bytes per launch of the test-particle sweep from the committed counter summary
    (two rocprofv3 --pmc runs, FETCH_SIZE then WRITE_SIZE, scaled by a copy of known size)."""
    try:
        with open(path) as f:
            t = json.load(f)
        return float(t["hbm_bytes_per_launch"]), t
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kahan", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--pmc-summary", default=os.path.join(ROOT, "profiles", "pmc_grav_mixed_ni_n131072.json"))
    ap.add_argument("--no-proposer", action="store_true")
    ap.add_argument("--proposer-years", type=float, default=5.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-days", type=float, default=100.0)
    ap.add_argument("--no-graph", action="store_true",
                    help="every section without HIP graphs (for kernel traces and counters)")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the mixed-vs-FP64 accuracy and reference-workload sections (for kernel traces "
                         "of the main sections alone)")
    ap.add_argument("--ensemble-years", type=float, default=100.0,
                    help="--only ensemble-span: years per disk (the ensemble workload's full span is 100)")
    ap.add_argument("--only", choices=["proposer", "ensemble", "ensemble-span", "drift", "smalldisk"], default=None,
                    help="run one auxiliary section alone and print its JSON (drift: the 16384-body "
                         "disk's energy drift over --drift-years, not part of the default run)")
    ap.add_argument("--fused-only", action="store_true",
                    help="with --only smalldisk: the fused leg alone (no step-by-step leg, no CPU leg)")
    ap.add_argument("--drift-years", type=float, default=1000.0)
    ap.add_argument("--drift-disk", choices=["powerlaw", "mixed", "mixed-disk-fp64"], default="powerlaw")
    ap.add_argument("--drift-tol", type=float, default=None,
                    help="step tolerance of the drift run (default 1e-12 for powerlaw, 1e-10 otherwise)")
    ap.add_argument("--drift-report-years", type=float, default=None)
    ap.add_argument("--drift-checkpoint-out", default=None,
                    help="write the drift run's state and totals there (npz) at its end")
    ap.add_argument("--drift-wall-limit", type=float, default=None,
                    help="with --drift-checkpoint-out: stop after this many seconds of the run")
    ap.add_argument("--drift-resume", default=None, help="continue the drift run of this checkpoint")
    ap.add_argument("--hbm-summary", default=os.path.join(ROOT, "profiles", "pmc_stage_err_eject_n131072.json"))
    ap.add_argument("--drift-summary", default=os.path.join(ROOT, "profiles", "drift_mixed_n131072_tol1e-11.json"))
    ap.add_argument("--no-bound", action="store_true", help="skip the section at tolerance 1e-11")
    ap.add_argument("--valu-rates", action="store_true",
                    help="print the vector-ALU issue rates of the GPU (red_valu_rates) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the force sweep would shard by sinks with an all-gather of source positions).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from rednative import generate, native
    native.load_library()
    if args.valu_rates:
        print(json.dumps({"valu_rates": native.valu_rates()}))
        return
    if args.only == "proposer":
        print(json.dumps({"proposer": proposer_section(native, args.proposer_years, not args.no_graph)}))
        return
    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_section(native, args.ensemble_days, 16, not args.no_graph)}))
        return
    if args.only == "ensemble-span":
        print(json.dumps({"ensemble_span": ensemble_span_section(native, args.ensemble_years,
                                                                 use_graph=not args.no_graph)}))
        return
    if args.only == "smalldisk":
        print(json.dumps({"small_disk": small_disk_section(native, with_cpu=not args.no_cpu,
                                                           use_graph=not args.no_graph,
                                                           fused_only=args.fused_only)}))
        return
    if args.only == "drift":
        print(json.dumps({"drift": drift_section(native, args.drift_years, args.drift_disk, tol=args.drift_tol,
                                                 report_years=args.drift_report_years,
                                                 checkpoint_out=args.drift_checkpoint_out,
                                                 resume=args.drift_resume, wall_limit_s=args.drift_wall_limit)}))
        return

    bodies = generate.to_internal(generate.disk_mixed(n_pl=65536, n_tp=65533, seed=303))
    pairs_eval, parts = pair_counts(bodies.counts)

    graphs = not args.no_graph
    mixed = run_path(native, bodies, native.PREC_MIXED, args.steps, args.warmup, kahan=args.kahan, use_graph=graphs)
    fp64 = None if args.no_fp64 else run_path(native, bodies, native.PREC_FP64, args.fp64_steps,
                                                   max(1, args.warmup // 2), use_graph=graphs)

    def summarize(r, steps):
        evals = steps + 5 * r["tries"]            # f0 once per step, 5 stages per attempt
        pairs = pairs_eval * evals
        yrs = r["sim_t"] / GAUSS / 365.25
        return dict(pairs_per_s=pairs / r["seconds"], sim_yr_per_hr=yrs / (r["seconds"] / 3600.0),
                    ms_per_step=1e3 * r["seconds"] / steps, attempts=r["tries"], steps=steps,
                    dE_over_E=abs((r["E1"] - r["E0"]) / r["E0"]), sim_days=r["sim_t"] / GAUSS)

    sm = summarize(mixed, args.steps)
    s64 = summarize(fp64, args.fp64_steps) if fp64 else None

    ni_ms, ni_n = mixed["prof"]["ni"]
    avg_ni_s = (ni_ms / max(ni_n, 1)) * 1e-3
    ni_pairs = bodies.counts[6] * sum(bodies.counts[:6])
    achieved = FLOP_PER_PAIR * ni_pairs / avg_ni_s / 1e12
    if fp64:
        ni64_ms, ni64_n = fp64["prof"]["ni"]
        achieved64 = FLOP_PER_PAIR * ni_pairs / ((ni64_ms / max(ni64_n, 1)) * 1e-3) / 1e12
    traffic, traffic_src = load_traffic(args.pmc_summary)
    # the rate of the sweep's own pair sequence alone on this GPU (sources in cache, no
    # memory traffic), measured here beside it
    rates = native.valu_rates()
    ni_pairs_per_s = ni_pairs / avg_ni_s

    acc = None if args.no_fp64 or args.no_aux else accuracy_section(native, bodies)
    refw = None if args.no_fp64 or args.no_aux else reference_workload_section(native)
    cpu = None if args.no_cpu else cpu_baseline(bodies)
    prop = None if args.no_proposer else proposer_section(native, args.proposer_years, graphs)
    ens = None if args.no_ensemble else ensemble_section(native, args.ensemble_days, 16, graphs)
    small = None if args.no_aux else small_disk_section(native, with_cpu=not args.no_cpu, use_graph=graphs)
    bound = None if args.no_bound or args.no_fp64 else bound_tolerance_section(
        native, bodies, max(10, args.steps // 3), max(4, args.fp64_steps // 3), args.drift_summary,
        use_graph=graphs)

    line = {
        "metric": "pair-interactions/s & sim-yr/hr at N=131072, mixed vs FP64, ΔE/E≤1e-9",
        "value": sm["pairs_per_s"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 differences + f32 r^-3 (mixed)",
        "data": "synthetic (seeded disk, 1 star + 2 giants + 65536 planetesimals + 65533 test particles)",
        "config": {"workload": "disk N=131072 (3 SI + 65536 NSI + 65533 NI), Cash-Karp RK5(4) tol 1e-10, "
                               "collisions 1x(R_i+R_j) merged, ejection 100 AU",
                   "pairs_per_eval": pairs_eval, "pairs_by_sweep": parts, "integrator": "rk5(4)"},
        "sim_yr_per_hr": sm["sim_yr_per_hr"],
        "dE_over_E_timed": sm["dE_over_E"],
        "sim_days_timed": sm["sim_days"],
        "attempts_timed": sm["attempts"],
        "fp64": None if not s64 else
                {"pairs_per_s": s64["pairs_per_s"], "sim_yr_per_hr": s64["sim_yr_per_hr"],
                 "ms_per_step": s64["ms_per_step"], "dE_over_E_timed": s64["dE_over_E"],
                 "steps": args.fp64_steps, "ni_sweep_tflops_20flop": achieved64,
                 "ni_sweep_frac_fp64_peak": achieved64 / PEAK_FP64_TFLOPS,
                 "ni_sweep_vs_term_sequence_alone": achieved64 * 1e12 / FLOP_PER_PAIR / rates["fp64_pairs_per_s"]},
        "mixed_vs_fp64": acc,
        "reference_workload": refw,
        "speedup_sim_yr_per_hr_vs_fp64": sm["sim_yr_per_hr"] / s64["sim_yr_per_hr"] if s64 else None,
        "speedup_pairs_per_s_vs_fp64": sm["pairs_per_s"] / s64["pairs_per_s"] if s64 else None,
        "roofline": {"bound": "valu", "kernel": "k_grav_mixed_bcast (test-particle sweep)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": None if not traffic_src else traffic_src.get("source"),
                     "launch_ms": avg_ni_s * 1e3, "pairs_per_launch": ni_pairs,
                     "flop_per_pair": FLOP_PER_PAIR,
                     "duration_source": "HIP events on the context stream around every launch, in a "
                                        "second pass over the same number of steps without graphs "
                                        "(the timed pass replays graphs with no events)",
                     "pair_sequence_alone": {"pairs_per_s": rates["mixed_pairs_per_s"],
                                             "sweep_ratio": ni_pairs_per_s / rates["mixed_pairs_per_s"],
                                             "lane_instr_per_s": {k: v for k, v in rates.items()
                                                                  if not k.endswith("pairs_per_s")},
                                             "source": "red_valu_rates: the sweep's pair sequence (the same "
                                                       "132 vector instructions per 8 pairs) alone, its "
                                                       "sources in cache, and single-kind instruction "
                                                       "chains; measured beside the sweep in this run"}},
        "roofline_hbm": hbm_rooflines(mixed, bodies, args.hbm_summary),
        "at_bound_tolerance": bound,
        "profile_ms": {k: v[0] for k, v in mixed["prof"].items() if v[1]},
        "cpu_baseline": cpu,
        "proposer": prop,
        "ensemble": ens,
        "small_disk": small,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
