"""The drop-in driver `red` (C++ over the C ABI).

CPU tests: the command line and parameter-file errors the reference raises
(options.cpp:97-195, parameter.cpp:55-285) are reported before any GPU call.
GPU tests: the driver reproduces the stored TwoBody run, and its result file equals
the Python driver loop over the same library.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG
from rednative import files, generate

RED = os.path.join(PKG, "lib", "red")
GAUSS = files.GAUSS


def run_red(args, cwd=None):
    return subprocess.run([RED] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


def test_help_and_bad_switch():
    r = run_red(["-h"])
    assert r.returncode == 0 and "-iDir" in r.stdout
    r = run_red(["-bogus"])
    assert r.returncode != 0 and "Invalid switch on command-line: -bogus." in r.stderr
    r = run_red(["-ic", "x.txt"])
    assert "Missing filename for -p" in r.stderr
    r = run_red(["-cpu", "-p", "a", "-ic", "b"])
    assert r.returncode != 0 and "GPU only" in r.stderr


@pytest.mark.parametrize("line,msg", [("foo = 1", "Invalid parameter: foo."),
                                      ("tolerance = 1e-x0", "Invalid number at: tolerance"),
                                      ("integrator = rk7", "Invalid integrator type: rk7"),
                                      ("close_encounter = maybe", "Invalid value at: close_encounter"),
                                      ("integrator = rkn", "not implemented")])
def test_parameter_errors(tmp_path, line, msg):
    (tmp_path / "p.txt").write_text("integrator = rk4\n" + line + "\n")
    (tmp_path / "b.txt").write_text((open(os.path.join(GOLDEN, "TwoBody2.txt")).read()))
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "b.txt"])
    assert r.returncode != 0 and msg in r.stderr, r.stderr
    p = files.parse_parameters if "Invalid" in msg else None
    if p:
        with pytest.raises(ValueError) as ei:
            p("integrator = rk4\n" + line + "\n")
        assert msg in str(ei.value)


def test_missing_body_file(tmp_path):
    (tmp_path / "p.txt").write_text("integrator = e\nlength = 1\n")
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "nothere.txt"])
    assert r.returncode != 0 and "Cannot open" in r.stderr


@pytest.mark.gpu
def test_driver_twobody_golden(tmp_path):
    for f in ("TwoBody2.txt", "parameters2.txt"):
        (tmp_path / f).write_text(open(os.path.join(GOLDEN, f)).read())
    print("case driver twobody", flush=True)
    r = run_red(["-iDir", str(tmp_path), "-p", "parameters2.txt", "-ic", "TwoBody2.txt", "-ef"])
    assert r.returncode == 0, r.stderr
    out = files.parse_result_ascii(str(tmp_path / "R_gpu_ns_fs_E_result.txt"))
    gold = files.parse_result_ascii(os.path.join(GOLDEN, "D_cpu_ns_fs_E_result.txt"))
    assert len(out) == len(gold) == 42
    for a, g in zip(out, gold):
        assert a[0] == g[0] and a[1] == g[1]
        # the stored file holds t and v in internal units; the driver prints t/k and v*k
        assert float(a[3]) == float(g[3]) / GAUSS
        for c in range(10, 13):
            assert float(a[c]) == float(g[c])
        for c in range(13, 16):
            assert float(a[c]) == float(g[c]) * GAUSS
    assert os.path.exists(tmp_path / "R_gpu_ns_fs_E_log.txt")


@pytest.mark.gpu
def test_driver_events_match_python_loop(tmp_path, native_mod):
    from rednative import mainloop
    b = generate.disk_allsi(150, seed=17, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    files.write_body_file(str(tmp_path / "disk.txt"), b)
    (tmp_path / "p.txt").write_text("integrator = rk4\nlength = 0.2\noutput_interval = 0.05\n"
                                    "collision_factor = 300\nejection = 50\ndt = 0.5\n")
    print("case driver events", flush=True)
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "disk.txt"])
    assert r.returncode == 0, r.stderr
    res = files.parse_result_ascii(str(tmp_path / "result.txt"))
    ev = [ln.split() for ln in open(tmp_path / "event.txt") if ln.strip()]
    # the same run through the Python driver loop
    prm = files.read_parameters(str(tmp_path / "p.txt"))
    bi = generate.to_internal(files.read_body_file(str(tmp_path / "disk.txt")))
    d = native_mod.Disk(bi, integrator=prm["integrator"], dt_day=prm["dt"],
                        radii_enhance_factor=prm["radii_enhance_factor"], ejection=prm["ejection"])
    d.t = prm["start_time"]
    snaps, events = [], []

    def snap(e):
        s = e.state()
        keep = (s["id"] > 0)
        snaps.append((e.t, s["id"][keep], s["r"][keep]))

    mainloop.run_main_loop(d, prm, snap, lambda evs: events.extend(evs))
    assert len(ev) == len(events) and len(events) > 0
    for line, e in zip(ev, events):
        assert int(line[3]) == e["id1"] and int(line[4]) == e["id2"]
        assert float(line[2]) == float(f"{e['d']:.16e}")
    rows = [row for row in res]
    n_last = len(snaps[-1][1])
    last = rows[-n_last:]
    for row, i, rr in zip(last, snaps[-1][1], snaps[-1][2]):
        assert int(row[0]) == i
        assert [float(x) for x in row[10:13]] == [float(f"{x:.16e}") for x in rr]


@pytest.mark.gpu
def test_driver_fused_matches_stepwise(tmp_path, native_mod):
    """A fixed-step run with collisions and an ejection: the driver's fused device loop
    (red_run_fixed between the iterations that need the host) and -stepwise (one red_step
    per iteration) write identical result and event files.  120 self-interacting bodies:
    14400 interacting pairs per evaluation, within the fused loop's limit."""
    b = generate.disk_allsi(120, seed=17, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    print("case driver fused vs stepwise", flush=True)
    out = {}
    for mode in ("fused", "stepwise"):
        d = tmp_path / mode
        d.mkdir()
        files.write_body_file(str(d / "disk.txt"), b)
        (d / "p.txt").write_text("integrator = rk4\nlength = 0.2\noutput_interval = 0.01\n"
                                 "collision_factor = 300\nejection = 50\ndt = 0.5\n")
        extra = ["-stepwise"] if mode == "stepwise" else []
        r = run_red(["-iDir", str(d), "-p", "p.txt", "-ic", "disk.txt"] + extra)
        assert r.returncode == 0, r.stderr
        out[mode] = ((d / "result.txt").read_bytes(), (d / "event.txt").read_bytes())
    assert out["fused"][0] == out["stepwise"][0]
    assert out["fused"][1] == out["stepwise"][1] and len(out["fused"][1]) > 0


@pytest.mark.gpu
def test_driver_astro_frame_matches_python_loop(tmp_path, native_mod):
    """frame_center = astro: the driver shifts a body list whose star is off the origin into
    the star's frame (tools.cpp:209-236) and integrates there; the Python loop over the
    library, given the same shift, writes the same final state."""
    from rednative import mainloop
    b = generate.disk_allsi(120, seed=19, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    b.r = b.r + np.array([1e-3, -2e-3, 5e-4])
    b.v = b.v + np.array([1e-5, 2e-5, -1e-5])
    files.write_body_file(str(tmp_path / "disk.txt"), b)
    (tmp_path / "p.txt").write_text("integrator = rkf8\ntolerance = 1e-10\nlength = 0.1\n"
                                    "output_interval = 0.05\nframe_center = astro\n"
                                    "collision_factor = 300\ndt = 0.5\n")
    print("case driver astro", flush=True)
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "disk.txt"])
    assert r.returncode == 0, r.stderr
    res = files.parse_result_ascii(str(tmp_path / "result.txt"))
    prm = files.read_parameters(str(tmp_path / "p.txt"))
    assert prm["frame_center"] == 0
    bf = files.read_body_file(str(tmp_path / "disk.txt"))
    native_mod.host_transform_to_ac(bf.types, bf.r, bf.v)
    bi = generate.to_internal(bf, transform_bc=False)
    d = native_mod.Disk(bi, integrator=prm["integrator"], adaptive=prm["adaptive"], tolerance=prm["tolerance"],
                        dt_day=prm["dt"], radii_enhance_factor=prm["radii_enhance_factor"], frame_center=0)
    d.t = prm["start_time"]
    snaps = []

    def snap(e):
        s = e.state()
        keep = (s["id"] > 0)
        snaps.append((s["id"][keep], s["r"][keep]))

    mainloop.run_main_loop(d, prm, snap)
    ids, rr = snaps[-1]
    last = list(res)[-len(ids):]
    assert int(last[0][0]) == ids[0] and float(last[0][10]) == 0.0     # the star at the origin
    for row, i, x in zip(last, ids, rr):
        assert int(row[0]) == i
        assert [float(v) for v in row[10:13]] == [float(f"{c:.16e}") for c in x]


def _write_snapshots(path, snaps):
    lines = []
    for t_day, b in snaps:
        for k in range(b.n):
            lines.append(files.format_result_line(b.ids[k], b.names[k], b.types[k], t_day, b.params4()[k],
                                                  b.mig_type[k], b.mig_stop_at[k], b.r[k], b.v[k]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_result_snapshot_reader(tmp_path):
    b0 = generate.disk_small(20, seed=4)
    b1 = generate.disk_small(12, seed=5)
    _write_snapshots(tmp_path / "r.txt", [(0.0, b0), (36.5, b1)])
    t, b = files.read_result_snapshot(str(tmp_path / "r.txt"))
    assert t == 36.5 and b.n == b1.n and list(b.counts) == list(b1.counts)
    assert np.array_equal(b.ids, b1.ids) and np.array_equal(b.r, b1.r) and np.array_equal(b.v, b1.v)
    assert np.array_equal(b.mass, b1.mass)
    bad = generate.disk_small(12, seed=5)
    bad.types = bad.types[::-1].copy()
    _write_snapshots(tmp_path / "bad.txt", [(1.0, bad)])
    with pytest.raises(ValueError, match="type order"):
        files.read_result_snapshot(str(tmp_path / "bad.txt"))


def test_restart_errors(tmp_path):
    (tmp_path / "p.txt").write_text("integrator = rk4\nlength = 0.1\n")
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-restart", "nothere.txt"])
    assert r.returncode != 0 and "Cannot open" in r.stderr
    (tmp_path / "r.txt").write_text("1 star 0 0.0 1.0\n")
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-restart", "r.txt"])
    assert r.returncode != 0 and "Invalid result line" in r.stderr
    bad = generate.disk_small(12, seed=5)
    bad.types = bad.types[::-1].copy()
    _write_snapshots(tmp_path / "bad.txt", [(1.0, bad)])
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-restart", "bad.txt"])
    assert r.returncode != 0 and "type order" in r.stderr


@pytest.mark.gpu
def test_driver_restart_from_snapshot(tmp_path, native_mod):
    """A run stopped at 0.1 yr and resumed with -restart from its last text snapshot ends
    where the uninterrupted 0.2 yr run ends.  Fixed-step RK4, so the step sequence does not
    depend on the controller state the snapshot lacks; the snapshot's v * k and t / k
    round trips leave differences of a few ulps, hence the tolerance."""
    b = generate.disk_allsi(100, seed=23, m_range=(1e-10, 1e-8))
    for sub in ("full", "half"):
        (tmp_path / sub).mkdir()
        files.write_body_file(str(tmp_path / sub / "disk.txt"), b)
    prm = "integrator = rk4\noutput_interval = 0.1\ndt = 0.5\nlength = {}\n"
    (tmp_path / "full" / "p.txt").write_text(prm.format(0.2))
    (tmp_path / "half" / "p1.txt").write_text(prm.format(0.1))
    (tmp_path / "half" / "p.txt").write_text(prm.format(0.2))
    print("case driver restart", flush=True)
    r = run_red(["-iDir", str(tmp_path / "full"), "-p", "p.txt", "-ic", "disk.txt"])
    assert r.returncode == 0, r.stderr
    r = run_red(["-iDir", str(tmp_path / "half"), "-p", "p1.txt", "-ic", "disk.txt"])
    assert r.returncode == 0, r.stderr
    t_mid, _ = files.read_result_snapshot(str(tmp_path / "half" / "result.txt"))
    assert t_mid > 36.0
    r = run_red(["-iDir", str(tmp_path / "half"), "-p", "p.txt", "-restart", "result.txt", "-result", "resumed"])
    assert r.returncode == 0, r.stderr
    t_full, full = files.read_result_snapshot(str(tmp_path / "full" / "result.txt"))
    t_res, res = files.read_result_snapshot(str(tmp_path / "half" / "resumed.txt"))
    first = files.parse_result_ascii(str(tmp_path / "half" / "resumed.txt"))[0]
    assert float(first[3]) == t_mid                     # resumed at the snapshot's time
    assert abs(t_res - t_full) <= 1e-12 * t_full, (t_res, t_full)
    assert np.array_equal(res.ids, full.ids)
    assert np.allclose(res.r, full.r, rtol=0, atol=1e-12)
    assert np.allclose(res.v, full.v, rtol=0, atol=1e-13)


def _oracle_loop(oracle_mod, body_path, prm, dt_day, padded=False):
    """The same run through the oracle and the Python restatement of the driver loop:
    snapshots (time, state) and the events handed to the event file, in order."""
    from rednative import mainloop
    bi = generate.to_internal(files.read_body_file(body_path))
    o = oracle_mod.Oracle(bi, integrator=prm["integrator"], adaptive=prm["adaptive"], tolerance=prm["tolerance"],
                          dt_day=dt_day, hit_centrum=prm["hit_centrum"], ejection=prm["ejection"],
                          radii_enhance_factor=prm["radii_enhance_factor"],
                          encounter_hill_factor=prm["encounter_hill_factor"], padded=padded)
    o.t = prm["start_time"]
    snaps, events = [], []
    mainloop.run_main_loop(o, prm, lambda e: snaps.append((e.t, e.state())), lambda evs: events.extend(evs))
    return bi, snaps, events


@pytest.mark.gpu
def test_driver_event_file_all_columns_match_oracle(tmp_path, oracle_mod):
    """Every one of the 50 columns of every event-file line (pp_disk.cu:1786-1837: name, t, d,
    id1, id2, then mass, density, radius, position, velocity and six orbital elements of the
    survivor before, the merger before and the survivor after) equals the oracle's record of
    the same run, printed in the reference's format."""
    b = generate.disk_allsi(150, seed=17, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    files.write_body_file(str(tmp_path / "disk.txt"), b)
    (tmp_path / "p.txt").write_text("integrator = rk4\nlength = 0.2\noutput_interval = 0.05\n"
                                    "collision_factor = 300\nejection = 50\ndt = 0.5\n")
    print("case driver event file columns", flush=True)
    r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "disk.txt"])
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in open(tmp_path / "event.txt") if ln.strip()]
    prm = files.read_parameters(str(tmp_path / "p.txt"))
    _, _, events = _oracle_loop(oracle_mod, str(tmp_path / "disk.txt"), prm, prm["dt"])
    assert len(lines) == len(events) > 0
    for k, (line, e) in enumerate(zip(lines, events)):
        want = files.format_event_line(e).split()
        assert len(line) == len(want) == 50, (k, len(line))
        assert line == want, (k, [(c, a, w) for c, (a, w) in enumerate(zip(line, want)) if a != w][:4])


@pytest.mark.gpu
def test_driver_binary_result_matches_text(tmp_path):
    """output_type = binary (pp_disk.cu:1727-1770): per snapshot the time in days, then per
    active body id, name (NUL-terminated), type, mass, radius, density, cd, migration type
    and stop, position and velocity (AU/day); the same run in text holds the same values."""
    import struct
    b = generate.disk_allsi(150, seed=17, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    print("case driver binary result", flush=True)
    out = {}
    for kind in ("text", "binary"):
        d = tmp_path / kind
        d.mkdir()
        files.write_body_file(str(d / "disk.txt"), b)
        (d / "p.txt").write_text(f"integrator = rk4\nlength = 0.2\noutput_interval = 0.05\ncollision_factor = 300\n"
                                 f"dt = 0.5\noutput_type = {kind}\n")
        r = run_red(["-iDir", str(d), "-p", "p.txt", "-ic", "disk.txt"])
        assert r.returncode == 0, r.stderr
        out[kind] = d
    rows = files.parse_result_ascii(str(out["text"] / "result.txt"))
    raw = (out["binary"] / "result.bin").read_bytes()
    # group the text rows into snapshots (consecutive rows sharing the time column)
    snaps = []
    for row in rows:
        if snaps and snaps[-1][0] == row[3]:
            snaps[-1][1].append(row)
        else:
            snaps.append((row[3], [row]))
    assert len(snaps) >= 5 and len(snaps[-1][1]) < len(snaps[0][1])   # collisions removed bodies
    pos = 0

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from("<" + fmt, raw, pos)
        pos += struct.calcsize("<" + fmt)
        return v

    for t, recs in snaps:
        (tb,) = take("d")
        assert tb == float(t)
        for row in recs:
            (bid,) = take("i")
            end = raw.index(b"\0", pos)
            name = raw[pos:end].decode()
            pos = end + 1
            (btype,) = take("i")
            p4 = take("4d")
            (mig,) = take("i")
            (mig_stop,) = take("d")
            rv = take("6d")
            assert (bid, name, btype, mig) == (int(row[0]), row[1], int(row[2]), int(row[8]))
            assert list(p4) == [float(x) for x in row[4:8]] and mig_stop == float(row[9])
            assert list(rv) == [float(x) for x in row[10:16]]
    assert pos == len(raw)


@pytest.mark.gpu
def test_driver_reference_rkf78_input(tmp_path, oracle_mod, native_mod):
    """The reference's own RKF7(8) input (TestRun/InputTest/Release/N1_massive_N3_test: star,
    10 protoplanets, 10 test particles, tol 1e-10, collision_factor 100, hit_centrum 0.1 AU,
    ejection 100 AU).  Its test particle testp0 starts exactly on protoplanet proto0 (same
    position, velocity differing by 1e-11 AU/day), so their pair term is 0/0 and the test
    particle's state is NaN after the first step, which still passes: test particles are
    outside the error norm (error_check_for_tp = false).  The collision pass merges it into
    proto0 with the NaN stage velocity recorded in the event, and the second step's error
    vector holds NaN.  The oracle passes NaN components over as the reference's CPU loop does
    (integrator.cu:237-253): estimate 0, an infinite next trial step (no clamp,
    int_rungekutta8.cu:776-777), t = inf after the third step.  The library follows the first
    step and the merge bit for bit, NaN bit patterns included, and reports the second step's
    NaN estimate as RED_ERR_NAN; the drop-in driver writes the same event line (all columns)
    and initial snapshot, then stops with that error."""
    src = os.path.join(GOLDEN, "N1_massive_N3_test")
    for f in ("N1_massive_N3_test.txt", "parameters.txt"):
        (tmp_path / f).write_text(open(os.path.join(src, f)).read())
    prm = files.read_parameters(str(tmp_path / "parameters.txt"))
    assert prm["integrator"] == 4 and prm["tolerance"] == 1e-10 and prm["radii_enhance_factor"] == 100.0
    print("case reference rkf78 input", flush=True)
    bi = generate.to_internal(files.read_body_file(str(tmp_path / "N1_massive_N3_test.txt")))
    assert np.array_equal(bi.r[1], bi.r[11]) and bi.types[11] == 6
    kw = dict(integrator=4, adaptive=True, tolerance=1e-10, dt_day=0.001, hit_centrum=0.1, ejection=100.0,
              radii_enhance_factor=100.0)
    o = oracle_mod.Oracle(bi, **kw)
    d = native_mod.Disk(bi, **kw)
    assert o.check_for_ejection_hit_centrum() == d.check_for_ejection_hit_centrum() == False
    assert o.step() == d.step()
    assert o.last_err() == d.last_step.err and o.dt_try == d.dt_try
    so, sg = o.state(), d.state()
    assert np.isnan(so["v"][11]).all() and np.isfinite(so["v"][:11]).all()
    for k in ("r", "v"):
        assert np.array_equal(so[k].view(np.uint64), sg[k].view(np.uint64)), k
    assert o.check_for_collision() and d.check_for_collision()
    assert [(e["id1"], e["id2"]) for e in d.sp_events()] == [(2, 12)]
    assert files.format_event_line(o.sp_events()[0]) == files.format_event_line(d.sp_events()[0])
    so, sg = o.state(), d.state()
    for k in ("r", "v", "p"):
        assert np.array_equal(so[k].view(np.uint64), sg[k].view(np.uint64)), k
    assert np.isnan(sg["v"][1]).all()
    o.clear_event_counter(); d.clear_event_counter()
    o.step()
    assert o.last_err() == 0.0 and o.dt_try == float("inf")
    with pytest.raises(native_mod.RedError) as ex:
        d.step()
    assert ex.value.code == 4
    d.close()
    # the drop-in driver on the same files
    r = run_red(["-iDir", str(tmp_path), "-p", "parameters.txt", "-ic", "N1_massive_N3_test.txt"])
    assert r.returncode != 0 and "non-finite error estimate" in r.stderr
    _, snaps, events = _oracle_loop(oracle_mod, str(tmp_path / "N1_massive_N3_test.txt"), prm, 0.001)
    lines = [ln.split() for ln in open(tmp_path / "event.txt") if ln.strip()]
    assert len(events) == 1 and lines == [files.format_event_line(events[0]).split()]
    t0, s0 = snaps[0]
    rows = files.parse_result_ascii(str(tmp_path / "result.txt"))
    assert len(rows) == bi.n
    for i, row in enumerate(rows):
        want = files.format_result_line(s0["id"][i], row[1], s0["type"][i], t0 / GAUSS, s0["p"][i], 0, 0.0,
                                        s0["r"][i], s0["v"][i] * GAUSS).split()
        assert row == want, (i, row, want)


@pytest.mark.gpu
def test_driver_padded_storage_matches_oracle(tmp_path, oracle_mod):
    """-ups: the outer bodies, beyond the ejection distance, are ejected at once and keep
    pulling and moving under the padded kernel's rules until the rebuild after the first step
    (no collisions: two merged bodies parked at one point would meet, a 0/0); every event line
    (all 50 columns) and the last snapshot equal the oracle's padded run, and differ from the
    run without -ups."""
    b = generate.disk_allsi(150, seed=17, a_range=(1.0, 1.05), m_range=(1e-10, 1e-8))
    files.write_body_file(str(tmp_path / "disk.txt"), b)
    (tmp_path / "p.txt").write_text("integrator = rk4\nlength = 0.2\noutput_interval = 0.1\n"
                                    "ejection = 1.04\ndt = 0.5\n")
    print("case driver padded storage", flush=True)
    out = {}
    for mode, extra in (("ups", ["-ups"]), ("plain", [])):
        r = run_red(["-iDir", str(tmp_path), "-p", "p.txt", "-ic", "disk.txt", "-result", f"result_{mode}",
                     "-event", f"event_{mode}", "-info", f"info_{mode}", "-log", f"log_{mode}"] + extra)
        assert r.returncode == 0, r.stderr
        out[mode] = ((tmp_path / f"result_{mode}.txt").read_text(), (tmp_path / f"event_{mode}.txt").read_text())
    assert out["ups"][0] != out["plain"][0]
    prm = files.read_parameters(str(tmp_path / "p.txt"))
    _, snaps, events = _oracle_loop(oracle_mod, str(tmp_path / "disk.txt"), prm, prm["dt"], padded=True)
    lines = [ln.split() for ln in out["ups"][1].splitlines() if ln.strip()]
    assert len(lines) == len(events) > 0
    assert sum(1 for e in events if e["name"] == 2) > 0   # ejections
    for k, (line, e) in enumerate(zip(lines, events)):
        assert line == files.format_event_line(e).split(), k
    t, s = snaps[-1]
    keep = s["id"] > 0
    rows = [ln.split() for ln in out["ups"][0].splitlines() if ln.strip()][-int(keep.sum()):]
    for row, i, rr in zip(rows, s["id"][keep], s["r"][keep]):
        assert int(row[0]) == i
        assert [float(x) for x in row[10:13]] == [float(f"{x:.16e}") for x in rr]
