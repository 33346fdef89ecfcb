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
