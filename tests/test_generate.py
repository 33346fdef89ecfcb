"""Benchmark workloads written in the reference's input layout (the generator's command
line): the body file reads back to the seeded disk bit for bit and the parameter file
parses with the reference's keys."""
import numpy as np
import pytest

from rednative import files, generate


@pytest.mark.parametrize("name,n", [("small", 256), ("powerlaw", 16384)])
def test_workload_files_round_trip(tmp_path, name, n):
    paths = generate.write_workload(name, str(tmp_path))
    b = files.read_body_file(paths[0])
    ref = generate.WORKLOADS[name]["make"](generate.WORKLOADS[name]["seed"])
    assert b.n == n and list(b.counts) == list(ref.counts)
    for k in ("ids", "types", "mass", "radius", "density", "r", "v"):
        assert np.array_equal(getattr(b, k), getattr(ref, k)), k
    assert b.names == ref.names
    prm = files.read_parameters(paths[1])
    assert prm["radii_enhance_factor"] == 1.0 and prm["ejection"] == 100.0
    if name == "small":
        assert prm["integrator"] == 2 and not prm["adaptive"] and prm["dt"] == 1.0
    else:
        assert prm["integrator"] == 4 and prm["adaptive"] and prm["tolerance"] == 1e-12


def test_small_workload_takes_10000_steps(tmp_path):
    # the driver loop steps while t <= stop: 10000 steps of one day need stop < 10000 days
    prm = files.read_parameters(generate.write_workload("small", str(tmp_path))[1])
    stop_day = (prm["start_time"] + prm["length"]) / files.GAUSS
    t, steps, dt = 0.0, 0, prm["dt"] * files.GAUSS
    while t <= prm["start_time"] + prm["length"]:
        t += dt
        steps += 1
    assert 9999.0 < stop_day < 10000.0 and steps == 10000


def test_ensemble_files(tmp_path):
    generate.main(["ensemble", "--disks", "3", "--out", str(tmp_path)])
    disks = generate.ensemble(n_disks=3)
    for k in range(3):
        b = files.read_body_file(str(tmp_path / f"disk_{k:03d}.txt"))
        assert np.array_equal(b.r, disks[k].r) and b.n == disks[k].n
    prm = files.read_parameters(str(tmp_path / "parameters.txt"))
    assert prm["integrator"] == 4 and prm["tolerance"] == 1e-11
