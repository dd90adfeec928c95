"""The committed HBM-traffic summary of the evolution kernel, as bench.py
reads it for the roofline's `traffic` field: its byte figure must follow from
its own counter readings and copy-calibration factors (FETCH_SIZE / fetch
factor + WRITE_SIZE / write factor, KiB), bench.py must report it for the
workload it was measured on, scale it per evaluation for another cycle count,
and report nothing for another workload. Runs without a GPU."""
import os

import pytest

import bench

PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                    "evo_persistent_kernel_traffic.txt")


def _fields():
    kv = {}
    for line in open(PATH):
        if "=" in line and not line.lstrip().startswith("#"):
            key, val = line.split("=", 1)
            kv[key.strip()] = val.strip()
    return kv


def test_bytes_follow_from_the_counter_readings():
    kv = _fields()
    kib = float(kv["fetch_kib"]) / float(kv["fetch_factor"]) + float(kv["write_kib"]) / float(kv["write_factor"])
    assert int(kv["hbm_bytes_per_launch"]) == pytest.approx(kib * 1024, abs=2)


def test_bench_reads_the_summary_for_its_workload():
    kv = _fields()
    runs, cycles, k, landmarks = int(kv["runs"]), int(kv["cycles"]), int(kv["k"]), int(kv["landmarks_total"])
    measured, basis = bench.measured_traffic(kv["kernel"], runs, cycles, k, landmarks)
    assert measured == float(kv["hbm_bytes_per_launch"])
    assert str(cycles) in basis
    scaled, basis2 = bench.measured_traffic(kv["kernel"], runs, 2 * cycles + 1, k, landmarks)
    assert scaled == pytest.approx(measured * (2 * cycles + 2) / (cycles + 1))
    assert "scaled" in basis2
    assert bench.measured_traffic(kv["kernel"], runs + 1, cycles, k, landmarks) == (None, None)
    assert bench.measured_traffic("another_kernel", runs, cycles, k, landmarks) == (None, None)
