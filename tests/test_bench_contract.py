"""The benchmark's command-line contract and its host-side pieces, on the
CPU: one GPU only, the workloads it names, the event options of the plain
and the full-pipeline runs, the sequence comparison, and the host cores the
CPU baseline uses."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import tapsim_modified_amd as tsg  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workloads_cover_the_five_tips():
    for name in ("c1", "c2", "c3", "c5"):
        w = bench.WORKLOADS[name]
        assert w["radius"] > 0 and w["shank"] > 0 and len(w["box"]) == 3
    # c3 and c5 are two-phase (solute clusters), c1 and c2 single-phase
    assert bench.WORKLOADS["c3"].get("clusters", 0) > 0 and bench.WORKLOADS["c5"].get("clusters", 0) > 0
    assert bench.WORKLOADS["c1"].get("clusters", 0) == 0 and bench.WORKLOADS["c2"].get("clusters", 0) == 0


def test_event_options_plain_and_full_pipeline():
    e = bench.event_options(tsg)
    assert (e.trajectory, e.integrator, e.stepper) == (1, 1, 1)  # Cash-Karp, error-restricted
    assert e.detector_z == 0.0 and e.ion_table == 0
    assert e.relax_threshold == float(np.float32(1e-2))
    f = bench.event_options(tsg, full_pipeline=True, zmax=1e-7)
    assert f.integrator == 2  # the classical RK4 step
    assert f.detector_z == float(np.float32(1e-7 + bench.DETECTOR_DISTANCE))
    assert (f.ion_table, f.ion_seed) == (1, bench.ION_SEED)


def test_ion_tables_are_normalisable():
    for sid, (mass, charge, weight) in bench.ION_TABLES.items():
        assert len(mass) == len(charge) == len(weight) <= 8
        assert all(m > 0 for m in mass) and all(q > 0 for q in charge) and sum(weight) > 0


def test_first_difference():
    assert bench.first_difference([1, 2, 3], [1, 2, 3]) is None
    assert bench.first_difference([1, 2, 3], [1, 5, 3]) == 1
    assert bench.first_difference([1, 2], [1, 2, 3]) is None  # compared over the shorter


def test_host_cores_are_counted():
    n = bench.host_cores()
    assert 1 <= n <= len(os.sched_getaffinity(0))
    assert isinstance(bench.cpu_model(), str)
