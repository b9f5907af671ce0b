"""bench.py's command line on a machine without a GPU: more than one GPU is refused with a message and exit
code 2 before any device work; the metric names the workload's headline."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def _ideal():
    sys.path.insert(0, ROOT)
    import bench
    return bench.ideal_server


def test_ideal_server_small_cases():
    ideal = _ideal()
    r = ideal([0.0, 0.0], [1.0, 2.0])
    assert list(r["fifo"]) == [1.0, 3.0] and list(r["srpt"]) == [1.0, 3.0] and list(r["ps"]) == [2.0, 3.0]
    # A long job, then a short one while it runs: SRPT preempts, sharing halves both rates until the
    # short one leaves.
    r = ideal([0.0, 1.0], [4.0, 1.0])
    assert list(r["fifo"]) == [4.0, 4.0]
    assert list(r["srpt"]) == [5.0, 1.0]
    assert list(r["ps"]) == [5.0, 2.0]
    r = ideal([0.0, 10.0], [1.0, 1.0])  # an idle gap
    assert all(list(v) == [1.0, 1.0] for v in r.values())


def test_ideal_server_queueing_laws():
    """M/M/1 at load 0.5: mean response E[S]/(1-rho) = 2 under FIFO and under processor sharing (where
    every size has mean slowdown 1/(1-rho) = 2); SRPT has the least mean response; every schedule is
    work-conserving, so the last job leaves at the same time."""
    import numpy as np
    ideal = _ideal()
    g = np.random.default_rng(5)
    n = 200000
    arr = np.cumsum(g.exponential(2.0, n))
    work = g.exponential(1.0, n)
    r = ideal(arr, work)
    assert abs(r["fifo"].mean() - 2.0) < 0.1
    assert abs(r["ps"].mean() - 2.0) < 0.1
    assert abs((r["ps"] / work).mean() - 2.0) < 0.1
    assert r["srpt"].mean() < r["ps"].mean() and r["srpt"].mean() < r["fifo"].mean()
    ends = [float((arr + v).max()) for v in r.values()]
    assert max(ends) - min(ends) < 1e-6 * ends[0]
