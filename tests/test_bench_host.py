"""Host-side checks of bench.py: the counter figures its line reads from profiles/ for the tunings the
autotuner picks, and its refusal of more than one GPU.  No GPU work."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALGORITHMIC = 40 * 7_420_808  # bytes per step of the 256^3, R = 120 droplet (read and write 5 FP32)


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


@pytest.mark.parametrize("workload, variant, planes, steps", [
    ("droplet_l1l2l3_256", 1, 18, 1),   # the headline line's tuning
    ("droplet_l1l2l3_256", 1, 12, 1),
    ("droplet_l1_256", "x2", 10, 2),    # the two-step kernel, two steps per launch
    ("droplet_l1_256", "x2", 8, 2),
    ("droplet_l1_256", "x2", 12, 2),
    ("droplet_l1_256", "x2", 16, 2),
])
def test_profile_figures_cover_the_picked_tunings(bench, workload, variant, planes, steps):
    """Each tuning has a traffic figure and a VALU-busy fraction, both from a committed summary
    under profiles/.  A launch moves at least one step's algorithmic bytes (the two-step kernel
    reads and writes the field once per pair of steps) and at most twice its steps' worth
    (re-reads of halos and lead-in planes)."""
    traffic, src = bench.measured_traffic(workload, variant, planes)
    assert src is not None and src.startswith("profiles/") and os.path.isfile(os.path.join(ROOT, src))
    assert ALGORITHMIC <= traffic <= 2 * steps * ALGORITHMIC
    busy, vsrc = bench.profiled("bench-valu:", "busy", workload, variant, planes)
    assert vsrc is not None and 0.0 < busy <= 1.0


def test_unprofiled_tuning_reads_as_missing(bench):
    assert bench.measured_traffic("droplet_l1l2l3_256", 1, 7) == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
