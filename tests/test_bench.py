"""bench.py's contract on the CPU: one GPU only, the algorithmic bytes of a 4001 x 4001 step, and
the measured traffic it reads back per step for the kernel and tile shape that ran."""
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_of_the_4001_step():
    # configuration 1's grid without its material (only the C-PML profiles matter here)
    prob = SimpleNamespace(nx=4000, nz=4000, dx=0.5, dz=0.5, dt=5.0e-5, npml=20, npower=2.0, rcoef=1.0e-3,
                           kmax=1.0, f0=80.0)
    cols, rows = bench.strip_counts(prob)
    # both strips of each axis: 20 layer points plus the reference's offsets of the left and top
    # layers (origins at L + 3 dx and L): a few points more than 2 x 20
    assert 40 < cols < 50 and 40 < rows < 50
    b = bench.algorithmic_bytes_per_step(4000, 4000, cols, rows)
    npts = 4001 * 4001
    assert b == 104.0 * npts + 48.0 * (cols + rows) * 4001
    assert b == pytest.approx(1.6819e9, rel=1e-4)


def test_recorded_traffic_is_per_time_step():
    lines = [ln.split(":", 1) for ln in open(bench.TRAFFIC_FILE) if ":" in ln and not ln.startswith("#")]
    kv = {k.strip(): float(v) for k, v in lines}
    one, src = bench.recorded_traffic((256, 8), 4000, 0)
    assert one == kv["tile 256x8"] and src.startswith("profiles")
    two, _ = bench.recorded_traffic((256, 24), 4000, 3)
    assert two == kv["two-step 256x24"] / 2  # the two-step kernel's launch covers two steps
    # another grid, or a shape that was not profiled: no figure
    assert bench.recorded_traffic((256, 8), 1200, 0) == (None, None)
    assert bench.recorded_traffic((128, 30), 4000, 3) == (None, None)
