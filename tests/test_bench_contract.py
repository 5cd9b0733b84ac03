"""bench.py's contract pieces that need no GPU: the algorithmic bytes per
stage of the benchmark workload, its metric name, and the refusal of more
than one GPU."""
import os
import subprocess
import sys

import bench
from dgcuda_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_stage_bytes_of_the_benchmark_workload():
    """8 K n_p N (2 + r) + 52 K + 40 S at 524,288 triangles, p = 2, Euler,
    SSP-RK3 (r = 2/3): the 327,155,712 B per stage of the roofline"""
    mesh = M.structured(512, lx=10.0, periodic=True, jitter=0.2, seed=bench.SEED)
    assert (mesh.K, mesh.S) == (524288, 786432)
    B = bench.stage_bytes(mesh.K, mesh.S, 6, 4, 2.0 / 3.0)
    assert round(B) == 327155712
    # 50 % of the 8 TB/s peak is a stage of at most 81.8 us
    assert abs(B / (0.5 * bench.HBM_PEAK_GBS * 1e9) * 1e6 - 81.79) < 0.01


def test_metric_name():
    assert bench.METRIC == ("DOF-updates/s per RK stage (Euler p=2, 524k tris) "
                            "+ adjoint/forward time ratio")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
