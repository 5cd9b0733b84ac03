"""dgrun, the reference's command line (get_input / run_dgcuda,
src/2d/conservation/main.cu:481-974) over the C ABI: argument handling
without a GPU, and a full run against the Python host API on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from dgcuda_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DGRUN = os.path.join(ROOT, "dgcuda_amd", "dgrun")
MESH = os.path.join(ROOT, "tests", "golden", "supersonic_A.pmsh")


@pytest.fixture(scope="module")
def dgrun():
    if not os.path.exists(DGRUN):
        from dgcuda_amd import build
        if not os.path.exists(L.LIB_PATH):
            build.build()
        else:
            build.build_cli()
    return DGRUN


def run(args, cwd):
    return subprocess.run([DGRUN] + args, capture_output=True, text=True, cwd=cwd, timeout=120)


def test_usage(dgrun, tmp_path):
    r = run(["-h"], tmp_path)
    assert r.returncode == 1 and "Usage: dgrun" in r.stdout and "[-c] Run to convergence" in r.stdout
    for bad in (["-n", "7", "-t", "1", MESH], ["-t", "1", MESH], ["-n", "1", MESH],
                ["-n", "1", "-t", "1", "--physics", "plasma", MESH], ["-n", "1", "-t", "1"]):
        r = run(bad, tmp_path)
        assert r.returncode == 1 and "Usage" in r.stdout, bad


def test_limiter_order_check(dgrun, tmp_path):
    """main.cu:647-650: the limiter is p = 1 only (the reference exits with 0)"""
    r = run(["-n", "2", "-t", "1", "--limiter", MESH], tmp_path)
    assert r.returncode == 0 and "Error: limiter only enabled for p = 1" in r.stdout


def test_missing_mesh(dgrun, tmp_path):
    r = run(["-n", "1", "-t", "1", str(tmp_path / "nope.pmsh")], tmp_path)
    assert r.returncode == 1 and "mesh file not found" in r.stdout


def _pos_values(path):
    vals = []
    with open(path) as f:
        assert f.readline().startswith('View "U')
        for line in f:
            m = re.match(r"ST \(([^)]*)\) \{([^}]*)\};", line)
            if m:
                vals.append([float(v) for v in m.group(2).split(",")])
    return np.array(vals)


@pytest.mark.gpu
def test_cli_run_matches_host_api(dgrun, tmp_path):
    """supersonic vortex (p = 1, RK4, 20 steps) on the reference's mesh A:
    printed L2 errors, U0-final.pos, the error/exact views and the
    coefficient dump agree with the same run through the Python host API"""
    from dgcuda_amd.solver import Solver
    r = run(["-n", "1", "-t", "20", "-e", "-b", "--problem", "supersonic", "--coeffs", "c.bin",
             MESH], tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    errs = [float(v) for v in re.findall(r"L2 U\d error     = (\S+)", r.stdout)]
    assert len(errs) == 4 and "Runtime:" in r.stdout
    s = Solver(L.EULER, 1, integrator=L.RK4, problem=L.PROB_EULER_SUPERSONIC)
    s.load_pmsh(MESH).project_ic()
    t, _ = s.step(20)
    ref = s.l2_error(t)
    assert np.allclose(errs, ref, rtol=1e-5, atol=0)
    C = np.fromfile(tmp_path / "c.bin", dtype="<f8").reshape(s.shape)
    assert np.array_equal(C, s.coeffs())
    u = _pos_values(tmp_path / "output" / "U0-final.pos")
    uh = np.stack(s.eval_u(0), 1)
    assert np.allclose(u, uh, atol=2e-15, rtol=0)
    for n in range(4):
        ex = _pos_values(tmp_path / "output" / ("U%dexact.msh" % n))
        er = _pos_values(tmp_path / "output" / ("U%derror.msh" % n))
        assert np.allclose(ex, np.stack(s.eval_exact(n, t), 1), atol=2e-15, rtol=0)
        assert np.allclose(er, np.stack(s.eval_u(n), 1) - ex, atol=4e-15, rtol=0)


@pytest.mark.gpu
def test_cli_endtime_and_convergence(dgrun, tmp_path):
    r = run(["-n", "1", "-T", "0.05", "--problem", "supersonic", MESH], tmp_path)
    assert r.returncode == 0 and " ? endtime = 0.050000" in r.stdout
    assert re.search(r"t = 0\.050000", r.stdout)
    r = run(["-n", "1", "-c", "1e-4", "--problem", "supersonic", MESH], tmp_path)
    assert r.returncode == 0 and " ? convergence = 0.000100" in r.stdout
    conv = [float(v) for v in re.findall(r"convergence = ([0-9.]+)", r.stdout)[1:]]
    assert conv and conv[-1] <= 1e-4
