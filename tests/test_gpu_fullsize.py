"""The two plain-method workloads at their stated sizes, HIP path against
the oracle (run on every host core) and against the exact solution.

- linear advection u_t + (1,1).grad u = 0 on the periodic unit square,
  64 x 64 quads split into 8,192 triangles (no jitter), p = 1, upwind flux,
  SSP-RK2, CFL 0.3, 100 steps; u0 = sin(2 pi x) sin(2 pi y), exact solution
  u0(x - t, y - t); the 32 x 32 mesh gives the convergence order
- the compressible Euler isentropic vortex (beta 5, centre (5, 5), free
  stream rho = p = 1, u = (1, 1)) on periodic [0,10]^2, 512 x 512 quads
  split into 524,288 triangles with interior vertices jittered +-0.2h from
  seed 20261202, p = 2, LLF, SSP-RK3, CFL 0.2, 200 steps; error against the
  exact advected vortex
Both must match the oracle's modal coefficients to 1e-12 relative L2."""
import os
import time

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd import mesh as M
from dgcuda_amd.solver import Solver
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12
THREADS = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _advection(n, steps):
    import oracle as O
    m = M.structured(n, periodic=True, jitter=0.0)
    s = Solver(L.ADVECTION, 1, integrator=L.SSPRK2, cfl=0.3, a=(1.0, 1.0),
               problem=L.PROB_ADV_SINSIN, problem_params=())
    s.set_mesh(m).project_ic()
    t, dt = s.step(steps)
    o = O.Oracle(O.ADVECTION, 1, integrator=3, cfl=0.3, a=(1.0, 1.0), problem=0,
                 threads=THREADS)
    o.set_mesh(m).project_ic()
    rc, to, dto = o.run(steps)
    assert rc == 0
    return m, s, o, t, dt, to, dto


def test_advection_64x64_100_steps():
    m, s, o, t, dt, to, dto = _advection(64, 100)
    assert m.K == 8192 and m.S == 12288
    assert abs(t - to) <= 1e-13 * to and abs(dt - dto) <= 1e-13 * dto
    r = rel_l2(s.coeffs(), o.coeffs())
    e64 = s.l2_error(t)[0]
    print("advection 64x64 p=1 SSP-RK2 100 steps: t", t, "rel vs oracle", r, "L2 vs exact", e64)
    assert r < TOL
    assert abs(e64 - o.l2_error(to)[0]) <= 1e-10 * e64
    # the same final time on the 32 x 32 mesh (dt doubles, 50 steps)
    m2, s2, _, t2, _, _, _ = _advection(32, 50)
    assert abs(t2 - t) <= 1e-13 * t
    e32 = s2.l2_error(t2)[0]
    rate = np.log2(e32 / e64)
    print("advection L2 error 32x32", e32, "64x64", e64, "order", rate)
    assert rate > 1.8
    s.close()
    s2.close()


@pytest.mark.timeout(900)
def test_euler_vortex_524k_200_steps():
    import oracle as O
    prm = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)
    m = M.structured(512, lx=10.0, periodic=True, jitter=0.2, seed=20261202)
    assert m.K == 524288 and m.S == 786432
    s = Solver(L.EULER, 2, integrator=L.SSPRK3, cfl=0.2, problem=L.PROB_EULER_VORTEX,
               problem_params=prm)
    s.set_mesh(m).project_ic()
    t0 = time.perf_counter()
    t, dt = s.step(200)
    gpu_s = time.perf_counter() - t0
    print("HIP: 200 steps in %.3f s, t = %.9f" % (gpu_s, t), flush=True)
    o = O.Oracle(O.EULER, 2, integrator=4, cfl=0.2, problem=2, problem_params=prm,
                 threads=THREADS)
    o.set_mesh(m).project_ic()
    t0 = time.perf_counter()
    done, to = 0, 0.0
    while done < 200:  # chunks, so the run reports progress
        rc, to, nst, _, dto = o.run_until(max_steps=25, t0=to)
        assert rc == 0 and nst == 25
        done += 25
        print("oracle: %d steps, %.1f s on %d threads" % (done, time.perf_counter() - t0, THREADS),
              flush=True)
    assert abs(t - to) <= 1e-13 * to and abs(dt - dto) <= 1e-13 * dto
    r = rel_l2(s.coeffs(), o.coeffs())
    err = s.l2_error(t)
    print("Euler vortex 524,288 triangles, 200 SSP-RK3 steps: rel L2 vs oracle %.3e; "
          "L2 error vs the exact vortex (rho, rho u, rho v, E) %s" % (r, err))
    assert r < TOL
    assert np.all(np.isfinite(err)) and err[0] < 1e-5
    s.close()
