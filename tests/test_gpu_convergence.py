"""Convergence of the HIP path itself against exact solutions (the oracle
shows the same orders on the CPU at smaller sizes; the parity tests tie the
two together at 1e-12):

- the steady supersonic vortex (euler/supersonic) on the reference's nested
  meshes A -> D (180 to 11,520 triangles), p = 1..3, RK4: the L2 error of
  rho falls at order ~p + 1 (the wall takes the exact curved normal at each
  point, supersonic.cu:65-90, so straight-sided elements keep high order)
- the isentropic vortex on periodic [0,10]^2, p = 1..3
- a translating shallow-water vortex in cyclostrophic balance, LLF and Roe
- a TE eigenmode of a PEC cavity with eps = 2.5
Each prints its errors and rates; the bar is p + 0.7 on the finest pair
(p + 0.5 for the shallow-water vortex with LLF at p = 2, whose oracle
rates at 24 -> 48 are 2.59-2.81)."""
import math
import os

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd import mesh as M
from dgcuda_amd.solver import Solver
from tests.test_oracle_physics import l2_error, project, swe_vortex, te_cavity_mode

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rates(errs):
    e = np.asarray(errs)
    return np.log2(e[:-1] / e[1:])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_supersonic_vortex_meshes_A_to_D(p):
    errs = []
    for lvl, name in enumerate("ABCD"):
        m = M.read_pmsh(os.path.join(GOLDEN, "supersonic_%s.pmsh" % name))
        s = Solver(L.EULER, p, integrator=L.RK4, cfl=0.5, problem=L.PROB_EULER_SUPERSONIC,
                   problem_params=())
        s.set_mesh(m).project_ic()
        # about the same physical time on every mesh (dt ~ h / (2p + 1))
        steps = int(400 * 2 ** lvl * (2 * p + 1) / 3)
        t, _ = s.step(steps)
        e = s.l2_error(t)
        assert np.all(np.isfinite(e))
        errs.append(e[0])
        print("supersonic p", p, name, "K", m.K, "steps", steps, "t %.4f" % t, "L2 rho", e[0])
        s.close()
    r = _rates(errs)
    print("supersonic p", p, "rates", r)
    assert r[-1] >= p + 0.7, (errs, r)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_isentropic_vortex_converges(p):
    errs = []
    ns = (32, 64, 128)
    for n in ns:
        m = M.structured(n, lx=10.0, periodic=True, jitter=0.1, seed=n)
        s = Solver(L.EULER, p, integrator=L.RK4, problem=L.PROB_EULER_VORTEX,
                   problem_params=(5.0, 5.0, 5.0, 1.0, 1.0, 10.0))
        s.set_mesh(m).project_ic()
        T, dt = 1.0, 0.0025 * 32 / n / (1 if p < 3 else 2)
        nsteps = int(round(T / dt))
        t, _ = s.step(nsteps, dt=T / nsteps)
        e = s.l2_error(t)
        errs.append(e[0])
        print("vortex p", p, "n", n, "steps", nsteps, "L2 rho", e[0])
        s.close()
    r = _rates(errs)
    print("vortex p", p, "rates", r)
    assert r[-1] >= p + 0.7, (errs, r)


@pytest.mark.parametrize("flux", [L.FLUX_LLF, L.FLUX_ROE])
@pytest.mark.parametrize("p,ns", [(1, (24, 48, 96)), (2, (24, 48, 96))])
def test_swe_moving_vortex_converges(flux, p, ns):
    import oracle as O
    G, T = 1.0, 0.1
    ex = swe_vortex(G, 1.0, 0.3, 0.1, 0.5, 0.5, 0.4, 0.2)
    errs = []
    for n in ns:
        m = M.structured(n, periodic=True, jitter=0.2, seed=n)
        o = O.Oracle(O.SWE, p).set_mesh(m)  # tables and geometry of the error integral
        s = Solver(L.SWE, p, integrator=L.SSPRK3, flux=flux, cfl=0.3, g=G,
                   problem=L.PROB_CONST, problem_params=(1, 0, 0))
        s.set_mesh(m)
        s.set_coeffs(project(o, m, ex(0.0)))
        t, nst, _ = s.run(t_end=T)
        assert t == T
        e = l2_error(o, m, s.coeffs(), ex(T))
        errs.append(e)
        print("swe vortex flux", flux, "p", p, "n", n, "steps", nst, "L2", e)
        s.close()
    r = np.array([_rates([a[v] for a in errs]) for v in range(3)])
    print("swe vortex flux", flux, "p", p, "rates", r)
    bar = p + (0.5 if (p == 2 and flux == L.FLUX_LLF) else 0.7)
    assert r[:, -1].min() >= bar, (errs, r)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_te_cavity_mode_converges(p):
    import oracle as O
    eps, T = 2.5, 0.6
    ex = te_cavity_mode(eps, 1.0, math.pi, 2 * math.pi)
    errs = []
    for n in (8, 16, 32):
        m = M.structured(n, periodic=False, jitter=0.2, seed=n, bc=(-1, -1, -1, -1))
        o = O.Oracle(O.MAXWELL_TE, p).set_mesh(m)
        s = Solver(L.MAXWELL_TE, p, integrator=L.RK4, cfl=0.5, problem=L.PROB_CONST,
                   problem_params=(0, 0, 0))
        s.set_mesh(m)
        s.set_eps(np.full(m.K, eps))
        s.set_coeffs(project(o, m, ex(0.0)))
        t, nst, _ = s.run(t_end=T)
        assert t == T
        e = l2_error(o, m, s.coeffs(), ex(T))
        errs.append(e)
        print("te cavity p", p, "n", n, "steps", nst, "L2", e)
        s.close()
    r = np.array([_rates([a[v] for a in errs]) for v in range(3)])
    print("te cavity p", p, "rates", r)
    assert r[:, -1].min() >= p + 0.7, (errs, r)
