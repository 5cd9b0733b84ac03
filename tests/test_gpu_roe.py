"""Shallow water with the Roe flux (the ensemble workload's flux) on the HIP
path against the oracle: the right-hand side for p = 0..5 on periodic and
walled meshes, SSP-RK3 runs, and every instance of an ensemble stepped in
one launch per stage.  The oracle's Roe flux is pinned by the Roe property
and the stationary hydraulic jump (tests/test_oracle_physics.py)."""
import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.ensemble import Ensemble
from tests.helpers import make_pair, rel_l2, unit_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12
PRM = (1.0, 0.1, 0.4, 0.55, 0.15, 1.0)


def _pair(p, mesh, integrator=L.SSPRK3, params=PRM, g=9.81):
    return make_pair(L.SWE, p, mesh, integrator=integrator, flux=L.FLUX_ROE, g=g,
                     problem=L.PROB_SWE_BUMP, params=params, cfl=0.3)


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("periodic", [True, False])
def test_roe_eval_rhs_parity(p, periodic):
    mesh = unit_mesh(6, periodic=periodic)
    s, o = _pair(p, mesh)
    s.project_ic()
    o.project_ic()
    u0 = o.coeffs()
    # a moving, perturbed state: every mode and both wave families
    u = u0 * (1.0 + 0.05 * np.random.default_rng(p).standard_normal(u0.shape))
    u[1] += 0.3 * u[0]
    u[2] -= 0.2 * u[0]
    k_gpu = s.eval_rhs(u, t=0.0, dt=0.01)
    k_ref = o.eval_rhs(u, t=0.0, dt=0.01)
    r = rel_l2(k_gpu, k_ref)
    print("roe rhs p", p, "periodic", periodic, "rel", r)
    assert r < TOL


@pytest.mark.parametrize("p,periodic", [(1, False), (2, False), (3, False), (2, True)])
def test_roe_steps_parity(p, periodic):
    """15 CFL steps of SSP-RK3 (the graphs of 8, 4, 2 and 1 steps in turn)"""
    mesh = unit_mesh(10, periodic=periodic)
    s, o = _pair(p, mesh)
    s.project_ic()
    o.project_ic()
    ts, dts = s.step(15)
    rc, to, dto = o.run(15)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * to and abs(dts - dto) <= 1e-13 * dto
    r = rel_l2(s.coeffs(), o.coeffs())
    print("roe steps p", p, "periodic", periodic, "rel", r)
    assert r < TOL


def test_ensemble_roe_instances_match_the_oracle():
    """six walled SWE bumps (p = 1, 2, 3, sizes 50-288 elements), Roe flux,
    SSP-RK3 with each instance's CFL dt: after 20 steps every instance's
    coefficients and time are the oracle's"""
    import oracle as O
    rng = np.random.default_rng(21)
    inst = []
    for n, p in ((6, 1), (9, 2), (12, 3), (7, 1), (10, 2), (5, 3)):
        prm = (1.0, 0.1, rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.05, 0.2), 1.0)
        inst.append((unit_mesh(n, periodic=False, seed=n), p, prm))
    E = Ensemble(physics=L.SWE, integrator=L.SSPRK3, flux=L.FLUX_ROE, g=9.81,
                 problem=L.PROB_SWE_BUMP)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(0)
    E.step(20)
    for i, (mesh, p, prm) in enumerate(inst):
        o = O.Oracle(O.SWE, p, integrator=4, flux=3, cfl=0.5, g=9.81, problem=L.PROB_SWE_BUMP,
                     problem_params=prm, posfix=True)
        o.set_mesh(mesh).project_ic()
        rc, to, dto = o.run(20)
        te, dte, n = E.state(i)
        r = rel_l2(E.coeffs(i), o.coeffs())
        print("ensemble instance", i, "K", mesh.K, "p", p, "rel vs oracle", r)
        assert rc == 0 and n == 20
        assert abs(te - to) <= 1e-13 * to and abs(dte - dto) <= 1e-13 * dto
        assert r < TOL
    E.close()
