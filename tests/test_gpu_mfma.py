"""The stage's volume integral on the FP64 matrix cores (dg_config.volume = 2,
Euler p = 4: v_mfma_f64_16x16x4f64 for the interpolation to the quadrature
points and the projection onto the test-function gradients) against the
oracle and against the vector-FMA kernel on the same inputs.

The matrix cores sum the products in another order than the vector FMAs, so
the two kernels agree to rounding (relative L2 <= 1e-14 on one evaluation),
not bit for bit; both stay within the 1e-12 parity bar of the oracle."""
import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd import mesh as M
from dgcuda_amd.solver import Solver
from tests.helpers import make_pair, rel_l2, unit_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12
VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)


@pytest.mark.parametrize("periodic", [True, False])
def test_mfma_eval_rhs_parity(periodic):
    mesh = unit_mesh(6, periodic=periodic)
    s, o = make_pair(L.EULER, 4, mesh, volume=L.VOLUME_MFMA)
    v, _ = make_pair(L.EULER, 4, mesh, volume=L.VOLUME_VALU)
    o.project_ic()
    u0 = o.coeffs()
    u = u0 * (1.0 + 0.05 * np.random.default_rng(4).standard_normal(u0.shape))
    k_m = s.eval_rhs(u, t=0.0, dt=0.01)
    k_v = v.eval_rhs(u, t=0.0, dt=0.01)
    k_ref = o.eval_rhs(u, t=0.0, dt=0.01)
    print("mfma eval_rhs periodic", periodic, "vs oracle", rel_l2(k_m, k_ref),
          "vs vector", rel_l2(k_m, k_v))
    assert rel_l2(k_m, k_ref) < TOL
    assert rel_l2(k_m, k_v) < 1e-14
    s.close()
    v.close()


def test_mfma_eval_rhs_positivity_fix():
    """every fifth element fails the point check: its modes above the mean
    are dropped on read in both kernels"""
    mesh = unit_mesh(6)
    s, o = make_pair(L.EULER, 4, mesh, volume=L.VOLUME_MFMA)
    o.project_ic()
    u = o.coeffs()
    u[:, 1:, ::5] *= 400.0
    k_m = s.eval_rhs(u, t=0.0, dt=0.01)
    k_ref = o.eval_rhs(u, t=0.0, dt=0.01)
    print("mfma positivity", rel_l2(k_m, k_ref))
    assert rel_l2(k_m, k_ref) < TOL
    s.close()


def test_mfma_steps_parity():
    """11 SSP-RK3 steps (graphs of 8, 2 and 1 steps) against the oracle"""
    mesh = unit_mesh(8)
    s, o = make_pair(L.EULER, 4, mesh, integrator=L.SSPRK3, volume=L.VOLUME_MFMA)
    s.project_ic()
    o.project_ic()
    ts, _ = s.step(11)
    rc, to, _ = o.run(11)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * abs(to)
    r = rel_l2(s.coeffs(), o.coeffs())
    print("mfma steps rel", r)
    assert r < TOL
    s.close()


def test_mfma_matches_vector_kernel_on_the_vortex():
    """64 x 64 periodic quads (8,192 triangles, 128 waves): 40 SSP-RK3 steps
    of the isentropic vortex with both kernels; the solutions agree to
    rounding and both meet the exact vortex equally"""
    mesh = M.structured(64, lx=10.0, periodic=True, jitter=0.2, seed=20261202)
    out = {}
    for vol in (L.VOLUME_VALU, L.VOLUME_MFMA):
        s = Solver(L.EULER, 4, integrator=L.SSPRK3, cfl=0.2, problem=L.PROB_EULER_VORTEX,
                   problem_params=VORTEX, volume=vol)
        s.set_mesh(mesh).project_ic()
        t, _ = s.step(40)
        out[vol] = (t, s.coeffs(), s.l2_error(t)[0], s.info().vgprs_stage)
        s.close()
    (tv, cv, ev, gv), (tm, cm, em, gm) = out[L.VOLUME_VALU], out[L.VOLUME_MFMA]
    print("vortex: t", tv, tm, "rel", rel_l2(cm, cv), "L2 rho error", ev, em, "vgprs", gv, gm)
    assert abs(tm - tv) <= 1e-13 * tv
    assert rel_l2(cm, cv) < 1e-12
    assert abs(em - ev) <= 1e-6 * ev
