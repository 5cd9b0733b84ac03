"""User boundary states (dg_boundary_points / dg_set_boundary_states /
dg_get_boundary_states): the reference's U_inflow / U_outflow / U_reflection
plugin hooks as data.  A handle whose problem has no boundary data of its own
(a constant state) runs exactly as the built-in supersonic vortex once it is
given that problem's boundary states; clearing them restores its own."""
import os

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd import mesh as M
from dgcuda_amd.solver import Solver

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _supersonic_mesh():
    return M.read_pmsh(os.path.join(GOLDEN, "supersonic_A.pmsh"))


@pytest.mark.parametrize("p", [1, 2, 4])
def test_user_states_reproduce_the_builtin_problem(p):
    m = _supersonic_mesh()
    a = Solver(L.EULER, p, integrator=L.RK4, problem=L.PROB_EULER_SUPERSONIC, problem_params=())
    a.set_mesh(m).project_ic()
    side, code, x, y = a.boundary_points()
    nb = len(side)
    assert nb == int(np.sum(m.right < 0)) and x.shape == (nb, p + 1)
    assert sorted(set(code.tolist())) == sorted(set(m.right[m.right < 0].tolist()))
    assert np.array_equal(code, m.right[side])
    # the points lie on their sides (between the side's end points)
    x1, y1, x2, y2 = (m.side_xy[i][side][:, None] for i in range(4))
    cross = (x2 - x1) * (y - y1) - (y2 - y1) * (x - x1)
    assert np.abs(cross).max() < 1e-12
    ub = a.boundary_states(0.0)
    # walls: the supersonic problem mirrors about the circle's normal (x, y) / r
    w = code == -1
    r = np.hypot(x[w], y[w])
    assert np.allclose(np.abs(ub[w, :, 0]), np.abs(x[w] / r), atol=1e-14)
    b = Solver(L.EULER, p, integrator=L.RK4, problem=L.PROB_EULER_CONST,
               problem_params=(1.0, 0.5, 0.0, 1.0))
    b.set_mesh(m)
    b.set_coeffs(a.coeffs())
    b.set_boundary_states(ub)
    np.testing.assert_array_equal(b.boundary_states(0.0), ub)
    ta, _ = a.step(40)
    tb, _ = b.step(40)
    assert ta == tb
    np.testing.assert_array_equal(a.coeffs(), b.coeffs())
    ka = a.eval_rhs(a.coeffs(), t=0.3, dt=1e-3)
    kb = b.eval_rhs(a.coeffs(), t=0.3, dt=1e-3)
    np.testing.assert_array_equal(ka, kb)
    # cleared: the constant state's own boundary data again
    b.set_boundary_states(None)
    assert not np.array_equal(b.boundary_states(0.0), ub)
    assert not np.array_equal(b.eval_rhs(a.coeffs(), t=0.3, dt=1e-3), ka)
    a.close()
    b.close()


def test_user_states_on_the_device_and_free_stream():
    """a free stream through inflow and outflow sides given as device data
    stays a free stream"""
    import torch
    m = M.structured(8, periodic=False, jitter=0.2, seed=3, bc=(-3, -2, -2, -3))
    U0 = np.array([1.2, 0.6, 0.3, 2.5])  # rho, rho u, rho v, E
    s = Solver(L.EULER, 2, integrator=L.SSPRK3, cfl=0.3, problem=L.PROB_CONST,
               problem_params=tuple(U0))
    s.set_mesh(m).project_ic()
    side, code, x, y = s.boundary_points()
    ub = np.broadcast_to(U0, (len(side), 3, 4)).copy()
    s.set_boundary_states(torch.from_numpy(ub).cuda())
    c0 = s.coeffs()
    s.step(20)
    assert np.max(np.abs(s.coeffs() - c0)) < 1e-12
    s.close()
