"""Stored traces: in runs of steps with the four-lane stage kernels (Euler
p >= 4) every stage stores its output's traces on the element's three sides,
and the next stage reads its neighbours' traces from there instead of
folding their modes itself.  The stored traces come from the same
multiply-add chains as the fold, so a run agrees with the gather path's run
(DG_STAGE_GATHER set when the handle is created) to rounding, also where
neighbours fail the positivity check and contribute their means only.  Not
bit for bit: the two paths are separate kernel instantiations, and with
HIP's default multiply-add contraction the compiler may fuse expressions
such as the RK combination differently in each (differences of 1e-15).  A
wrong side, point order or flag would differ at 1e-3 or more."""
import os

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.solver import Solver
from tests.helpers import CASES, make_pair, rel_l2, unit_mesh

pytestmark = pytest.mark.gpu


def _solver(p, integrator, gather, problem=None, params=None):
    prob, prm, fl = CASES[L.EULER]
    if gather:
        os.environ["DG_STAGE_GATHER"] = "1"
    try:
        s = Solver(L.EULER, p, integrator=integrator, flux=fl, cfl=0.5,
                   problem=problem if problem is not None else prob,
                   problem_params=params if params is not None else prm)
    finally:
        os.environ.pop("DG_STAGE_GATHER", None)
    return s


@pytest.mark.parametrize("p,integrator,periodic", [
    (4, L.SSPRK3, True), (4, L.RK4, False), (4, L.SSPRK2, True), (5, L.SSPRK3, False)])
def test_stored_traces_equal_the_gather_path(p, integrator, periodic):
    mesh = unit_mesh(8, periodic=periodic)
    out = []
    for gather in (False, True):
        s = _solver(p, integrator, gather)
        s.set_mesh(mesh).project_ic()
        t1, _ = s.step(7)       # graphs of 4, 2 and 1 steps
        t2, dt = s.step(5, t=t1)  # a second run starts from the traces of a fresh fill
        out.append((t2, dt, s.coeffs()))
        s.close()
    (ta, da, ca), (tb, db, cb) = out
    print("p", p, "integrator", integrator, "periodic", periodic, "rel", rel_l2(ca, cb),
          "dt", da, db)
    assert abs(ta - tb) <= 1e-14 * abs(tb) and abs(da - db) <= 1e-13 * abs(db)
    assert rel_l2(ca, cb) < 1e-13


def test_stored_traces_with_flagged_neighbours():
    """every fifth element's state fails the point check: its neighbours read
    its mean only, from its coefficients, in both paths"""
    mesh = unit_mesh(6)
    out = []
    for gather in (False, True):
        s = _solver(4, L.SSPRK3, gather)
        s.set_mesh(mesh).project_ic()
        u = s.coeffs()
        u[:, 1:, ::5] *= 400.0
        s.set_coeffs(u)
        s.step(3, dt=1e-4)
        out.append(s.coeffs())
        s.close()
    print("flagged rel", rel_l2(out[0], out[1]))
    assert np.all(np.isfinite(out[0]))
    assert rel_l2(out[0], out[1]) < 1e-13


def test_stored_traces_parity_with_the_oracle():
    """20 SSP-RK3 steps of the p = 4 vortex against the oracle"""
    mesh = unit_mesh(10)
    s, o = make_pair(L.EULER, 4, mesh, integrator=L.SSPRK3)
    s.project_ic()
    o.project_ic()
    ts, _ = s.step(20)
    rc, to, _ = o.run(20)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * abs(to)
    r = rel_l2(s.coeffs(), o.coeffs())
    print("stored traces vs oracle rel", r)
    assert r < 1e-12
    s.close()
