"""The C ABI's input checks on a live handle: malformed meshes are rejected
on the host with DG_ERR_ARG and a message (no kernel ever sees them), and
the handle stays usable afterwards."""
import dataclasses

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.solver import DgError, Solver
from tests.helpers import CASES, unit_mesh

pytestmark = pytest.mark.gpu


def _copy(mesh, **kw):
    fields = {f.name: np.array(getattr(mesh, f.name), copy=True)
              for f in dataclasses.fields(mesh)}
    fields.update(kw)
    return type(mesh)(**fields)


def _with_extra_side(mesh, left, right, lsn, rsn):
    """the mesh plus one more side that no element lists"""
    xy = mesh.side_xy[:, :1]
    return _copy(mesh, side_xy=np.concatenate([mesh.side_xy, xy], axis=1),
                 left=np.append(mesh.left, left).astype(np.int32),
                 right=np.append(mesh.right, right).astype(np.int32),
                 lsn=np.append(mesh.lsn, lsn).astype(np.int32),
                 rsn=np.append(mesh.rsn, rsn).astype(np.int32))


def _bad_meshes():
    m = unit_mesh(4, periodic=False)
    out = {}
    out["left out of range"] = _copy(m, left=np.where(np.arange(m.S) == 3, m.K, m.left).astype(np.int32))
    out["boundary code"] = _copy(m, right=np.where(m.right < 0, -4, m.right).astype(np.int32))
    out["local side number"] = _copy(m, lsn=np.where(np.arange(m.S) == 0, 3, m.lsn).astype(np.int32))
    es = m.elem_sides.copy()
    es[1, 0] = m.S
    out["element side id"] = _copy(m, elem_sides=es)
    es = m.elem_sides.copy()
    far = int(np.nonzero((m.left != 0) & (m.right != 0))[0][0])
    es[0, 0] = far
    out["side not touching its element"] = _copy(m, elem_sides=es)
    V = m.V.copy()
    V[[2, 3, 4, 5]] = V[[4, 5, 2, 3]]  # every element clockwise
    out["clockwise"] = _copy(m, V=V)
    out["orphan boundary side"] = _with_extra_side(m, 0, -1, 0, 0)
    out["interior side listed by neither element"] = _with_extra_side(m, 0, 1, 0, 0)
    return out


@pytest.mark.parametrize("case", sorted(_bad_meshes()))
def test_malformed_mesh_rejected(case):
    bad = _bad_meshes()[case]
    prob, prm, fl = CASES[L.MAXWELL_TE]
    s = Solver(L.MAXWELL_TE, 2, integrator=L.RK4, flux=fl, problem=prob, problem_params=prm)
    with pytest.raises(DgError) as ei:
        s.set_mesh(bad)
    assert ei.value.code == L.DG_ERR_ARG
    assert "dg_set_mesh" in str(ei.value)
    # the handle takes a good mesh afterwards and steps
    good = unit_mesh(4, periodic=False)
    s.set_mesh(good).project_ic()
    t, dt = s.step(3)
    assert t > 0 and dt > 0 and np.all(np.isfinite(s.coeffs()))
    s.close()


@pytest.mark.parametrize("case", ["clockwise", "orphan boundary side", "left out of range",
                                  "side not touching its element"])
def test_rejected_mesh_keeps_the_previous_one(case):
    """a mesh rejected on a handle that already has one leaves that mesh, its
    state and its buffers in place: validation runs before anything is freed,
    so no later call can launch on freed buffers"""
    bad = _bad_meshes()[case]
    good = unit_mesh(4, periodic=False)
    prob, prm, fl = CASES[L.EULER]
    s = Solver(L.EULER, 2, integrator=L.SSPRK3, flux=fl, problem=prob, problem_params=prm)
    s.set_mesh(good).project_ic()
    c0 = s.coeffs()
    with pytest.raises(DgError) as ei:
        s.set_mesh(bad)
    assert ei.value.code == L.DG_ERR_ARG
    assert s.info().K == good.K and s.info().S == good.S
    np.testing.assert_array_equal(s.coeffs(), c0)
    t, dt = s.step(2)
    assert t > 0 and np.all(np.isfinite(s.coeffs()))
    # the same steps on a fresh handle give the same state
    r = Solver(L.EULER, 2, integrator=L.SSPRK3, flux=fl, problem=prob, problem_params=prm)
    r.set_mesh(good).project_ic()
    r.step(2)
    np.testing.assert_array_equal(s.coeffs(), r.coeffs())
    s.close()
    r.close()
