"""Mesh generator and .pmsh boundary (read_mesh main.cu:205-277, genmesh.py:286-313)."""
import os

import numpy as np
import pytest

from dgcuda_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _local_side_pts(mesh, e, k):
    V = mesh.V[:, e].reshape(3, 2)
    return V[k], V[(k + 1) % 3]


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_structured_connectivity(periodic, jitter):
    n, L = 7, 2.5
    m = M.structured(n, lx=L, periodic=periodic, jitter=jitter, seed=11)
    m.check()
    assert m.K == 2 * n * n
    assert m.S == (3 * n * n if periodic else 3 * n * n + 2 * n)
    assert abs(m.areas().sum() - L * L) < 1e-12
    # interior sides: left element traverses the side as stored, right the other
    # way (up to a periodic shift)
    for s in range(m.S):
        a, b = _local_side_pts(m, m.left[s], m.lsn[s])
        assert np.array_equal(a, m.side_xy[:2, s]) and np.array_equal(b, m.side_xy[2:, s])
        if m.right[s] >= 0:
            c, d = _local_side_pts(m, m.right[s], m.rsn[s])
            shift = c - b
            assert np.allclose(d - a, shift, atol=1e-12)
            assert np.allclose(np.round(shift / L), shift / L, atol=1e-12)
    if not periodic:
        bnd = m.right < 0
        assert bnd.sum() == 4 * n
        assert np.all(np.diff(bnd.astype(int)) <= 0), "boundary sides first"


def test_jitter_seeded_and_bounded():
    a = M.structured(16, periodic=True, jitter=0.2, seed=5)
    b = M.structured(16, periodic=True, jitter=0.2, seed=5)
    c = M.structured(16, periodic=True, jitter=0.0)
    assert np.array_equal(a.V, b.V)
    h = 1.0 / 16
    assert np.max(np.abs(a.V - c.V)) <= 0.2 * h + 1e-15
    assert np.max(np.abs(a.V - c.V)) > 0.1 * h


def test_pmsh_roundtrip(tmp_path):
    m = M.structured(5, periodic=False, jitter=0.2, seed=2, bc=(-1, -2, -1, -3))
    p = tmp_path / "m.pmsh"
    M.write_pmsh(m, p)
    r = M.read_pmsh(p)
    r.check()
    assert np.allclose(r.V, m.V, atol=1e-15, rtol=0)
    for f in ("elem_sides", "left", "right", "lsn", "rsn"):
        assert np.array_equal(getattr(r, f), getattr(m, f))


@pytest.mark.parametrize("name,K,S", [("A", 180, 293), ("B", 720, 1126), ("C", 2880, 4412)])
def test_reference_supersonic_meshes(name, K, S):
    """the reference's own meshes: CCW, consistent local side numbers, boundary
    sides first (-1 wall, -2 outflow, -3 inflow)"""
    m = M.read_pmsh(os.path.join(GOLDEN, "supersonic_%s.pmsh" % name))
    assert (m.K, m.S) == (K, S)
    m.check()
    bnd = m.right < 0
    assert np.all(np.diff(bnd.astype(int)) <= 0)
    assert set(np.unique(m.right[bnd])) == {-1, -2, -3}
