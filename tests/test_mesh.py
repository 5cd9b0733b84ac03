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


def test_gmsh_ingestion_matches_reference_genmesh():
    """rotatinghill A: the reference's .msh and the .pmsh its genmesh wrote;
    ingestion reproduces elements, sides, codes and ordering exactly"""
    from dgcuda_amd import gmsh as G
    m = G.genmesh(os.path.join(GOLDEN, "rotatinghill_A.msh"))
    r = M.read_pmsh(os.path.join(GOLDEN, "rotatinghill_A.pmsh"))
    assert (m.K, m.S) == (r.K, r.S) == (1080, 1664)
    assert np.max(np.abs(m.V - r.V)) < 1e-15
    assert np.max(np.abs(m.side_xy - r.side_xy)) < 1e-15
    for f in ("left", "right", "lsn", "elem_sides"):
        assert np.array_equal(getattr(m, f), getattr(r, f)), f
    inner = r.right >= 0
    assert np.array_equal(m.rsn[inner], r.rsn[inner])


def test_gmsh_periodic_and_orientation(tmp_path):
    """a gmsh file written from a structured mesh: clockwise triangles are
    reoriented, opposite faces join when periodic, untagged faces are refused"""
    from dgcuda_amd import gmsh as G
    ref = M.structured(4, lx=2.0, periodic=True)
    V = ref.V
    pts = {}
    tris = []
    for e in range(ref.K):
        ids = []
        for q in range(3):
            key = (round(V[2 * q, e], 12), round(V[2 * q + 1, e], 12))
            ids.append(pts.setdefault(key, len(pts) + 1))
        if e % 3 == 0:
            ids = [ids[1], ids[0], ids[2]]  # clockwise on purpose
        tris.append(ids)
    path = tmp_path / "box.msh"
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(pts))
        for (x, y), i in sorted(pts.items(), key=lambda t: t[1]):
            f.write("%d %.17g %.17g 0\n" % (i, x, y))
        f.write("$EndNodes\n$Elements\n%d\n" % len(tris))
        for q, t in enumerate(tris):
            f.write("%d 2 2 100 6 %d %d %d\n" % (q + 1, *t))
        f.write("$EndElements\n")
    with pytest.raises(ValueError):
        G.genmesh(str(path))
    m = G.genmesh(str(path), periodic=(2.0, 2.0))
    assert m.K == ref.K and m.S == 3 * 16 and np.all(m.right >= 0)
    assert np.all(m.areas() > 0)
    w = G.genmesh(str(path), default_bc=-1)
    assert int((w.right == -1).sum()) == 16
