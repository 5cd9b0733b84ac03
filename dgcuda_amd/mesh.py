"""Triangle meshes in the reference's .pmsh form.

A Mesh holds exactly the arrays read_mesh produces (src/2d/conservation/
main.cu:205-277): per element the three vertices V1..V3 (counter-clockwise)
and the ids of its sides 0 (V1V2), 1 (V2V3), 2 (V3V1); per side its two end
points (taken from the left element), the left and right elements and their
local side numbers.  right < 0 marks a boundary: -1 wall, -2 outflow,
-3 inflow.  Boundary sides come first, as conservation/genmesh.py:239-278
orders them.

structured() builds the synthetic meshes of the benchmark configurations
(n x n quads, each split into two triangles, optional seeded jitter of the
interior vertices, periodic or bounded); read_pmsh / write_pmsh are the
on-disk boundary (write format of genmesh.py:286-313, %.015lf).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Mesh:
    V: np.ndarray           # float64 [6, K]: V1x, V1y, V2x, V2y, V3x, V3y
    elem_sides: np.ndarray  # int32 [3, K]: side id of local side 0, 1, 2
    side_xy: np.ndarray     # float64 [4, S]: x1, y1, x2, y2
    left: np.ndarray        # int32 [S]
    right: np.ndarray       # int32 [S]
    lsn: np.ndarray         # int32 [S]
    rsn: np.ndarray         # int32 [S]

    @property
    def K(self):
        return self.V.shape[1]

    @property
    def S(self):
        return self.left.shape[0]

    def areas(self):
        V = self.V
        return 0.5 * ((V[2] - V[0]) * (V[5] - V[1]) - (V[4] - V[0]) * (V[3] - V[1]))

    def centroids(self):
        V = self.V
        return (V[0] + V[2] + V[4]) / 3.0, (V[1] + V[3] + V[5]) / 3.0

    def check(self):
        """Consistency checks of the connectivity (raises ValueError)."""
        K, S = self.K, self.S
        if np.any(self.areas() <= 0):
            raise ValueError("mesh has clockwise or degenerate elements")
        es = self.elem_sides
        if es.min() < 0 or es.max() >= S:
            raise ValueError("element side id out of range")
        if self.left.min() < 0 or self.left.max() >= K or self.right.max() >= K:
            raise ValueError("side element id out of range")
        # every side is referenced by its left element and, when interior,
        # by its right element, at the recorded local side numbers
        e = np.arange(K)
        for k in range(3):
            s = es[k]
            is_left = self.left[s] == e
            is_right = self.right[s] == e
            if not np.all(is_left | is_right):
                raise ValueError("element lists a side that does not touch it")
            loc = np.where(is_left, self.lsn[s], self.rsn[s])
            if not np.all(loc == k):
                raise ValueError("local side numbers disagree with the element side order")
        return True


def _sort_boundary_first(left, right, lsn, rsn, side_xy):
    key = np.where(right < 0, -right, 4)
    order = np.argsort(key, kind="stable")
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    return order, inv


def structured(nx, ny=None, lx=1.0, ly=None, x0=0.0, y0=0.0, periodic=True, jitter=0.0,
               seed=0, bc=(-1, -1, -1, -1)):
    """nx x ny quads on [x0, x0+lx] x [y0, y0+ly], each split into two CCW
    triangles (a, b, c) and (a, c, d) with a = (i, j), b = (i+1, j),
    c = (i+1, j+1), d = (i, j+1).  Interior vertices move by a seeded uniform
    offset in [-jitter*h, jitter*h] per coordinate (boundary and periodic
    image vertices stay fixed).  periodic: opposite edges are joined;
    otherwise bc = (bottom, right, top, left) boundary codes."""
    ny = nx if ny is None else ny
    ly = lx if ly is None else ly
    hx, hy = lx / nx, ly / ny
    X = x0 + hx * np.arange(nx + 1, dtype=np.float64)
    Y = y0 + hy * np.arange(ny + 1, dtype=np.float64)
    XV, YV = np.meshgrid(X, Y, indexing="ij")  # [nx+1, ny+1]
    if jitter > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=(2, nx - 1, ny - 1))
        XV[1:-1, 1:-1] += d[0] * hx
        YV[1:-1, 1:-1] += d[1] * hy
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    I = I.T.ravel()  # quad q = j*nx + i (row-major, i fastest)
    J = J.T.ravel()
    nq = nx * ny
    ax, ay = XV[I, J], YV[I, J]
    bx, by = XV[I + 1, J], YV[I + 1, J]
    cx, cy = XV[I + 1, J + 1], YV[I + 1, J + 1]
    dx, dy = XV[I, J + 1], YV[I, J + 1]
    K = 2 * nq
    V = np.empty((6, K))
    V[:, 0::2] = [ax, ay, bx, by, cx, cy]  # T0 = (a, b, c)
    V[:, 1::2] = [ax, ay, cx, cy, dx, dy]  # T1 = (a, c, d)
    q = np.arange(nq)
    t0, t1 = 2 * q, 2 * q + 1
    # owned sides per quad: bottom B (a->b), diagonal D (c->a), left L (d->a)
    sB, sD, sL = 3 * q, 3 * q + 1, 3 * q + 2
    S = 3 * nq
    extra_top = extra_right = None
    if not periodic:
        extra_top = S + np.arange(nx)          # top edge of row ny-1 (c->d)
        extra_right = S + nx + np.arange(ny)   # right edge of column nx-1 (b->c)
        S += nx + ny
    left = np.empty(S, np.int32)
    right = np.empty(S, np.int32)
    lsn = np.zeros(S, np.int32)
    rsn = np.zeros(S, np.int32)
    sxy = np.empty((4, S))
    # diagonal: T0 side 2 (c->a) / T1 side 0 (a->c)
    left[sD], lsn[sD], right[sD], rsn[sD] = t0, 2, t1, 0
    sxy[:, sD] = [cx, cy, ax, ay]
    # bottom: T0 side 0 (a->b); right neighbour T1 of the quad below, its side 1
    left[sB], lsn[sB] = t0, 0
    sxy[:, sB] = [ax, ay, bx, by]
    below = ((J - 1) % ny) * nx + I
    right[sB], rsn[sB] = 2 * below + 1, 1
    # left: T1 side 2 (d->a); right neighbour T0 of the quad to the left, its side 1
    left[sL], lsn[sL] = t1, 2
    sxy[:, sL] = [dx, dy, ax, ay]
    leftq = J * nx + (I - 1) % nx
    right[sL], rsn[sL] = 2 * leftq, 1
    es = np.empty((3, K), np.int32)
    es[0, 0::2], es[2, 0::2] = sB, sD
    es[0, 1::2], es[2, 1::2] = sD, sL
    # local side 1 of T0 is the left edge of the quad to the right; of T1 the
    # bottom edge of the quad above
    es[1, 0::2] = 3 * (J * nx + (I + 1) % nx) + 2
    es[1, 1::2] = 3 * (((J + 1) % ny) * nx + I)
    if not periodic:
        bot, rgt, top, lft = bc
        right[sB[J == 0]] = bot
        rsn[sB[J == 0]] = 0
        right[sL[I == 0]] = lft
        rsn[sL[I == 0]] = 0
        # extra sides
        ii = np.arange(nx)
        qt = (ny - 1) * nx + ii
        left[extra_top], lsn[extra_top], right[extra_top] = 2 * qt + 1, 1, top
        sxy[:, extra_top] = [XV[ii + 1, ny], YV[ii + 1, ny], XV[ii, ny], YV[ii, ny]]
        jj = np.arange(ny)
        qr = jj * nx + (nx - 1)
        left[extra_right], lsn[extra_right], right[extra_right] = 2 * qr, 1, rgt
        sxy[:, extra_right] = [XV[nx, jj], YV[nx, jj], XV[nx, jj + 1], YV[nx, jj + 1]]
        es[1, 0::2][I == nx - 1] = extra_right[J[I == nx - 1]]
        es[1, 1::2][J == ny - 1] = extra_top[I[J == ny - 1]]
        order, inv = _sort_boundary_first(left, right, lsn, rsn, sxy)
        left, right, lsn, rsn = left[order], right[order], lsn[order], rsn[order]
        sxy = sxy[:, order]
        es = inv[es].astype(np.int32)
    return Mesh(V=np.ascontiguousarray(V), elem_sides=np.ascontiguousarray(es),
                side_xy=np.ascontiguousarray(sxy), left=left, right=right, lsn=lsn, rsn=rsn)


def read_pmsh(path):
    """read_mesh of main.cu:205-277."""
    with open(path) as f:
        K = int(f.readline().split()[0])
        el = np.loadtxt(f, max_rows=K, ndmin=2)
        S = int(f.readline().split()[0])
        sd = np.loadtxt(f, max_rows=S, ndmin=2)
    if el.shape != (K, 9) or sd.shape != (S, 8):
        raise ValueError("%s: not enough items" % path)
    return Mesh(V=np.ascontiguousarray(el[:, :6].T), elem_sides=np.ascontiguousarray(el[:, 6:9].T.astype(np.int32)),
                side_xy=np.ascontiguousarray(sd[:, :4].T), left=sd[:, 4].astype(np.int32),
                right=sd[:, 5].astype(np.int32), lsn=sd[:, 6].astype(np.int32),
                rsn=sd[:, 7].astype(np.int32))


def write_pmsh(mesh, path):
    """The .pmsh grammar written by genmesh.py:286-313 (%.015lf)."""
    with open(path, "w") as f:
        f.write("%d\n" % mesh.K)
        for e in range(mesh.K):
            f.write("%.015f %.015f %.015f %.015f %.015f %.015f %d %d %d\n" % (
                tuple(mesh.V[:, e]) + tuple(int(s) for s in mesh.elem_sides[:, e])))
        f.write("%d\n" % mesh.S)
        for s in range(mesh.S):
            f.write("%.015f %.015f %.015f %.015f %d %d %d %d\n" % (
                tuple(mesh.side_xy[:, s]) + (int(mesh.left[s]), int(mesh.right[s]),
                                              int(mesh.lsn[s]), int(mesh.rsn[s]))))
