"""gmsh v2.2 ASCII -> .pmsh mesh ingestion (replaces the reference's Python 2
conservation/genmesh.py:13-315).

Same mapping rules as the reference:
  - triangles in file order are the elements; their sides are 0 = (v0, v1),
    1 = (v1, v2), 2 = (v2, v0) (genmesh.py:100)
  - a side belongs first to the element that creates it (its left element)
    and side ids follow creation order; the second element to list a side is
    its right element (genmesh.py:100-118)
  - boundary lines with physical tag 10000 / 20000 / 30000 make the side a
    wall / outflow / inflow, right = -1 / -2 / -3 (genmesh.py:166-174)
  - boundary sides come first: -1, then -2, then -3, in the order the
    reference's in-place swap loop leaves them (genmesh.py:239-278), so
    the .pmsh connectivity is the reference's side for side
Differences, on purpose:
  - clockwise triangles are reoriented (v0 <-> v1) *before* their sides are
    built, so local side numbers stay consistent (the reference swaps the
    vertices only when writing, genmesh.py:296-301)
  - the swap loop's update of the right element of a boundary side (a
    negative index, genmesh.py:255-261) is skipped; order="stable" gives a
    plain stable sort instead
  - an untagged open boundary side is an error unless `default_bc` is given
    (the reference leaves such a side's right element at 0)
  - periodic=(LX, LY): vertices on the max faces are identified with their
    images on the min faces, so opposite boundary sides become one interior
    side (its end points are the left element's own vertices)
CLI: python -m dgcuda_amd.gmsh in.msh out.pmsh [--periodic LX LY] [--default-bc CODE]
"""
import argparse
import sys

import numpy as np

from .mesh import Mesh, write_pmsh

TAG_CODE = {10000: -1, 20000: -2, 30000: -3}
_OPEN = np.iinfo(np.int32).min


def parse_msh(path):
    """nodes [n, 2], triangles [K, 3] (0-based), boundary lines [(a, b, tag)]"""
    with open(path) as f:
        lines = f.read().splitlines()

    def find(tag, start):
        for q in range(start, len(lines)):
            if lines[q].strip() == tag:
                return q
        raise ValueError("%s: no %s section" % (path, tag))

    fmt = find("$MeshFormat", 0)
    version = lines[fmt + 1].split()
    if not version[0].startswith("2") or version[1] != "0":
        raise ValueError("%s: gmsh format %s, only 2.x ASCII is read" % (path, " ".join(version)))
    i = find("$Nodes", fmt)
    nn = int(lines[i + 1])
    index = {}
    xy = np.empty((nn, 2))
    for q in range(nn):
        s = lines[i + 2 + q].split()
        index[int(s[0])] = q
        xy[q] = float(s[1]), float(s[2])
    i = find("$Elements", i)
    ne = int(lines[i + 1])
    tris, bnd = [], []
    for q in range(ne):
        s = [int(v) for v in lines[i + 2 + q].split()]
        etype, ntags = s[1], s[2]
        tags = s[3:3 + ntags]
        nodes = [index[v] for v in s[3 + ntags:]]
        if etype == 2:
            tris.append(nodes[:3])
        elif etype == 1:
            bnd.append((nodes[0], nodes[1], tags[0] if tags else 0))
    return xy, np.array(tris, np.int64).reshape(-1, 3), bnd


def _periodic_ids(xy, period, tol):
    """canonical vertex ids: a vertex on a max face maps to its image"""
    LX, LY = period
    x, y = xy[:, 0], xy[:, 1]
    x0, y0 = x.min(), y.min()
    h = tol * max(LX, LY)
    key = {(int(round(a / h)), int(round(b / h))): q for q, (a, b) in enumerate(xy)}
    cid = np.arange(len(xy))
    for q in range(len(xy)):
        a, b = x[q], y[q]
        if abs(a - (x0 + LX)) < h:
            a -= LX
        if abs(b - (y0 + LY)) < h:
            b -= LY
        t = key.get((int(round(a / h)), int(round(b / h))))
        if t is None:
            raise ValueError("vertex %d has no periodic image" % q)
        cid[q] = t
    return cid


def _reference_partition(right):
    """side order after genmesh.py:239-278: for N = -1, -2, -3 every side
    with right == N is swapped into the next slot j (j grows from 0)"""
    S = right.size
    perm = np.arange(S)
    r = right.copy()
    j = 0
    for N in (-1, -2, -3):
        for i in range(S):
            if r[i] == N:
                perm[i], perm[j] = perm[j], perm[i]
                r[i], r[j] = r[j], r[i]
                j += 1
    return perm


def build_mesh(xy, tris, bnd, periodic=None, default_bc=None, tol=1e-9, order="reference"):
    tris = np.array(tris, np.int64)
    K = tris.shape[0]
    x, y = xy[:, 0], xy[:, 1]
    J = ((x[tris[:, 1]] - x[tris[:, 0]]) * (y[tris[:, 2]] - y[tris[:, 0]]) -
         (x[tris[:, 2]] - x[tris[:, 0]]) * (y[tris[:, 1]] - y[tris[:, 0]]))
    cw = J < 0
    tris[cw, 0], tris[cw, 1] = tris[cw, 1], tris[cw, 0].copy()
    cid = _periodic_ids(xy, periodic, tol) if periodic is not None else np.arange(len(xy))
    nv = len(xy)
    # every (element, local side) in creation order: element-major, side minor
    a = np.stack([tris[:, 0], tris[:, 1], tris[:, 2]], 1).ravel()
    b = np.stack([tris[:, 1], tris[:, 2], tris[:, 0]], 1).ravel()
    ca, cb = cid[a], cid[b]
    key = np.minimum(ca, cb) * nv + np.maximum(ca, cb)
    pos = np.arange(3 * K)
    byk = np.lexsort((pos, key))
    ks = key[byk]
    first = np.ones(3 * K, bool)
    first[1:] = ks[1:] != ks[:-1]
    grp = np.cumsum(first) - 1
    if np.bincount(grp).max() > 2:
        raise ValueError("a side is shared by more than two triangles")
    creator = byk[first]                         # position of each group's creator
    sid_of_grp = np.empty(creator.size, np.int64)
    sid_of_grp[np.argsort(creator, kind="stable")] = np.arange(creator.size)
    S = creator.size
    creator_by_sid = np.empty(S, np.int64)
    creator_by_sid[sid_of_grp] = creator
    left = (creator_by_sid // 3).astype(np.int32)
    lsn = (creator_by_sid % 3).astype(np.int32)
    right = np.full(S, _OPEN, np.int32)
    rsn = np.zeros(S, np.int32)
    sec = byk[~first]
    sec_sid = sid_of_grp[grp[~first]]
    right[sec_sid] = sec // 3
    rsn[sec_sid] = sec % 3
    # element side lists
    sid_of_pos = np.empty(3 * K, np.int64)
    sid_of_pos[byk] = sid_of_grp[grp]
    es = sid_of_pos.reshape(K, 3).T.astype(np.int64)
    sxy = np.stack([x[a[creator_by_sid]], y[a[creator_by_sid]],
                    x[b[creator_by_sid]], y[b[creator_by_sid]]])
    # boundary codes from the tagged lines
    open_key = {}
    for s in np.nonzero(right == _OPEN)[0]:
        pa, pb = cid[a[creator_by_sid[s]]], cid[b[creator_by_sid[s]]]
        open_key[(min(pa, pb), max(pa, pb))] = s
    for (pa, pb, tag) in bnd:
        s = open_key.get((min(cid[pa], cid[pb]), max(cid[pa], cid[pb])))
        if s is not None and tag in TAG_CODE:
            right[s] = TAG_CODE[tag]
    untagged = right == _OPEN
    if np.any(untagged):
        if default_bc is None:
            raise ValueError("%d boundary sides carry no physical tag 10000/20000/30000"
                             % int(untagged.sum()))
        right[untagged] = default_bc
    rsn[right < 0] = 0
    # boundary sides first: -1, -2, -3, then interior
    if order == "stable":
        code = np.where(right < 0, -right, 4)
        srt = np.argsort(code, kind="stable")
    else:
        srt = _reference_partition(right)
    inv = np.empty(S, np.int64)
    inv[srt] = np.arange(S)
    V = np.stack([x[tris[:, 0]], y[tris[:, 0]], x[tris[:, 1]], y[tris[:, 1]],
                  x[tris[:, 2]], y[tris[:, 2]]])
    return Mesh(V=np.ascontiguousarray(V), elem_sides=np.ascontiguousarray(inv[es].astype(np.int32)),
                side_xy=np.ascontiguousarray(sxy[:, srt]), left=np.ascontiguousarray(left[srt]),
                right=np.ascontiguousarray(right[srt]), lsn=np.ascontiguousarray(lsn[srt]),
                rsn=np.ascontiguousarray(rsn[srt]))


def genmesh(inp, out=None, periodic=None, default_bc=None, order="reference"):
    """read a gmsh v2.2 file, build the .pmsh arrays, optionally write them"""
    xy, tris, bnd = parse_msh(inp)
    m = build_mesh(xy, tris, bnd, periodic=periodic, default_bc=default_bc, order=order)
    m.check()
    if out:
        write_pmsh(m, out)
    return m


def main(argv=None):
    ap = argparse.ArgumentParser(description="gmsh v2.2 ASCII -> .pmsh")
    ap.add_argument("inp")
    ap.add_argument("out")
    ap.add_argument("--periodic", nargs=2, type=float, metavar=("LX", "LY"))
    ap.add_argument("--default-bc", type=int, choices=(-1, -2, -3))
    a = ap.parse_args(argv)
    m = genmesh(a.inp, a.out, periodic=tuple(a.periodic) if a.periodic else None,
                default_bc=a.default_bc)
    print("%d elements, %d sides (%d boundary)" % (m.K, m.S, int((m.right < 0).sum())))


if __name__ == "__main__":
    sys.exit(main())
