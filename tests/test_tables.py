"""Pins the reference-element tables (quadrature, basis) against closed forms
and an exact Gram-Schmidt, and the product's compile-time tables against
the oracle's."""
import math
import os
import re
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PMAX = 5


def _parse_ref_tables(path):
    txt = open(path).read()
    num = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
    body = txt[txt.index("dgt_basis[DGT_NPMAX][DGT_NPMAX]"):]
    body = body[body.index("{") + 1:body.index("};")]
    rows = [[float(v) for v in re.findall(num, r)] for r in re.findall(r"\{([^{}]*)\}", body)]
    q2, q1 = [], []
    for p in range(PMAX + 1):
        m = re.search(r"dgt_q2_p%d\[\d+\] = \{([^}]*)\}" % p, txt)
        q2.append(np.array([float(v) for v in re.findall(num, m.group(1))]))
        m = re.search(r"dgt_q1_p%d\[\d+\] = \{([^}]*)\}" % p, txt)
        q1.append(np.array([float(v) for v in re.findall(num, m.group(1))]))
    return rows, q2, q1


REF = _parse_ref_tables(os.path.join(ROOT, "oracle", "ref_tables.h"))


def monomials(n_p):
    """graded monomial exponents (a, b): 1, r, s, r^2, rs, s^2, ..."""
    out, d = [], 0
    while len(out) < n_p:
        for j in range(d + 1):
            out.append((d - j, j))
        d += 1
    return out[:n_p]


def tri_moment(a, b):
    """integral of r^a s^b over the reference triangle (0,0),(1,0),(0,1)"""
    return Fraction(math.factorial(a) * math.factorial(b), math.factorial(a + b + 2))


def test_two_table_copies_identical():
    a = open(os.path.join(ROOT, "oracle", "ref_tables.h")).read()
    b = open(os.path.join(ROOT, "dgcuda_amd", "csrc", "dg_ref_tables.h")).read()
    strip = lambda s: s[s.index("#define DGT_PMAX"):s.rindex("#endif")]  # noqa: E731
    assert strip(a) == strip(b)


@pytest.mark.parametrize("p", range(PMAX + 1))
def test_quadrature_2d_moments(p):
    """the degree-2p rule (weights halved, main.cu:181) integrates r^a s^b
    exactly for a + b <= 2p"""
    q = REF[1][p].reshape(-1, 3)
    r, s, w = q[:, 0], q[:, 1], q[:, 2] / 2.0
    for d in range(2 * p + 1):
        for a in range(d + 1):
            b = d - a
            got = np.sum(w * r ** a * s ** b)
            assert abs(got - float(tri_moment(a, b))) < 2e-15, (p, a, b)


@pytest.mark.parametrize("p", range(PMAX + 1))
def test_quadrature_1d_gauss(p):
    q = REF[2][p].reshape(-1, 2)
    x, w = q[:, 0], q[:, 1]
    assert len(x) == p + 1
    for k in range(2 * (p + 1)):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 2e-15, (p, k)
    # the reversed right trace (conserv_kernels.cu:885) needs an exactly symmetric rule
    assert np.array_equal(x, -x[::-1])
    assert np.array_equal(w, w[::-1])


def gram_schmidt(n_p):
    """orthonormal basis on the reference triangle by Gram-Schmidt on the
    graded monomials, exact rational arithmetic until the normalisation"""
    mons = monomials(n_p)
    G = [[tri_moment(a + c, b + d) for (c, d) in mons] for (a, b) in mons]
    qs, norms = [], []
    for k in range(n_p):
        v = [Fraction(0)] * n_p
        v[k] = Fraction(1)
        for q, nq in zip(qs, norms):
            ip = sum(v[i] * G[i][j] * q[j] for i in range(n_p) for j in range(n_p))
            v = [v[i] - ip / nq * q[i] for i in range(n_p)]
        nv = sum(v[i] * G[i][j] * v[j] for i in range(n_p) for j in range(n_p))
        qs.append(v)
        norms.append(nv)
    return np.array([[float(c) / math.sqrt(float(n)) for c in q] for q, n in zip(qs, norms)])


@pytest.mark.parametrize("p", range(PMAX + 1))
def test_basis_is_reference_gram_schmidt(p):
    """basis.cu:19-66 tabulates Gram-Schmidt on graded monomials; its
    rounding grows with the order (Gram error ~1e-12 at p=3, ~1e-9 at p=4)"""
    n_p = (p + 1) * (p + 2) // 2
    GS = gram_schmidt(n_p)
    tol = {0: 1e-15, 1: 1e-15, 2: 5e-14, 3: 2e-11, 4: 2e-9, 5: 5e-7}[p]
    for i in range(n_p):
        row = np.zeros(n_p)
        row[:len(REF[0][i])] = REF[0][i]
        assert np.max(np.abs(row - GS[i])) <= tol * np.max(np.abs(GS[i])), (p, i)


@pytest.mark.parametrize("p", range(PMAX + 1))
def test_basis_gram_matrix(p):
    n_p = (p + 1) * (p + 2) // 2
    mons = monomials(n_p)
    C = np.zeros((n_p, n_p))
    for i in range(n_p):
        C[i, :len(REF[0][i])] = REF[0][i]
    M = np.array([[float(tri_moment(a + c, b + d)) for (c, d) in mons] for (a, b) in mons])
    Gram = C @ M @ C.T
    tol = {0: 1e-15, 1: 1e-14, 2: 1e-13, 3: 1e-11, 4: 1e-9, 5: 1e-6}[p]
    assert np.max(np.abs(Gram - np.eye(n_p))) < tol


def _poly(coef, r, s):
    mons = monomials(len(coef))
    return sum(c * r ** a * s ** b for c, (a, b) in zip(coef, mons))


def _poly_dr(coef, r, s):
    mons = monomials(len(coef))
    return sum(c * a * r ** max(a - 1, 0) * s ** b for c, (a, b) in zip(coef, mons) if a > 0)


def _poly_ds(coef, r, s):
    mons = monomials(len(coef))
    return sum(c * b * r ** a * s ** max(b - 1, 0) for c, (a, b) in zip(coef, mons) if b > 0)


@pytest.mark.parametrize("p", range(PMAX + 1))
def test_oracle_tables_match_definition(p, oracle_lib):
    """oracle tables = basis values / weighted gradients at the quadrature
    points; side k points run counter-clockwise along edge k of the
    reference triangle (V1V2, V2V3, V3V1)"""
    o = oracle_lib.Oracle(oracle_lib.EULER, p)
    T = o.tables()
    n_p = o.n_p
    q = REF[1][p].reshape(-1, 3)
    r, s, w = q[:, 0], q[:, 1], q[:, 2] / 2
    x1 = REF[2][p].reshape(-1, 2)[:, 0]
    assert np.array_equal(T["w"], w) and np.array_equal(T["r1"], r)
    corners = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    for i in range(n_p):
        c = REF[0][i]
        scale = max(1.0, np.max(np.abs(c)))
        assert np.allclose(T["basis"][i], _poly(c, r, s), rtol=0, atol=1e-14 * scale)
        assert np.allclose(T["gx"][i], w * _poly_dr(c, r, s), rtol=0, atol=1e-14 * scale)
        assert np.allclose(T["gy"][i], w * _poly_ds(c, r, s), rtol=0, atol=1e-14 * scale)
        for k in range(3):
            frac = (1 + x1) / 2  # position along the edge from V_k to V_k+1
            pts = corners[k][None, :] * (1 - frac[:, None]) + corners[k + 1][None, :] * frac[:, None]
            assert np.allclose(T["side"][k, i], _poly(c, pts[:, 0], pts[:, 1]), rtol=0,
                               atol=2e-14 * scale), (i, k)
        assert np.allclose(T["vertex"][i], _poly(c, corners[:3, 0], corners[:3, 1]), rtol=0,
                           atol=1e-14 * scale)


def _parse_tabvals():
    txt = open(os.path.join(ROOT, "dgcuda_amd", "csrc", "dg_tabvals.h")).read()
    out = {}
    for p, body in re.findall(r"struct TabC<(\d)> \{(.*?)\n\};", txt, flags=re.S):
        d = {}
        for name, vals in re.findall(r"double (\w+)\[\d+\] = \{([^}]*)\}", body):
            d[name] = np.array([float.fromhex(v.strip()) for v in vals.split(",")])
        out[int(p)] = d
    return out


def test_product_tables_bitwise_equal_oracle(oracle_lib):
    """the kernels' compile-time tables carry the same bits as the oracle's"""
    tv = _parse_tabvals()
    for p in range(PMAX + 1):
        T = oracle_lib.Oracle(oracle_lib.EULER, p).tables()
        d = tv[p]
        assert np.array_equal(d["B"], T["basis"].ravel())
        assert np.array_equal(d["GX"], T["gx"].ravel())
        assert np.array_equal(d["GY"], T["gy"].ravel())
        assert np.array_equal(d["SIDE"], T["side"].ravel())
        assert np.array_equal(d["VERT"], T["vertex"].ravel())
        assert np.array_equal(d["W"], T["w"]) and np.array_equal(d["W1"], T["w1"])
        assert np.array_equal(d["X1"], T["r1d"])
