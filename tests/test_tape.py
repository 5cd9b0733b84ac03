"""The FP64 tape (tests/tape.py) against the oracle, on the CPU.

The tape is what the GPU adjoint's gradients are compared with, so it is
pinned here to the oracle three ways: its right-hand side equals the
oracle's (oracle/dg_oracle.c eval_rhs), its RK steps reproduce the oracle's
runs, and its reverse-mode gradients match central (u0) and fourth-order
(eps) differences of the oracle's own forward solves."""
import numpy as np
import pytest
import torch

import oracle as O
from dgcuda_amd import _lib as L
from tests.helpers import CASES, rel_l2, unit_mesh
from tests.tape import Tape

TE_PULSE = (L.PROB_MAXWELL_PULSE, (1.0, 0.5, 0.5, 0.15))


def _setup(physics, p, n, integ, periodic, seed=4, vortex_at_rest=False):
    mesh = unit_mesh(n, periodic=periodic)
    prob, prm, fl = CASES[physics]
    if physics == L.MAXWELL_TE:
        prob, prm = TE_PULSE
    if vortex_at_rest:
        prm = (5.0, 0.5, 0.5, 0.0, 0.0, 1.0)
    o = O.Oracle(physics, p, integrator=integ, flux=fl, cfl=0.5, posfix=True,
                 problem=prob, problem_params=prm)
    o.set_mesh(mesh)
    eps = None
    if physics == L.MAXWELL_TE:
        eps = np.random.default_rng(seed).uniform(1, 4, mesh.K)
        o.set_eps(eps)
    o.project_ic()
    u0 = o.coeffs()
    rng = np.random.default_rng(seed + 1)
    u0 = u0 * (1.0 + 0.02 * rng.standard_normal(u0.shape))
    if physics == L.MAXWELL_TE:  # the pulse starts with E = 0: give E a random field
        u0[:2] = 0.05 * np.abs(u0[2]).max() * rng.standard_normal(u0[:2].shape)
    return mesh, o, Tape(o, mesh, physics), u0, eps


def _tensor(a):
    return None if a is None else torch.as_tensor(a, dtype=torch.float64)


@pytest.mark.parametrize("physics,periodic", [(L.MAXWELL_TE, False), (L.EULER, True),
                                              (L.EULER, False)])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tape_rhs_matches_oracle(physics, periodic, p):
    mesh, o, tp, u0, eps = _setup(physics, p, 5, L.RK4, periodic,
                                  vortex_at_rest=not periodic)
    dt = 1e-3
    k_orc = o.eval_rhs(u0, 0.0, dt)
    k_tape = tp.rhs(_tensor(u0), dt, _tensor(eps)).numpy()
    assert rel_l2(k_tape, k_orc) <= 1e-13


@pytest.mark.parametrize("physics,integ,periodic", [(L.MAXWELL_TE, L.RK4, False),
                                                    (L.EULER, L.SSPRK3, True),
                                                    (L.EULER, L.RK2MID, True)])
def test_tape_steps_match_oracle(physics, integ, periodic):
    mesh, o, tp, u0, eps = _setup(physics, 2, 5, integ, periodic)
    nsteps, dt = 8, 2e-3
    o.set_coeffs(u0)
    o.run(nsteps, dt)
    ref = o.coeffs()
    got = tp.run(_tensor(u0), nsteps, dt, integ, _tensor(eps)).numpy()
    assert rel_l2(got, ref) <= 1e-12


def _orc_objective(tp, o, kind, params):
    return float(tp.objective(_tensor(o.coeffs()), kind, params))


def test_tape_gradient_te_matches_oracle_differences():
    """J is quadratic in u0 (linear dynamics, squared objective): central
    differences are exact up to rounding; in eps they are fourth-order"""
    mesh, o, tp, u0, eps = _setup(L.MAXWELL_TE, 2, 4, L.RK4, False)
    nsteps, dt, prm = 8, 2e-3, (0.5, 0.9, 0.4, 0.8, 2)
    J, gu, ge = tp.gradient(u0, nsteps, dt, L.RK4, L.OBJ_PATCH_SQ, prm, eps)
    assert J > 0
    rng = np.random.default_rng(11)

    def solve(u, e):
        o.set_eps(e)
        o.set_coeffs(u)
        o.run(nsteps, dt)
        return _orc_objective(tp, o, L.OBJ_PATCH_SQ, prm)

    d = rng.standard_normal(u0.shape)
    d /= np.linalg.norm(d)
    h = 1e-3 * np.linalg.norm(u0)
    fd = (solve(u0 + h * d, eps) - solve(u0 - h * d, eps)) / (2 * h)
    assert abs(fd - np.vdot(gu, d)) <= 1e-8 * abs(fd)
    de = rng.standard_normal(eps.shape)
    de /= np.linalg.norm(de)
    h = 1e-3
    Jp = [solve(u0, eps + s * h * de) for s in (1, -1, 2, -2)]
    fd = (8 * (Jp[0] - Jp[1]) - (Jp[2] - Jp[3])) / (12 * h)
    assert abs(fd - np.vdot(ge, de)) <= 1e-6 * abs(fd)


def test_tape_gradient_euler_matches_oracle_differences():
    mesh, o, tp, u0, eps = _setup(L.EULER, 2, 4, L.SSPRK3, True)
    nsteps, dt = 6, 2e-3
    J, gu, _ = tp.gradient(u0, nsteps, dt, L.SSPRK3, L.OBJ_ENTROPY)
    rng = np.random.default_rng(12)
    for _ in range(2):
        d = rng.standard_normal(u0.shape)
        d /= np.linalg.norm(d)
        h = 1e-6
        Jp = []
        for s in (1, -1):
            o.set_coeffs(u0 + s * h * d)
            o.run(nsteps, dt)
            Jp.append(_orc_objective(tp, o, L.OBJ_ENTROPY, ()))
        fd = (Jp[0] - Jp[1]) / (2 * h)
        assert abs(fd - np.vdot(gu, d)) <= 1e-6 * abs(fd)
