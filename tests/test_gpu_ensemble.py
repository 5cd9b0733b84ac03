"""Ensembles (dg_ens_*): independent meshes of different sizes and orders
stepped together with one stage launch.  Each instance must evolve as it
does alone on a handle (the same per-element arithmetic, its own CFL dt; the
compiler may contract a multiply-add differently in the merged kernel, so
the bound is a few ulps, 1e-14 relative) and match the oracle to the parity
tolerance."""
import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.ensemble import Ensemble
from dgcuda_amd.solver import Solver
from tests.helpers import rel_l2, unit_mesh

pytestmark = pytest.mark.gpu

SWE_ROE = dict(physics=L.SWE, integrator=L.SSPRK3, flux=L.FLUX_ROE, g=9.81,
               problem=L.PROB_SWE_BUMP)


def _instances(seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for n, p in ((6, 1), (9, 2), (12, 3), (7, 1), (10, 2), (5, 3)):
        prm = (1.0, 0.1, rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.05, 0.2), 1.0)
        out.append((unit_mesh(n, periodic=False, seed=n), p, prm))
    return out


def _alone(kw, mesh, p, prm, nsteps, dt=None):
    s = Solver(kw["physics"], p, integrator=kw["integrator"], flux=kw.get("flux", 0),
               g=kw.get("g", 9.8), problem=kw["problem"], problem_params=prm)
    s.set_mesh(mesh).project_ic()
    t, dtl = s.step(nsteps, dt=dt)
    return s, t, dtl


@pytest.mark.parametrize("block,split", [(0, False), (64, False), (128, False), (256, False),
                                         (64, True), (128, True), (256, True)])
def test_ensemble_matches_instances_alone(block, split):
    """SWE bumps with the Roe flux, SSP-RK3, walls, p in {1, 2, 3}, six
    meshes of different sizes: after 20 CFL steps (two calls) every
    instance's coefficients, time and dt equal the instance stepped alone"""
    inst = _instances()
    E = Ensemble(**SWE_ROE)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(block, split)
    info = E.info()
    assert info.ninst == len(inst) and info.block in (64, 128, 256)
    # three orders: one launch per stage, or one per order group and stage
    assert info.launches_per_step == (3 * 3 if info.split else 3) + 1
    if block:
        assert info.block == block and bool(info.split) == split
        assert list(info.block_per_order) == [0, block, block, block, 0, 0]
    else:
        print("tuned: block", info.block, "split", info.split, "merged", list(info.tune_ms),
              "split", list(info.tune_split_ms), "per order", list(info.block_per_order),
              "best", info.tune_best_ms)
        per = list(info.block_per_order)
        assert per[0] == 0 and per[4] == per[5] == 0
        assert all(b in (64, 128, 256) for b in per[1:4])
        if not info.split:
            assert per[1:4] == [info.block] * 3
    E.step(12)
    E.step(8)
    for i, (mesh, p, prm) in enumerate(inst):
        s, t, dtl = _alone(SWE_ROE, mesh, p, prm, 12)
        t, dtl = s.step(8, t=t)
        te, dte, n = E.state(i)
        print("instance", i, "K", mesh.K, "p", p, "rel", rel_l2(E.coeffs(i), s.coeffs()))
        assert n == 20 and abs(te - t) <= 1e-14 * t and abs(dte - dtl) <= 1e-14 * dtl
        assert rel_l2(E.coeffs(i), s.coeffs()) <= 1e-14
        s.close()
    E.close()


def test_ensemble_oracle_and_positivity():
    """Euler vortices in walled boxes (positivity check on), RK4, p = 0..3,
    and one instance against the oracle"""
    import oracle as O
    kw = dict(physics=L.EULER, integrator=L.RK4, problem=L.PROB_EULER_VORTEX)
    prm = (5.0, 0.5, 0.5, 1.0, 1.0, 1.0)
    meshes = [(unit_mesh(n, periodic=False, seed=3 + n), p) for n, p in ((6, 0), (7, 1), (8, 2), (6, 3))]
    E = Ensemble(**kw, problem_params=prm)
    for mesh, p in meshes:
        E.add(mesh, p)
    E.build(128)
    E.step(10)
    for i, (mesh, p) in enumerate(meshes):
        s, t, dtl = _alone(kw, mesh, p, prm, 10)
        assert rel_l2(E.coeffs(i), s.coeffs()) <= 1e-14, i
        assert abs(E.state(i)[0] - t) <= 1e-14 * t
        s.close()
    mesh, p = meshes[2]
    o = O.Oracle(O.EULER, p, integrator=1, posfix=True, problem=L.PROB_EULER_VORTEX,
                 problem_params=prm)
    o.set_mesh(mesh).project_ic()
    rc, to, _ = o.run(10)
    assert rc == 0 and abs(E.state(2)[0] - to) <= 1e-13 * to
    assert rel_l2(E.coeffs(2), o.coeffs()) < 1e-12
    E.close()


def test_ensemble_fixed_dt_and_errors():
    """a fixed dt for every instance; the L2 error of an instance is its
    handle's; adding after build and a sixth-order instance are refused"""
    inst = _instances(5)[:3]
    E = Ensemble(**SWE_ROE)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(256)
    E.step(6, dt=1e-3)
    for i, (mesh, p, prm) in enumerate(inst):
        s, t, _ = _alone(SWE_ROE, mesh, p, prm, 6, dt=1e-3)
        assert rel_l2(E.coeffs(i), s.coeffs()) <= 1e-14
        te, dte, n = E.state(i)
        assert abs(te - 6e-3) < 1e-15 and dte == 1e-3 and n == 6
        assert rel_l2(E.l2_error(i, te), s.l2_error(t)) <= 1e-13
        s.close()
    with pytest.raises(Exception):
        E.add(inst[0][0], 1)
    E.close()
    E2 = Ensemble(**SWE_ROE)
    with pytest.raises(Exception):
        E2.add(inst[0][0], 6)
    E2.close()


@pytest.mark.parametrize("case", ["swe_roe", "advection"])
def test_ensemble_high_orders_match_instances_alone(case):
    """instances of order 4 and 5 (split launches at 64-element blocks, the
    one candidate the tuner times then) beside lower orders: every instance
    evolves as it does alone"""
    rng = np.random.default_rng(5)
    if case == "swe_roe":
        kw = SWE_ROE
        orders = ((6, 4), (5, 5), (8, 1), (7, 3))
        inst = [(unit_mesh(n, periodic=False, seed=n), p,
                 (1.0, 0.1, rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.05, 0.2), 1.0))
                for n, p in orders]
    else:
        kw = dict(physics=L.ADVECTION, integrator=L.RK4, problem=L.PROB_ADV_SINSIN)
        inst = [(unit_mesh(n, periodic=True, seed=n), p, ()) for n, p in ((6, 5), (9, 2), (5, 4))]
    E = Ensemble(**kw)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(0)
    info = E.info()
    print(case, "block", info.block, "split", info.split, "split ms", list(info.tune_split_ms),
          "merged ms", list(info.tune_ms))
    assert info.block == 64 and info.split == 1
    assert list(info.tune_ms) == [-1.0, -1.0, -1.0] and info.tune_split_ms[1] == -1.0
    E.step(6)
    E.step(4)
    for i, (mesh, p, prm) in enumerate(inst):
        s, t, dtl = _alone(kw, mesh, p, prm, 10)
        te, dte, n = E.state(i)
        r = rel_l2(E.coeffs(i), s.coeffs())
        print("instance", i, "p", p, "rel", r)
        assert n == 10 and abs(te - t) <= 1e-14 * t and abs(dte - dtl) <= 1e-14 * dtl
        assert r <= 1e-14
        s.close()
    E.close()


def test_ensemble_euler_high_orders_match_instances_alone():
    """Euler instances of order 4 and 5 run the four-lane kernel's ensemble
    form (gather path, 64-element blocks) beside one-thread groups.  Alone,
    a run of steps at p >= 4 reads stored traces instead of folding the
    neighbours' modes: the same values to rounding (1e-13, as the stored-trace
    test), not bit for bit"""
    rng = np.random.default_rng(21)
    kw = dict(physics=L.EULER, integrator=L.SSPRK3, problem=L.PROB_EULER_VORTEX, cfl=0.2)
    inst = [(unit_mesh(n, periodic=True, seed=n), p,
             (5.0, rng.uniform(0.4, 0.6), rng.uniform(0.4, 0.6), 1.0, 1.0, 1.0))
            for n, p in ((6, 4), (5, 5), (7, 2), (4, 4))]
    E = Ensemble(**kw)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(0)
    info = E.info()
    assert info.split == 1 and info.block_per_order[4] == 64 and info.block_per_order[5] == 64
    E.step(5)
    E.step(4)
    for i, (mesh, p, prm) in enumerate(inst):
        s = Solver(L.EULER, p, integrator=L.SSPRK3, problem=L.PROB_EULER_VORTEX,
                   problem_params=prm, cfl=0.2)
        s.set_mesh(mesh).project_ic()
        t, dtl = s.step(9)
        te, dte, n = E.state(i)
        r = rel_l2(E.coeffs(i), s.coeffs())
        print("euler instance", i, "p", p, "rel", r)
        assert n == 9 and abs(te - t) <= 1e-13 * t and abs(dte - dtl) <= 1e-13 * dtl
        assert r <= (1e-13 if p >= 4 else 1e-14)
        s.close()
    E.close()


def test_ensemble_order_limits():
    """instances of order 4 or 5 take block 64 | split only"""
    E = Ensemble(**SWE_ROE)
    E.add(unit_mesh(5, periodic=False), 4, (1.0, 0.1, 0.5, 0.5, 0.1, 1.0))
    with pytest.raises(L.DgError, match="order 4 or 5"):
        E.build(128)
    with pytest.raises(L.DgError, match="order 4 or 5"):
        E.build(64)
    E.close()


@pytest.mark.parametrize("case", ["euler_vortex_inflow", "advection_sinsin_dirichlet"])
def test_ensemble_time_dependent_boundaries(case):
    """boundary data that depend on time (the exact solution on inflow /
    Dirichlet sides) are evaluated per instance at its own stage times, with
    its own problem parameters: every instance evolves as it does alone"""
    rng = np.random.default_rng(9)
    if case == "euler_vortex_inflow":
        kw = dict(physics=L.EULER, integrator=L.SSPRK3, problem=L.PROB_EULER_VORTEX, cfl=0.3)
        inst = [(unit_mesh(n, periodic=False, seed=n, bc=(-3, -2, -3, -2)), p,
                 (5.0, rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), 1.0, 1.0, 1.0))
                for n, p in ((6, 1), (5, 2), (7, 3), (4, 2))]
    else:
        kw = dict(physics=L.ADVECTION, integrator=L.RK4, problem=L.PROB_ADV_SINSIN)
        inst = [(unit_mesh(n, periodic=False, seed=n, bc=(-2, -2, -2, -2)), p, ())
                for n, p in ((6, 1), (5, 3), (8, 2), (5, 4))]
    E = Ensemble(**kw)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    E.build(0)
    E.step(7)
    E.step(5)
    for i, (mesh, p, prm) in enumerate(inst):
        s = Solver(kw["physics"], p, integrator=kw["integrator"], problem=kw["problem"],
                   problem_params=prm, cfl=kw.get("cfl", 0.0))
        s.set_mesh(mesh).project_ic()
        t, dtl = s.step(12)
        te, dte, n = E.state(i)
        r = rel_l2(E.coeffs(i), s.coeffs())
        print(case, "instance", i, "p", p, "rel", r, "t", te, t)
        assert n == 12 and abs(te - t) <= 1e-14 * t and abs(dte - dtl) <= 1e-14 * dtl
        assert r <= 1e-14
        s.close()
    E.close()


@pytest.mark.parametrize("blocks", ["64,256,128,64,64,64", "256,128,256,128,64,64"])
def test_ensemble_per_order_blocks_match_instances_alone(blocks):
    """the split layout with a different block size per order group (what the
    tuner's refinement may pick; fixed here through DG_ENS_BLOCKS): every
    instance still evolves as it does alone"""
    import os
    inst = _instances(17)
    E = Ensemble(**SWE_ROE)
    for mesh, p, prm in inst:
        E.add(mesh, p, prm)
    os.environ["DG_ENS_BLOCKS"] = blocks
    try:
        E.build(0)
    finally:
        os.environ.pop("DG_ENS_BLOCKS", None)
    info = E.info()
    want = [int(b) for b in blocks.split(",")]
    assert info.split == 1 and list(info.block_per_order)[1:4] == want[1:4]
    print("per order", list(info.block_per_order), "split", info.split, "best", info.tune_best_ms)
    E.step(9)
    for i, (mesh, p, prm) in enumerate(inst):
        s, t, dtl = _alone(SWE_ROE, mesh, p, prm, 9)
        te, dte, n = E.state(i)
        assert n == 9 and abs(te - t) <= 1e-14 * t
        assert rel_l2(E.coeffs(i), s.coeffs()) <= 1e-14
        s.close()
    E.close()
