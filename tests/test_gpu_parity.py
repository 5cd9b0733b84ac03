"""HIP path vs the oracle on the same seeded inputs (bit-for-bit inputs,
FP64 tolerance 1e-12 relative L2 on the modal coefficients).

Every case calls the library through its C ABI (ctypes); the oracle is the
CPU restatement of the reference's eval_surface / eval_volume / eval_rhs /
RK stages (oracle/dg_oracle.c)."""
import numpy as np
import pytest

from dgcuda_amd import _lib as L
from tests.helpers import make_pair, rel_l2, unit_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative L2 on modal coefficients (FP64)

PHYS = [L.ADVECTION, L.EULER, L.SWE, L.MAXWELL_TM, L.MAXWELL_TE]


def _eps(mesh, seed=3):
    return np.random.default_rng(seed).uniform(1.0, 4.0, mesh.K)


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("periodic", [True, False])
def test_eval_rhs_parity(physics, p, periodic):
    mesh = unit_mesh(6, periodic=periodic)
    eps = _eps(mesh) if physics == L.MAXWELL_TE else None
    s, o = make_pair(physics, p, mesh, eps_elem=eps)
    s.project_ic()
    o.project_ic()
    u0 = o.coeffs()
    assert rel_l2(s.coeffs(), u0) < TOL, "projection of the initial condition"
    # a perturbed state exercises every mode
    u = u0 * (1.0 + 0.05 * np.random.default_rng(p).standard_normal(u0.shape))
    k_gpu = s.eval_rhs(u, t=0.0, dt=0.01)
    k_ref = o.eval_rhs(u, t=0.0, dt=0.01)
    print("physics", physics, "p", p, "periodic", periodic, "rel", rel_l2(k_gpu, k_ref))
    assert rel_l2(k_gpu, k_ref) < TOL


@pytest.mark.parametrize("physics,p", [(L.EULER, 2), (L.EULER, 4), (L.EULER, 5), (L.SWE, 5)])
def test_eval_rhs_positivity_fix(physics, p):
    """states whose point values go negative in every fifth element: those
    elements (and their neighbours' view of them) take their means only, in
    the register-resident and the streamed (N * n_p > 48) stage kernels"""
    mesh = unit_mesh(6)
    s, o = make_pair(physics, p, mesh)
    s.project_ic()
    o.project_ic()
    u = o.coeffs()
    u[:, 1:, ::5] *= 400.0
    k_gpu = s.eval_rhs(u, t=0.0, dt=0.01)
    k_ref = o.eval_rhs(u, t=0.0, dt=0.01)
    print("positivity eval_rhs", physics, p, rel_l2(k_gpu, k_ref))
    assert rel_l2(k_gpu, k_ref) < TOL


@pytest.mark.parametrize("physics,p,integrator", [
    (L.EULER, 4, L.SSPRK3), (L.EULER, 5, L.RK4), (L.SWE, 3, L.SSPRK3),
    (L.MAXWELL_TE, 3, L.RK4), (L.MAXWELL_TM, 5, L.SSPRK2), (L.ADVECTION, 5, L.RK2MID)])
def test_steps_parity_high_order(physics, p, integrator):
    """11 steps: the graphs of 8, 2 and 1 steps in turn"""
    mesh = unit_mesh(8)
    eps = _eps(mesh) if physics == L.MAXWELL_TE else None
    s, o = make_pair(physics, p, mesh, integrator=integrator, eps_elem=eps)
    s.project_ic()
    o.project_ic()
    ts, dts = s.step(11)
    rc, to, dto = o.run(11)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * abs(to)
    r = rel_l2(s.coeffs(), o.coeffs())
    print("physics", physics, "p", p, "integrator", integrator, "rel", r)
    assert r < TOL


@pytest.mark.parametrize("integrator", [L.RK4, L.RK2MID, L.SSPRK2, L.SSPRK3])
def test_steps_parity_euler_vortex(integrator):
    mesh = unit_mesh(12)
    s, o = make_pair(L.EULER, 2, mesh, integrator=integrator)
    s.project_ic()
    o.project_ic()
    ts, dts = s.step(20)
    rc, to, dto = o.run(20)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * abs(to)
    assert abs(dts - dto) <= 1e-13 * abs(dto)
    r = rel_l2(s.coeffs(), o.coeffs())
    print("integrator", integrator, "rel", r, "t", ts)
    assert r < TOL


@pytest.mark.parametrize("physics", PHYS)
def test_steps_parity_walls(physics):
    mesh = unit_mesh(10, periodic=False)
    eps = _eps(mesh) if physics == L.MAXWELL_TE else None
    s, o = make_pair(physics, 2, mesh, integrator=L.SSPRK3, cfl=0.3, eps_elem=eps)
    s.project_ic()
    o.project_ic()
    s.step(15)
    rc, _, _ = o.run(15)
    assert rc == 0
    r = rel_l2(s.coeffs(), o.coeffs())
    print("walls physics", physics, "rel", r)
    assert r < TOL


def test_limiter_parity_p1():
    mesh = unit_mesh(10)
    s, o = make_pair(L.EULER, 1, mesh, integrator=L.RK4, limiter=True)
    s.project_ic()
    o.project_ic()
    ts, _ = s.step(10)
    rc, to, _ = o.run(10)
    assert rc == 0 and abs(ts - to) <= 1e-13 * to
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL


def test_fixed_dt_and_time():
    mesh = unit_mesh(8)
    s, o = make_pair(L.ADVECTION, 1, mesh, integrator=L.SSPRK2, cfl=0.3)
    s.project_ic()
    o.project_ic()
    t, dt = s.step(7, dt=1e-3, t=0.25)
    rc, to, _ = o.run(7, dt=1e-3)
    assert abs(t - (0.25 + 7e-3)) < 1e-14 and dt == 1e-3
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL


def _supersonic(p, integrator, cfl, limiter=False):
    import os
    from dgcuda_amd import mesh as M
    mesh = M.read_pmsh(os.path.join(os.path.dirname(__file__), "golden", "supersonic_A.pmsh"))
    return make_pair(L.EULER, p, mesh, integrator=integrator, cfl=cfl, limiter=limiter,
                     problem=L.PROB_EULER_SUPERSONIC, params=())


@pytest.mark.parametrize("p,integrator,limiter", [(2, L.SSPRK3, False), (1, L.RK4, True),
                                                  (1, L.RK2MID, False)])
def test_run_to_endtime(p, integrator, limiter):
    """-T: the step that would pass t_end is shortened to land on it
    (time_integrator.cu:225-231); steps, time and state as the oracle's loop"""
    s, o = _supersonic(p, integrator, 0.2 if integrator == L.SSPRK3 else 0.0, limiter)
    s.project_ic()
    o.project_ic()
    t, n, _ = s.run(t_end=0.0123)
    rc, to, no, _, _ = o.run_until(t_end=0.0123)
    print("t_end", p, integrator, limiter, t, n, to, no)
    assert rc == 0 and n == no and t == 0.0123 and to == 0.0123
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL
    # already there: no step
    t2, n2, _ = s.run(t_end=0.0123, t=0.0123)
    assert n2 == 0 and t2 == 0.0123


@pytest.mark.parametrize("limiter", [False, True])
def test_run_convergence(limiter):
    """-c tol: steady supersonic vortex marched until max |c_new - c_old| over
    the first K*N entries is <= tol (time_integrator.cu:403-417)"""
    s, o = _supersonic(1, L.RK4, 0.0, limiter)
    s.project_ic()
    o.project_ic()
    t, n, cv = s.run(max_steps=5000, conv_tol=1e-4)
    rc, to, no, cvo, _ = o.run_until(max_steps=5000, conv_tol=1e-4)
    print("conv", limiter, n, no, cv, cvo)
    assert rc == 0 and n == no and 10 < n < 5000
    assert abs(t - to) <= 1e-12 * to and abs(cv - cvo) <= 1e-9 * cvo and cv <= 1e-4
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL


def test_run_max_steps_matches_step():
    mesh = unit_mesh(8)
    s, o = make_pair(L.EULER, 2, mesh, integrator=L.SSPRK3, cfl=0.2)
    s.project_ic()
    c0 = s.coeffs()
    t1, n1, _ = s.run(max_steps=9)
    c1 = s.coeffs()
    s.set_coeffs(c0)
    t2, _ = s.step(9)
    assert n1 == 9 and t1 == t2 and np.array_equal(c1, s.coeffs())


@pytest.mark.parametrize("physics", [L.ADVECTION, L.EULER])
def test_output_side(physics):
    """eval_u, plot_exact and the L1 / L2 error integrals after a few steps"""
    mesh = unit_mesh(8, periodic=False, bc=(-2, -2, -3, -3))
    prob = L.PROB_ADV_LINEAR if physics == L.ADVECTION else L.PROB_EULER_VORTEX
    prm = () if physics == L.ADVECTION else (5.0, 0.5, 0.5, 1.0, 1.0, 1.0)
    s, o = make_pair(physics, 2, mesh, integrator=L.SSPRK3, cfl=0.2, problem=prob, params=prm)
    s.project_ic()
    o.project_ic()
    t, _ = s.step(4)
    rc, to, _ = o.run(4)
    assert abs(t - to) <= 1e-14 * to
    for n in range(s.N):
        for a, b in zip(s.eval_u(n), o.eval_u(n)):
            assert rel_l2(a, b) < TOL
        for a, b in zip(s.eval_exact(n, t), o.eval_exact(n, t)):
            assert rel_l2(a, b) < 1e-14
    assert rel_l2(s.l2_error(t), o.l2_error(t)) < 1e-10
    assert rel_l2(s.l1_error(t), o.l1_error(t)) < 1e-10


@pytest.mark.parametrize("integrator,p", [(L.RK4, 2), (L.SSPRK3, 2), (L.RK4, 4)])
def test_positivity_fix_parity(integrator, p):
    """a strong vortex in a walled unit box: the pressure at a check point of
    one element goes negative after a few steps, so the positivity fix
    (check_physical, element-local) acts; the state returned after every
    step is the oracle's, stepped one at a time and in one graph call"""
    mesh = unit_mesh(6, periodic=False)
    s, o = make_pair(L.EULER, p, mesh, integrator=integrator)
    s.project_ic()
    o.project_ic()
    u0 = s.coeffs()
    for k in range(10):
        s.step(1, dt=1e-3, t=k * 1e-3)
        o.run(1, dt=1e-3)
        rel = rel_l2(s.coeffs(), o.coeffs())
        print("positivity", integrator, "step", k + 1, rel)
        assert rel < TOL, k + 1
    s.set_coeffs(u0)
    o.set_coeffs(u0)
    s.step(10, dt=1e-3)
    o.run(10, dt=1e-3)
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL


@pytest.mark.parametrize("p", [1, 2, 4])
def test_rotating_hill_parity(p):
    """advection/rotatinghill on the reference's mesh A: the velocity
    2 pi (-y, x) at every volume and side point, the time-dependent inflow
    and the CFL speed at quadrature point n_q / 2, stepped with RK4 and the
    CFL rule as the oracle's loop; then to t_end (the last step shortened)"""
    import os
    from dgcuda_amd import mesh as M
    mesh = M.read_pmsh(os.path.join(os.path.dirname(__file__), "golden", "rotatinghill_A.pmsh"))
    s, o = make_pair(L.ADVECTION, p, mesh, integrator=L.RK4, problem=L.PROB_ADV_ROTHILL,
                     params=())
    s.project_ic()
    o.project_ic()
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL
    assert rel_l2(s.lam(), o.lam()) < 1e-15
    ts, dts = s.step(30)
    rc, to, dto = o.run(30)
    assert rc == 0 and abs(ts - to) <= 1e-13 * to and abs(dts - dto) <= 1e-13 * dto
    r = rel_l2(s.coeffs(), o.coeffs())
    print("rotating hill p", p, "rel", r, "t", ts)
    assert r < TOL
    t, n, _ = s.run(t_end=0.05, t=ts)
    rc, to, no, _, _ = o.run_until(t_end=0.05, t0=to)
    assert rc == 0 and n == no and t == 0.05 and to == 0.05
    assert rel_l2(s.coeffs(), o.coeffs()) < TOL
    assert rel_l2(s.l2_error(t), o.l2_error(t)) < 1e-10


@pytest.mark.parametrize("physics,p,n,integrator,periodic,nsteps", [
    (L.EULER, 2, 128, L.SSPRK3, True, 20),       # 32,768 elements: 128 blocks of 256
    (L.EULER, 4, 60, L.SSPRK3, True, 8),         # 7,200 elements, four-lane kernel (partial block)
    (L.MAXWELL_TE, 3, 90, L.RK4, False, 10),     # 16,200 elements, walls, per-element eps
    (L.SWE, 1, 100, L.SSPRK2, False, 20)])       # 20,000 elements, walls
def test_parity_many_blocks(physics, p, n, integrator, periodic, nsteps):
    """meshes of thousands of elements: many blocks (and partial ones) in
    every launch, the XCD remap of blocks to element ranges and the
    multi-block step reduction (CFL dt from every block's partial)"""
    mesh = unit_mesh(n, periodic=periodic)
    eps = _eps(mesh) if physics == L.MAXWELL_TE else None
    s, o = make_pair(physics, p, mesh, integrator=integrator, eps_elem=eps)
    s.project_ic()
    o.project_ic()
    ts, dts = s.step(nsteps)
    rc, to, dto = o.run(nsteps)
    assert rc == 0
    assert abs(ts - to) <= 1e-13 * abs(to)
    assert abs(dts - dto) <= 1e-13 * abs(dto)
    r = rel_l2(s.coeffs(), o.coeffs())
    print("physics", physics, "p", p, "K", mesh.K, "rel", r)
    assert r < TOL
