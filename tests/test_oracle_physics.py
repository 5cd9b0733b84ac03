"""Pins the oracle's shallow-water, Maxwell and CFL physics against what the
mathematics fixes (the reference ships no golden vectors for them).

- shallow water: the Roe flux reduces to the exact upwind flux when every
  wave moves one way (this needs the Roe property sum_k lambda_k a_k r_k =
  (F(UR) - F(UL)).n), and it passes a stationary hydraulic jump built from
  the Rankine-Hugoniot relation h_r / h_l = (sqrt(1 + 8 Fr^2) - 1) / 2 with
  no dissipation (a flux with G h^2 in place of G h^2 / 2, or any wave
  strength that does not split the jump exactly, fails it); a translating
  steady vortex in cyclostrophic balance (u_theta^2 / r = G dh/dr) converges
  at about p + 1 with the LLF and the Roe flux
- the LLF speed of Euler, shallow water and Maxwell TM equals the larger
  spectral radius of the two states' normal-flux Jacobians (computed here by
  differences of the oracle's physical flux, whose form the convergence
  tests pin)
- the CFL speed of Euler and shallow water (eval_global_lambda,
  euler.cu:248-277, shallowwater.cu) equals the spectral radius of the flux
  Jacobian at the cell mean in the direction of the mean velocity, and the
  first step's dt is cfl min_r / (max lambda (2p + 1)) (time_integrator.cu:217)
- Maxwell TE: a PEC cavity eigenmode with constant eps != 1; Maxwell TM: a
  plane wave of the reference's system (U = (Hx, Hy, Ez), fluxes of
  maxwells.cu:67-84), both converging at about p + 1
"""
import math

import numpy as np
import pytest

from dgcuda_amd import mesh as M


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# ------------------------------------------------------------ helpers

def _quad_points(o, mesh):
    """physical quadrature points [n_q, K] of every element and the tables"""
    T = o.tables()
    V = mesh.V
    r, s = T["r1"][:, None], T["r2"][:, None]
    x = V[2] * r + V[4] * s + V[0] * (1 - r - s)
    y = V[3] * r + V[5] * s + V[1] * (1 - r - s)
    return x, y, T


def project(o, mesh, f):
    """L2 projection (init_conditions' formula c_i = sum_j w_j u(x_j) phi_i(r_j))
    of f(x, y) -> [N, n_q, K] onto the oracle's basis, computed here"""
    x, y, T = _quad_points(o, mesh)
    U = f(x, y)  # [N, n_q, K]
    return np.einsum("j,ij,njk->nik", T["w"], T["basis"], U)


def l2_error(o, mesh, C, f):
    """per-variable L2 error of the coefficients C against f, by the
    element quadrature (w sums to 1/2, so J w integrates over the element)"""
    x, y, T = _quad_points(o, mesh)
    U = f(x, y)
    Uh = np.einsum("ij,nik->njk", T["basis"], C)
    J = o.geometry()["J"]
    return np.sqrt(np.einsum("j,k,njk->n", T["w"], J, (U - Uh) ** 2))


def normal_flux(o, U, nx, ny):
    fx, fy = o.flux_phys(U)
    return np.asarray(fx) * nx + np.asarray(fy) * ny


def jac_radius(o, U, nx, ny, h=1e-6):
    """spectral radius of d(F.n)/dU by central differences"""
    U = np.asarray(U, float)
    N = len(U)
    A = np.empty((N, N))
    for j in range(N):
        d = np.zeros(N)
        d[j] = h * max(1.0, abs(U[j]))
        A[:, j] = (normal_flux(o, U + d, nx, ny) - normal_flux(o, U - d, nx, ny)) / (2 * d[j])
    return np.max(np.abs(np.linalg.eigvals(A)))


def rate(e0, e1):
    return math.log(e0 / e1) / math.log(2.0)


SWE, EULER, TM, TE = 2, 1, 3, 4
LLF, ROE = 1, 3


def _swe_state(h, u, v):
    return np.array([h, h * u, h * v])


# ------------------------------------------------------------ Roe flux

@pytest.mark.parametrize("G", [1.0, 9.81])
def test_roe_is_upwind_when_every_wave_moves_one_way(O, G):
    rng = np.random.default_rng(11)
    o = O.Oracle(SWE, 1, flux=ROE, g=G)
    for _ in range(200):
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = math.cos(th), math.sin(th)
        hl, hr = rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)
        c = math.sqrt(G * max(hl, hr))
        sgn = rng.choice([-1.0, 1.0])
        # normal speeds above 2c in one direction, any tangential speed
        unl, unr = sgn * rng.uniform(2, 4) * c, sgn * rng.uniform(2, 4) * c
        utl, utr = rng.uniform(-2, 2, 2)
        UL = _swe_state(hl, unl * nx - utl * ny, unl * ny + utl * nx)
        UR = _swe_state(hr, unr * nx - utr * ny, unr * ny + utr * nx)
        F = o.flux_point(UL, UR, nx, ny)
        want = normal_flux(o, UL if sgn > 0 else UR, nx, ny)
        assert np.allclose(F, want, rtol=1e-13, atol=1e-13 * np.abs(want).max()), (F, want)


@pytest.mark.parametrize("G", [1.0, 9.81])
def test_roe_passes_a_stationary_hydraulic_jump(O, G):
    """a shock at rest (Rankine-Hugoniot with speed 0) is an eigenvector of
    the Roe matrix with eigenvalue 0: the Roe flux is F(UL).n = F(UR).n"""
    rng = np.random.default_rng(12)
    o = O.Oracle(SWE, 1, flux=ROE, g=G)
    for _ in range(100):
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = math.cos(th), math.sin(th)
        hl = rng.uniform(0.3, 1.5)
        Fr = rng.uniform(1.2, 4.0)  # supercritical upstream
        un_l = Fr * math.sqrt(G * hl)
        hr = hl * 0.5 * (math.sqrt(1 + 8 * Fr * Fr) - 1)
        un_r = hl * un_l / hr
        ut = rng.uniform(-1, 1)  # tangential velocity, continuous across the jump
        sgn = rng.choice([-1.0, 1.0])  # flow along n or against it (left/right swapped)
        a = _swe_state(hl, sgn * un_l * nx - ut * ny, sgn * un_l * ny + ut * nx)
        b = _swe_state(hr, sgn * un_r * nx - ut * ny, sgn * un_r * ny + ut * nx)
        UL, UR = (a, b) if sgn > 0 else (b, a)
        FL, FR = normal_flux(o, UL, nx, ny), normal_flux(o, UR, nx, ny)
        scale = np.abs(FL).max()
        assert np.allclose(FL, FR, rtol=0, atol=1e-12 * scale), "Rankine-Hugoniot"
        F = o.flux_point(UL, UR, nx, ny)
        assert np.allclose(F, FL, rtol=0, atol=1e-12 * scale), (F, FL)


def test_roe_and_llf_consistency_and_symmetry(O):
    """F(U, U) = F(U).n; F(UL, UR, n) = -F(UR, UL, -n) (conservation across a side)"""
    rng = np.random.default_rng(13)
    for flux in (LLF, ROE):
        o = O.Oracle(SWE, 1, flux=flux, g=9.81)
        for _ in range(50):
            th = rng.uniform(0, 2 * np.pi)
            nx, ny = math.cos(th), math.sin(th)
            UL = _swe_state(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 2))
            UR = _swe_state(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 2))
            assert np.allclose(o.flux_point(UL, UL, nx, ny), normal_flux(o, UL, nx, ny),
                               rtol=1e-14, atol=1e-14)
            assert np.allclose(o.flux_point(UL, UR, nx, ny), -o.flux_point(UR, UL, -nx, -ny),
                               rtol=1e-13, atol=1e-14)


# ------------------------------------------------------------ LLF and CFL speeds

@pytest.mark.parametrize("phys,states", [
    (EULER, lambda rng: np.array([rng.uniform(0.5, 2), *rng.uniform(-1, 1, 2),
                                  rng.uniform(2, 4)])),
    (SWE, lambda rng: _swe_state(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 2))),
    (TM, lambda rng: rng.uniform(-1, 1, 3)),
])
def test_llf_speed_is_the_spectral_radius(O, phys, states):
    rng = np.random.default_rng(14)
    o = O.Oracle(phys, 1, flux=LLF, g=9.81, gamma=1.4, eps=2.0, mu=1.5)
    for _ in range(40):
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = math.cos(th), math.sin(th)
        UL, UR = states(rng), states(rng)
        F = o.flux_point(UL, UR, nx, ny)
        central = 0.5 * (normal_flux(o, UL, nx, ny) + normal_flux(o, UR, nx, ny))
        lam = 2 * (F - central) / (UL - UR)  # F = central + lam (UL - UR) / 2
        want = max(jac_radius(o, UL, nx, ny), jac_radius(o, UR, nx, ny))
        assert np.allclose(lam, want, rtol=1e-7), (phys, lam, want)


@pytest.mark.parametrize("phys,prob,prm", [
    (EULER, 2, (5.0, 0.5, 0.5, 1.0, 1.0, 1.0)),   # isentropic vortex
    (SWE, 5, (1.0, 0.3, 0.4, 0.55, 0.15, 1.0)),   # bump
])
def test_cfl_speed_and_dt(O, phys, prob, prm):
    m = M.structured(6, periodic=phys == EULER, jitter=0.2, seed=5, bc=(-1, -1, -1, -1))
    p, cfl = 2, 0.3
    o = O.Oracle(phys, p, integrator=4, cfl=cfl, g=9.81, problem=prob,
                 problem_params=prm).set_mesh(m)
    o.project_ic()
    C = o.coeffs()
    lam = o.lam()
    phi0 = o.tables()["basis"][0, 0]
    rng = np.random.default_rng(15)
    for e in range(m.K):
        U = C[:, 0, e] * phi0  # cell mean
        vel = U[1:3] / U[0]
        sp = np.hypot(*vel)
        nx, ny = vel / sp if sp > 0 else (1.0, 0.0)  # at rest every direction is alike
        want = jac_radius(o, U, nx, ny)
        assert abs(lam[e] - want) <= 1e-7 * want, (e, lam[e], want)
        th = rng.uniform(0, 2 * np.pi)  # no direction has a larger radius
        assert jac_radius(o, U, math.cos(th), math.sin(th)) <= lam[e] * (1 + 1e-7)
    rc, t, dt = o.run(1)
    assert rc == 0
    assert abs(dt - cfl * o.min_r() / lam.max() / (2 * p + 1)) <= 1e-14 * dt


# ------------------------------------------------------------ exact solutions

def swe_vortex(G, h0, A, sig, x0, y0, U0, V0, L=1.0):
    """translating steady vortex of the shallow-water equations: u_theta =
    A s exp(-s^2/2) (s = r / sig) balanced by h = h0 - A^2/(2G) exp(-s^2),
    carried by the uniform velocity (U0, V0) on the periodic [0, L]^2"""
    def f(t):
        def u(x, y):
            dx = (x - x0 - U0 * t + L / 2) % L - L / 2
            dy = (y - y0 - V0 * t + L / 2) % L - L / 2
            s2 = (dx * dx + dy * dy) / sig ** 2
            g = A * np.exp(-s2 / 2) / sig
            h = h0 - A * A / (2 * G) * np.exp(-s2)
            return np.stack([h, h * (U0 - g * dy), h * (V0 + g * dx)])
        return u
    return f


@pytest.mark.parametrize("flux", [LLF, ROE])
@pytest.mark.parametrize("p,ns,min_rate", [(1, (24, 48), 1.7), (2, (24, 48), 2.5)])
def test_swe_moving_vortex_converges(O, flux, p, ns, min_rate):
    G, T = 1.0, 0.1
    ex = swe_vortex(G, 1.0, 0.3, 0.1, 0.5, 0.5, 0.4, 0.2)
    errs = []
    for n in ns:
        m = M.structured(n, periodic=True, jitter=0.2, seed=n)
        o = O.Oracle(SWE, p, integrator=4, flux=flux, cfl=0.3, g=G, problem=7,
                     problem_params=(1, 0, 0), threads=8).set_mesh(m)
        o.set_coeffs(project(o, m, ex(0.0)))
        rc, t, steps, _, _ = o.run_until(t_end=T)
        assert rc == 0 and t == T
        errs.append(l2_error(o, m, o.coeffs(), ex(T)))
    rates = [rate(a, b) for a, b in zip(errs[0], errs[1])]
    assert min(rates) >= min_rate, (errs, rates)
    assert errs[1].max() < 2e-3


def te_cavity_mode(eps, mu, kx, ky):
    """TE mode (Ex, Ey, Hz) of the PEC unit square with constant eps, mu:
    Hz = cos(kx x) cos(ky y) cos(w t), w = |k| / sqrt(eps mu)"""
    w = math.hypot(kx, ky) / math.sqrt(eps * mu)

    def f(t):
        def u(x, y):
            return np.stack([-ky / (eps * w) * np.cos(kx * x) * np.sin(ky * y) * math.sin(w * t),
                             kx / (eps * w) * np.sin(kx * x) * np.cos(ky * y) * math.sin(w * t),
                             np.cos(kx * x) * np.cos(ky * y) * math.cos(w * t)])
        return u
    return f


@pytest.mark.parametrize("p,min_rate", [(1, 1.7), (2, 2.7), (3, 3.6)])
def test_te_cavity_mode_converges(O, p, min_rate):
    eps, T = 2.5, 0.6
    ex = te_cavity_mode(eps, 1.0, math.pi, 2 * math.pi)
    errs = []
    for n in (6, 12):
        m = M.structured(n, periodic=False, jitter=0.2, seed=n, bc=(-1, -1, -1, -1))
        o = O.Oracle(TE, p, integrator=1, cfl=0.5, problem=7, problem_params=(0, 0, 0),
                     threads=8).set_mesh(m)
        o.set_eps(np.full(m.K, eps))
        o.set_coeffs(project(o, m, ex(0.0)))
        rc, t, _, _, _ = o.run_until(t_end=T)
        assert rc == 0 and t == T
        errs.append(l2_error(o, m, o.coeffs(), ex(T)))
    rates = [rate(a, b) for a, b in zip(errs[0], errs[1])]
    assert min(rates) >= min_rate, (errs, rates)


def tm_plane_wave(eps, mu):
    """plane wave of the reference's TM system along n = (1, 1)/sqrt 2 on the
    periodic unit square: U = r sin(2 pi (x + y) - w t) with r the
    eigenvector (-ny/mu, nx/mu, c) of the normal-flux matrix for the speed
    c = 1/sqrt(eps mu), w = 2 pi sqrt 2 c"""
    c = 1 / math.sqrt(eps * mu)
    nx = ny = 1 / math.sqrt(2)
    r = np.array([-ny / mu, nx / mu, c])
    w = 2 * math.pi * math.sqrt(2) * c

    def f(t):
        def u(x, y):
            return r[:, None, None] * np.sin(2 * math.pi * (x + y) - w * t)[None]
        return u
    return f


@pytest.mark.parametrize("p,min_rate", [(1, 1.7), (2, 2.7)])
def test_tm_plane_wave_converges(O, p, min_rate):
    eps, mu, T = 2.0, 1.5, 0.3
    ex = tm_plane_wave(eps, mu)
    errs = []
    for n in (8, 16):
        m = M.structured(n, periodic=True, jitter=0.2, seed=n)
        o = O.Oracle(TM, p, integrator=1, cfl=0.5, eps=eps, mu=mu, problem=7,
                     problem_params=(0, 0, 0), threads=8).set_mesh(m)
        o.set_coeffs(project(o, m, ex(0.0)))
        rc, t, _, _, _ = o.run_until(t_end=T)
        assert rc == 0 and t == T
        errs.append(l2_error(o, m, o.coeffs(), ex(T)))
    rates = [rate(a, b) for a, b in zip(errs[0], errs[1])]
    assert min(rates) >= min_rate, (errs, rates)
    # the opposite sign convention (r with its H part negated) is not a solution
    o2 = O.Oracle(TM, p, integrator=1, cfl=0.5, eps=eps, mu=mu, problem=7,
                  problem_params=(0, 0, 0)).set_mesh(m)
    flip = lambda t: (lambda x, y: ex(t)(x, y) * np.array([-1, -1, 1])[:, None, None])
    o2.set_coeffs(project(o2, m, flip(0.0)))
    o2.run_until(t_end=T)
    assert l2_error(o2, m, o2.coeffs(), flip(T)).max() > 100 * errs[1].max()
