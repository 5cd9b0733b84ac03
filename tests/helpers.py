"""Shared builders for the parity tests: the same configuration is set up on
the HIP library (through the C ABI) and on the oracle."""
import numpy as np

from dgcuda_amd import mesh as M
from dgcuda_amd import _lib as L

# physics -> (problem, params, flux) used by the parity cases
CASES = {
    L.ADVECTION: (L.PROB_ADV_SINSIN, (), L.FLUX_DEFAULT),
    L.EULER: (L.PROB_EULER_VORTEX, (5.0, 0.5, 0.5, 1.0, 1.0, 1.0), L.FLUX_DEFAULT),
    L.SWE: (L.PROB_SWE_BUMP, (1.0, 0.1, 0.4, 0.55, 0.15, 1.0), L.FLUX_DEFAULT),
    L.MAXWELL_TM: (L.PROB_MAXWELL_PULSE, (1.0, 0.5, 0.45, 0.2), L.FLUX_DEFAULT),
    L.MAXWELL_TE: (L.PROB_MAXWELL_PULSE, (1.0, 0.5, 0.45, 0.2), L.FLUX_DEFAULT),
}


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_pair(physics, p, mesh, integrator=L.RK4, flux=None, problem=None, params=None,
              cfl=0.0, limiter=False, g=9.8, eps_elem=None, vortex_scale=1.0, volume=0):
    """(Solver, Oracle) configured identically; vortex parameters are scaled
    for the mesh's box when vortex_scale != 1."""
    from dgcuda_amd.solver import Solver
    import oracle as O
    prob, prm, fl = CASES[physics]
    if problem is not None:
        prob = problem
    if params is not None:
        prm = params
    if flux is not None:
        fl = flux
    if cfl == 0.0:
        cfl = 0.7 if integrator == L.RK2MID else 0.5
    s = Solver(physics, p, integrator=integrator, flux=fl, limiter=limiter, cfl=cfl, g=g,
               problem=prob, problem_params=prm, volume=volume)
    s.set_mesh(mesh)
    o = O.Oracle(physics, p, integrator=integrator, flux=fl, cfl=cfl, limiter=limiter,
                 posfix=True, g=g, problem=prob, problem_params=prm)
    o.set_mesh(mesh)
    if eps_elem is not None:
        s.set_eps(eps_elem)
        o.set_eps(eps_elem)
    return s, o


def unit_mesh(n, periodic=True, jitter=0.2, seed=7, bc=(-1, -1, -1, -1)):
    return M.structured(n, periodic=periodic, jitter=jitter, seed=seed, bc=bc)
