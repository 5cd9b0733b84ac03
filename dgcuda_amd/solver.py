"""Host-side mirror of the reference's driver API on top of libdgamd.so.

The reference drives the path through run_dgcuda (src/2d/conservation/
main.cu:600-974): read_mesh -> init_gpu -> set_quadrature/preval_* ->
init_conditions -> time_integrate_rk4/rk2 -> eval_error_L2 -> write_U.
Solver keeps that order with explicit calls; every call goes through the
C ABI (include/dg.h) and raises DgError with the library's message on
failure.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import DgConfig, DgError, DgInfo, ptr

_DEFAULT_CFL = {_lib.RK4: 0.5, _lib.RK2MID: 0.7}


class Solver:
    def __init__(self, physics=_lib.EULER, p=1, integrator=_lib.RK4, flux=_lib.FLUX_DEFAULT,
                 limiter=False, positivity=True, cfl=0.0, gamma=1.4, g=9.8, eps=1.0, mu=1.0,
                 a=(1.0, 1.0), problem=_lib.PROB_EULER_CONST, problem_params=(1, 1, 1, 1),
                 device=0):
        self.lib = _lib.load()
        cfg = DgConfig()
        self.lib.dg_config_default(ctypes.byref(cfg))
        cfg.physics, cfg.p, cfg.flux, cfg.integrator = physics, p, flux, integrator
        cfg.limiter, cfg.positivity, cfg.cfl = int(limiter), int(positivity), cfl
        cfg.gamma, cfg.g, cfg.eps, cfg.mu = gamma, g, eps, mu
        cfg.ax, cfg.ay = a
        cfg.problem = problem
        for i in range(8):
            cfg.problem_params[i] = float(problem_params[i]) if i < len(problem_params) else 0.0
        cfg.device = device
        h = ctypes.c_void_p()
        rc = self.lib.dg_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != 0:
            raise DgError(rc, "dg_create rejected the configuration")
        self.h = h
        self.cfg = cfg
        self.physics, self.p = physics, p
        self.N = _lib.NVARS[physics]
        self.n_p = (p + 1) * (p + 2) // 2
        self.K = 0

    # -- errors
    def _chk(self, rc):
        if rc != 0:
            msg = self.lib.dg_last_error(self.h)
            raise DgError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.dg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- mesh and state
    def set_mesh(self, mesh):
        arrs = [np.ascontiguousarray(mesh.V, np.float64),
                np.ascontiguousarray(mesh.elem_sides, np.int32),
                np.ascontiguousarray(mesh.side_xy, np.float64),
                np.ascontiguousarray(mesh.left, np.int32),
                np.ascontiguousarray(mesh.right, np.int32),
                np.ascontiguousarray(mesh.lsn, np.int32),
                np.ascontiguousarray(mesh.rsn, np.int32)]
        V, es, sxy, L, R, ls, rs = arrs
        self._chk(self.lib.dg_set_mesh(self.h, mesh.K, ptr(V), ptr(es), mesh.S, ptr(sxy), ptr(L),
                                       ptr(R), ptr(ls), ptr(rs)))
        self.K = mesh.K
        self.mesh = mesh
        return self

    def load_pmsh(self, path):
        self._chk(self.lib.dg_load_pmsh(self.h, str(path).encode()))
        self.K = self.info().K
        return self

    def set_eps(self, eps):
        e = np.ascontiguousarray(eps, np.float64)
        self._chk(self.lib.dg_set_eps(self.h, ptr(e)))

    @property
    def shape(self):
        return (self.N, self.n_p, self.K)

    def project_ic(self):
        self._chk(self.lib.dg_project_ic(self.h))
        return self

    def set_coeffs(self, C):
        if isinstance(C, np.ndarray):
            C = np.ascontiguousarray(C, np.float64).reshape(self.shape)
            self._chk(self.lib.dg_set_coeffs(self.h, ptr(C), _lib.HOST))
        else:
            self._chk(self.lib.dg_set_coeffs(self.h, ptr(C.contiguous()), _lib.DEVICE))

    def coeffs(self, out=None):
        """Modal coefficients [N, n_p, K] (the reference's SoA order)."""
        if out is None:
            out = np.empty(self.shape)
        if isinstance(out, np.ndarray):
            self._chk(self.lib.dg_get_coeffs(self.h, ptr(out), _lib.HOST))
        else:
            self._chk(self.lib.dg_get_coeffs(self.h, ptr(out), _lib.DEVICE))
        return out

    # -- the right-hand side and time stepping
    def eval_rhs(self, u, t=0.0, dt=1.0, out=None):
        """k = dt * L(u) (eval_surface + eval_volume + eval_rhs)."""
        if isinstance(u, np.ndarray):
            u = np.ascontiguousarray(u, np.float64).reshape(self.shape)
            out = np.empty(self.shape) if out is None else out
            self._chk(self.lib.dg_eval_rhs(self.h, ptr(u), t, dt, ptr(out), _lib.HOST))
        else:
            import torch
            out = torch.empty_like(u) if out is None else out
            self._chk(self.lib.dg_eval_rhs(self.h, ptr(u.contiguous()), t, dt, ptr(out),
                                           _lib.DEVICE))
        return out

    def step(self, nsteps, dt=None, t=0.0):
        """nsteps RK steps; dt=None uses the CFL rule.  Returns (t, dt_last)."""
        tt = ctypes.c_double(t)
        dd = ctypes.c_double(0.0)
        self._chk(self.lib.dg_step(self.h, int(nsteps), float(dt or 0.0), ctypes.byref(tt),
                                   ctypes.byref(dd)))
        return tt.value, dd.value

    def run(self, max_steps=-1, t_end=None, conv_tol=None, dt=None, t=0.0):
        """The reference's time loop: stop after max_steps steps (-1: no limit),
        at t_end (the last step shortened to land on it) or once the
        step-to-step change of the first K*N state entries is <= conv_tol.
        Returns (t, steps, last_change)."""
        tt = ctypes.c_double(t)
        ns = ctypes.c_int(0)
        cv = ctypes.c_double(0.0)
        self._chk(self.lib.dg_run(self.h, int(max_steps), math.inf if t_end is None else float(t_end),
                                  -1.0 if conv_tol is None else float(conv_tol), float(dt or 0.0),
                                  ctypes.byref(tt), ctypes.byref(ns), ctypes.byref(cv)))
        return tt.value, ns.value, cv.value

    def timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._chk(self.lib.dg_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- discrete adjoint
    def set_objective(self, kind, params=(), weights=None):
        """kind: _lib.OBJ_LINEAR (weights [N, n_p, K]), OBJ_ENTROPY (Euler),
        OBJ_PATCH_SQ (params x0, x1, y0, y1, variable)"""
        prm = np.ascontiguousarray(np.asarray(params, np.float64).ravel())
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, np.float64).reshape(self.shape))
            self._chk(self.lib.dg_adjoint_set_objective(self.h, kind, ptr(prm) if prm.size else None,
                                                        prm.size, ptr(w), _lib.HOST))
        else:
            self._chk(self.lib.dg_adjoint_set_objective(self.h, kind, ptr(prm) if prm.size else None,
                                                        prm.size, None, _lib.HOST))

    def adjoint_solve(self, nsteps, dt, every=0, want_eps=False):
        """forward nsteps of fixed dt from the current state, then the reverse
        sweep: returns (J, dJ/du0 [N, n_p, K], dJ/deps [K] or None)"""
        J = ctypes.c_double()
        gu = np.empty(self.shape)
        ge = np.empty(self.K) if want_eps else None
        self._chk(self.lib.dg_adjoint_solve(self.h, int(nsteps), float(dt), int(every),
                                            ctypes.byref(J), ptr(gu), ptr(ge), _lib.HOST))
        return J.value, gu, ge

    def adjoint_rhs(self, u, gk, t=0.0, dt=1.0):
        """(dk/du)^T gk for k = dt * L(u)"""
        u = np.ascontiguousarray(u, np.float64).reshape(self.shape)
        gk = np.ascontiguousarray(gk, np.float64).reshape(self.shape)
        gu = np.empty(self.shape)
        self._chk(self.lib.dg_adjoint_rhs(self.h, ptr(u), float(t), float(dt), ptr(gk), ptr(gu),
                                          _lib.HOST))
        return gu

    def objective(self):
        """the objective at the current coefficients"""
        J = ctypes.c_double()
        self._chk(self.lib.dg_adjoint_objective(self.h, ctypes.byref(J)))
        return J.value

    def adjoint_timing(self):
        f, r = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.dg_adjoint_timing(self.h, ctypes.byref(f), ctypes.byref(r)))
        return f.value, r.value

    # -- outputs
    def l2_error(self, t):
        err = np.empty(self.N)
        self._chk(self.lib.dg_l2_error(self.h, float(t), ptr(err)))
        return err

    def l1_error(self, t):
        err = np.empty(self.N)
        self._chk(self.lib.dg_l1_error(self.h, float(t), ptr(err)))
        return err

    def eval_exact(self, n=0, t=0.0):
        """variable n of the exact state at the vertices (plot_exact)"""
        u = [np.empty(self.K) for _ in range(3)]
        self._chk(self.lib.dg_eval_exact(self.h, n, float(t), ptr(u[0]), ptr(u[1]), ptr(u[2])))
        return u

    def eval_u(self, n=0):
        u = [np.empty(self.K) for _ in range(3)]
        self._chk(self.lib.dg_eval_u(self.h, n, ptr(u[0]), ptr(u[1]), ptr(u[2])))
        return u

    def lam(self):
        out = np.empty(self.K)
        self._chk(self.lib.dg_lambda(self.h, ptr(out)))
        return out

    def limit(self):
        self._chk(self.lib.dg_limit(self.h))

    def info(self):
        inf = DgInfo()
        self._chk(self.lib.dg_info(self.h, ctypes.byref(inf)))
        return inf

    def geometry(self):
        inf = self.info()
        J, diam = np.empty(inf.K), np.empty(inf.K)
        ln, nx, ny = np.empty(inf.S), np.empty(inf.S), np.empty(inf.S)
        self._chk(self.lib.dg_geometry(self.h, ptr(J), ptr(ln), ptr(nx), ptr(ny), ptr(diam)))
        return dict(J=J, len=ln, nx=nx, ny=ny, diam=diam, min_r=inf.min_r)

    def write_pos(self, path, n=0, view=None):
        """gmsh view of variable n at the element vertices (write_U main.cu:69-116)."""
        u1, u2, u3 = self.eval_u(n)
        V = self.mesh.V
        with open(path, "w") as f:
            f.write('View "U%i " {\n' % (n if view is None else view))
            for e in range(self.K):
                f.write("ST (%.015f,%.015f,0,%.015f,%.015f,0,%.015f,%.015f,0) {%.015f,%.015f,%.015f};\n"
                        % (V[0, e], V[1, e], V[2, e], V[3, e], V[4, e], V[5, e], u1[e], u2[e], u3[e]))
            f.write("};")
