"""Host-side mirror of the reference's driver API on top of libdgamd.so.

The reference drives the path through run_dgcuda (src/2d/conservation/
main.cu:600-974): read_mesh -> init_gpu -> set_quadrature/preval_* ->
init_conditions -> time_integrate_rk4/rk2 -> eval_error_L2 -> write_U.
Solver keeps that order with explicit calls; every call goes through the
C ABI (include/dg.h) and raises DgError with the library's message on
failure.
"""
import ctypes
import warnings
import math

import numpy as np

from . import _lib
from ._lib import DgConfig, DgError, DgInfo, ptr


def _torch():
    import torch
    return torch

_DEFAULT_CFL = {_lib.RK4: 0.5, _lib.RK2MID: 0.7}


def _is_tensor(x):
    return type(x).__module__.split(".")[0] == "torch"


class Solver:
    def __init__(self, physics=_lib.EULER, p=1, integrator=_lib.RK4, flux=_lib.FLUX_DEFAULT,
                 limiter=False, positivity=True, cfl=0.0, gamma=1.4, g=9.8, eps=1.0, mu=1.0,
                 a=(1.0, 1.0), problem=_lib.PROB_EULER_CONST, problem_params=(1, 1, 1, 1),
                 device=0, volume=_lib.VOLUME_DEFAULT):
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
        cfg.volume = volume
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

    # -- array arguments of the C ABI
    def _in(self, x, count, name):
        """(address, DG_HOST / DG_DEVICE, object to keep alive) of an input
        of `count` doubles: numpy arrays, sequences and CPU tensors take the
        host path (copied to a contiguous float64 array if needed); CUDA
        tensors must be float64 and on this solver's device, and are passed
        as device pointers (made contiguous if they are not)."""
        if _is_tensor(x):
            if x.dtype != _torch().float64:
                raise TypeError("%s: tensor must be float64, got %s" % (name, x.dtype))
            if x.numel() != count:
                raise ValueError("%s: %d elements, expected %d" % (name, x.numel(), count))
            if x.device.type == "cuda":
                self._check_device(x, name)
                x = x.contiguous()
                return x.data_ptr(), _lib.DEVICE, x
            x = x.detach().contiguous().numpy()
        a = np.ascontiguousarray(x, np.float64)
        if a.size != count:
            raise ValueError("%s: %d elements, expected %d" % (name, a.size, count))
        return a.ctypes.data, _lib.HOST, a

    def _out(self, out, count, name):
        """(address, DG_HOST / DG_DEVICE) of an output of `count` doubles,
        written in place: a C-contiguous float64 numpy array, a contiguous
        float64 CPU tensor, or a contiguous float64 CUDA tensor on this
        solver's device"""
        if _is_tensor(out):
            if out.dtype != _torch().float64:
                raise TypeError("%s: tensor must be float64, got %s" % (name, out.dtype))
            if out.numel() != count or not out.is_contiguous():
                raise ValueError("%s: need a contiguous tensor of %d elements" % (name, count))
            if out.device.type == "cuda":
                self._check_device(out, name)
                return out.data_ptr(), _lib.DEVICE
            return out.data_ptr(), _lib.HOST
        if not isinstance(out, np.ndarray) or out.dtype != np.float64 or \
                not out.flags["C_CONTIGUOUS"] or out.size != count:
            raise ValueError("%s: need a C-contiguous float64 array of %d elements" % (name, count))
        return out.ctypes.data, _lib.HOST

    def _check_device(self, t, name):
        if t.device.index is not None and t.device.index != self.cfg.device:
            raise ValueError("%s: tensor on cuda:%d, solver on device %d"
                             % (name, t.device.index, self.cfg.device))

    def _empty(self, like, shape):
        """a new output of `shape` on the same side of the boundary as `like`"""
        if _is_tensor(like) and like.device.type == "cuda":
            return _torch().empty(shape, dtype=_torch().float64, device=like.device)
        return np.empty(shape)

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
        """per-element permittivity [K] (host data; a CUDA tensor is copied
        to the host, K doubles)"""
        if _is_tensor(eps):
            eps = eps.detach().to("cpu", _torch().float64)
        e = np.ascontiguousarray(eps, np.float64)
        if e.size != self.K:
            raise ValueError("eps: %d elements, expected K = %d" % (e.size, self.K))
        self._chk(self.lib.dg_set_eps(self.h, ptr(e)))

    @property
    def shape(self):
        return (self.N, self.n_p, self.K)

    def project_ic(self):
        self._chk(self.lib.dg_project_ic(self.h))
        return self

    @property
    def size(self):
        return self.N * self.n_p * self.K

    def set_coeffs(self, C):
        """the state [N, n_p, K]: host data, or a CUDA tensor copied on the device"""
        a, where, _keep = self._in(C, self.size, "set_coeffs")
        self._chk(self.lib.dg_set_coeffs(self.h, a, where))

    def coeffs(self, out=None):
        """Modal coefficients [N, n_p, K] (the reference's SoA order), into a
        new numpy array or into `out` (numpy array, CPU or CUDA tensor)."""
        if out is None:
            out = np.empty(self.shape)
        a, where = self._out(out, self.size, "coeffs")
        self._chk(self.lib.dg_get_coeffs(self.h, a, where))
        return out

    # -- the right-hand side and time stepping
    def eval_rhs(self, u, t=0.0, dt=1.0, out=None):
        """k = dt * L(u) (eval_surface + eval_volume + eval_rhs).  A CUDA
        tensor u gives a CUDA tensor k, computed with no host copy; host
        data gives a numpy array."""
        ua, uw, _keep = self._in(u, self.size, "eval_rhs u")
        if out is None:
            out = self._empty(u, self.shape)
        oa, ow = self._out(out, self.size, "eval_rhs out")
        if uw != ow:
            raise ValueError("eval_rhs: u and out must both be host data or both CUDA tensors")
        self._chk(self.lib.dg_eval_rhs(self.h, ua, float(t), float(dt), oa, uw))
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
            w, where, _keep = self._in(weights, self.size, "objective weights")
            self._chk(self.lib.dg_adjoint_set_objective(self.h, kind, ptr(prm) if prm.size else None,
                                                        prm.size, w, where))
        else:
            self._chk(self.lib.dg_adjoint_set_objective(self.h, kind, ptr(prm) if prm.size else None,
                                                        prm.size, None, _lib.HOST))

    def adjoint_solve(self, nsteps, dt, every=0, want_eps=False, device=None):
        """forward nsteps of fixed dt from the current state, then the reverse
        sweep: returns (J, dJ/du0 [N, n_p, K], dJ/deps [K] or None).  With
        device (a torch device) the gradients are CUDA tensors written on the
        device; otherwise numpy arrays."""
        J = ctypes.c_double()
        if device is not None:
            torch = _torch()
            gu = torch.empty(self.shape, dtype=torch.float64, device=device)
            ge = torch.empty(self.K, dtype=torch.float64, device=device) if want_eps else None
        else:
            gu = np.empty(self.shape)
            ge = np.empty(self.K) if want_eps else None
        ga, where = self._out(gu, self.size, "grad_u0")
        ea = self._out(ge, self.K, "grad_eps")[0] if want_eps else None
        self._chk(self.lib.dg_adjoint_solve(self.h, int(nsteps), float(dt), int(every),
                                            ctypes.byref(J), ga, ea, where))
        if self.info().adj_fix_fired:
            warnings.warn("dg_adjoint_solve: the positivity fix fired in the forward map; the "
                          "gradient takes it as inactive and is not exact there",
                          RuntimeWarning, stacklevel=2)
        return J.value, gu, ge

    def adjoint_rhs(self, u, gk, t=0.0, dt=1.0):
        """(dk/du)^T gk for k = dt * L(u); CUDA tensors in give a CUDA tensor
        out (no host copy), host data a numpy array"""
        ua, uw, _k1 = self._in(u, self.size, "adjoint_rhs u")
        ka, kw, _k2 = self._in(gk, self.size, "adjoint_rhs gk")
        if uw != kw:
            raise ValueError("adjoint_rhs: u and gk must both be host data or both CUDA tensors")
        gu = self._empty(u, self.shape)
        oa, _ = self._out(gu, self.size, "adjoint_rhs out")
        self._chk(self.lib.dg_adjoint_rhs(self.h, ua, float(t), float(dt), ka, oa, uw))
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

    # -- user boundary states (the U_inflow / U_outflow / U_reflection hooks)
    def boundary_points(self):
        """the mesh's boundary sides and their quadrature points: (side [nb],
        code [nb] (-1 wall, -2 outflow, -3 inflow), x [nb, n_q1], y [nb, n_q1])"""
        nb = ctypes.c_int()
        self._chk(self.lib.dg_boundary_points(self.h, ctypes.byref(nb), None, None, None, None))
        n, q = nb.value, self.p + 1
        side, code = np.empty(n, np.int32), np.empty(n, np.int32)
        x, y = np.empty((n, q)), np.empty((n, q))
        self._chk(self.lib.dg_boundary_points(self.h, ctypes.byref(nb), ptr(side), ptr(code),
                                              ptr(x), ptr(y)))
        return side, code, x, y

    def set_boundary_states(self, U):
        """[nb, n_q1, N]: outside states at outflow / inflow points, the
        reflection direction in entries 0, 1 at wall points; None restores
        the problem's boundary data"""
        if U is None:
            self._chk(self.lib.dg_set_boundary_states(self.h, None, _lib.HOST))
            return
        nb = ctypes.c_int()
        self._chk(self.lib.dg_boundary_points(self.h, ctypes.byref(nb), None, None, None, None))
        a, where, _keep = self._in(U, nb.value * (self.p + 1) * self.N, "boundary states")
        self._chk(self.lib.dg_set_boundary_states(self.h, a, where))

    def boundary_states(self, t=0.0):
        """the boundary data in use [nb, n_q1, N] (the caller's, or the problem's at t)"""
        nb = ctypes.c_int()
        self._chk(self.lib.dg_boundary_points(self.h, ctypes.byref(nb), None, None, None, None))
        out = np.empty((nb.value, self.p + 1, self.N))
        self._chk(self.lib.dg_get_boundary_states(self.h, float(t), ptr(out)))
        return out

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
