"""Python binding of the oracle (oracle/dg_oracle.c).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference path that the
HIP library is checked against.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

ADVECTION, EULER, SWE, MAXWELL_TM, MAXWELL_TE = range(5)
NVARS = {ADVECTION: 1, EULER: 4, SWE: 3, MAXWELL_TM: 3, MAXWELL_TE: 3}

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def build_baseline(out_dir, cflags="-O3 -march=native -fPIC -fopenmp -ffp-contract=off"):
    """The oracle compiled for timing on this host (the CPU baseline's build:
    -O3 -march=native, still -ffp-contract=off, so its arithmetic is the
    checker's) into out_dir; returns the library path."""
    r = subprocess.run(["make", "-s", "-B", "-C", HERE, "OUT=" + os.path.abspath(out_dir),
                        "CFLAGS=" + cflags], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle baseline build failed:\n" + r.stdout + r.stderr)
    return os.path.join(os.path.abspath(out_dir), "liboracle.so")


def load(path=None):
    """the checker's library (oracle/_build), or the library at `path`"""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is None and not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(path or LIB_PATH)
    P, I, D, VP = ctypes.POINTER, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    sig = {
        "orc_new": (VP, [I, I, I, I, D, I, I]),
        "orc_free": (None, [VP]),
        "orc_set_threads": (None, [I]),
        "orc_set_params": (I, [VP, D, D, D, D, D, D]),
        "orc_set_problem": (I, [VP, I, VP, I]),
        "orc_set_mesh": (I, [VP, I, VP, VP, I, VP, VP, VP, VP, VP]),
        "orc_set_eps": (I, [VP, VP]),
        "orc_project_ic": (I, [VP]),
        "orc_set_coeffs": (I, [VP, VP]),
        "orc_get_coeffs": (I, [VP, VP]),
        "orc_dofs": (I, [VP]),
        "orc_eval_rhs": (I, [VP, VP, D, D, VP]),
        "orc_run": (I, [VP, I, D, P(D), P(D)]),
        "orc_run_until": (I, [VP, I, D, D, D, D, P(D), P(I), P(D), P(D)]),
        "orc_l1_error": (I, [VP, D, VP]),
        "orc_eval_exact": (I, [VP, I, D, VP, VP, VP]),
        "orc_min_r": (D, [VP]),
        "orc_geometry": (I, [VP] + [VP] * 9),
        "orc_tables": (I, [VP] + [VP] * 10),
        "orc_nq": (I, [VP]),
        "orc_nq1": (I, [VP]),
        "orc_l2_error": (I, [VP, D, VP]),
        "orc_lambda": (I, [VP, VP]),
        "orc_limit": (I, [VP, VP]),
        "orc_eval_u": (I, [VP, I, VP, VP, VP]),
        "orc_flux_point": (I, [VP, VP, VP, D, D, D, D, VP]),
        "orc_flux_phys": (I, [VP, VP, VP, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """One configuration of the reference path on the CPU (SoA [N, n_p, K])."""

    def __init__(self, physics, p, integrator=1, flux=0, cfl=0.5, limiter=False, posfix=True,
                 gamma=1.4, g=9.8, eps=1.0, mu=1.0, a=(1.0, 1.0), problem=0, problem_params=(),
                 threads=0, lib=None):
        self.lib = lib or load()
        if threads:
            self.lib.orc_set_threads(int(threads))
        self.o = self.lib.orc_new(physics, p, flux, integrator, cfl, int(limiter), int(posfix))
        if not self.o:
            raise ValueError("oracle rejected the configuration")
        self.lib.orc_set_params(self.o, gamma, g, eps, mu, a[0], a[1])
        pp = np.zeros(8)
        pp[:len(problem_params)] = problem_params
        self._pp = pp
        self.lib.orc_set_problem(self.o, problem, _p(pp), 8)
        self.N = NVARS[physics]
        self.n_p = (p + 1) * (p + 2) // 2
        self.p = p
        self.K = 0

    def __del__(self):
        if getattr(self, "o", None):
            self.lib.orc_free(self.o)
            self.o = None

    def set_mesh(self, mesh):
        self._mesh = [np.ascontiguousarray(mesh.V, np.float64),
                      np.ascontiguousarray(mesh.elem_sides, np.int32),
                      np.ascontiguousarray(mesh.side_xy, np.float64),
                      np.ascontiguousarray(mesh.left, np.int32),
                      np.ascontiguousarray(mesh.right, np.int32),
                      np.ascontiguousarray(mesh.lsn, np.int32),
                      np.ascontiguousarray(mesh.rsn, np.int32)]
        V, es, sxy, L, R, ls, rs = self._mesh
        self.lib.orc_set_mesh(self.o, mesh.K, _p(V), _p(es), mesh.S, _p(sxy), _p(L), _p(R),
                              _p(ls), _p(rs))
        self.K, self.S = mesh.K, mesh.S
        return self

    @property
    def shape(self):
        return (self.N, self.n_p, self.K)

    def set_eps(self, eps):
        self._eps = np.ascontiguousarray(eps, np.float64)
        self.lib.orc_set_eps(self.o, _p(self._eps))

    def project_ic(self):
        self.lib.orc_project_ic(self.o)
        return self

    def set_coeffs(self, C):
        C = np.ascontiguousarray(C, np.float64)
        self.lib.orc_set_coeffs(self.o, _p(C))

    def coeffs(self):
        out = np.empty(self.shape)
        self.lib.orc_get_coeffs(self.o, _p(out))
        return out

    def eval_rhs(self, u, t=0.0, dt=1.0):
        u = np.array(u, np.float64, copy=True).reshape(self.shape)
        k = np.empty(self.shape)
        self.lib.orc_eval_rhs(self.o, _p(u), t, dt, _p(k))
        return k

    def run(self, nsteps, dt=None):
        t = ctypes.c_double()
        d = ctypes.c_double()
        rc = self.lib.orc_run(self.o, int(nsteps), float(dt or 0.0), ctypes.byref(t),
                              ctypes.byref(d))
        return rc, t.value, d.value

    def run_until(self, max_steps=-1, t_end=None, conv_tol=None, dt=None, t0=0.0):
        """the reference's loop with -T / -c stops: (rc, t, steps, conv, dt_last)"""
        t = ctypes.c_double()
        n = ctypes.c_int()
        cv = ctypes.c_double()
        d = ctypes.c_double()
        rc = self.lib.orc_run_until(self.o, int(max_steps), float(dt or 0.0), float(t0),
                                    float("inf") if t_end is None else float(t_end),
                                    -1.0 if conv_tol is None else float(conv_tol),
                                    ctypes.byref(t), ctypes.byref(n), ctypes.byref(cv),
                                    ctypes.byref(d))
        return rc, t.value, n.value, cv.value, d.value

    def min_r(self):
        return self.lib.orc_min_r(self.o)

    def geometry(self):
        K, S = self.K, self.S
        out = dict(J=np.empty(K), xr=np.empty(K), yr=np.empty(K), xs=np.empty(K), ys=np.empty(K),
                   len=np.empty(S), nx=np.empty(S), ny=np.empty(S), diam=np.empty(K))
        self.lib.orc_geometry(self.o, *[_p(out[k]) for k in
                                        ("J", "xr", "yr", "xs", "ys", "len", "nx", "ny", "diam")])
        return out

    def tables(self):
        np_, nq, nq1 = self.n_p, self.lib.orc_nq(self.o), self.lib.orc_nq1(self.o)
        out = dict(basis=np.empty((np_, nq)), gx=np.empty((np_, nq)), gy=np.empty((np_, nq)),
                   side=np.empty((3, np_, nq1)), vertex=np.empty((np_, 3)), w=np.empty(nq),
                   r1=np.empty(nq), r2=np.empty(nq), w1=np.empty(nq1), r1d=np.empty(nq1))
        self.lib.orc_tables(self.o, *[_p(out[k]) for k in
                                      ("basis", "gx", "gy", "side", "vertex", "w", "r1", "r2",
                                       "w1", "r1d")])
        return out

    def l2_error(self, t):
        err = np.empty(self.N)
        self.lib.orc_l2_error(self.o, t, _p(err))
        return err

    def lam(self):
        out = np.empty(self.K)
        self.lib.orc_lambda(self.o, _p(out))
        return out

    def limit(self, C):
        C = np.array(C, np.float64, copy=True)
        self.lib.orc_limit(self.o, _p(C))
        return C

    def flux_point(self, UL, UR, nx, ny, epsL=1.0, epsR=1.0):
        UL = np.zeros(4) + np.pad(np.asarray(UL, float), (0, 4 - len(UL)))
        UR = np.zeros(4) + np.pad(np.asarray(UR, float), (0, 4 - len(UR)))
        F = np.zeros(4)
        self.lib.orc_flux_point(self.o, _p(UL), _p(UR), nx, ny, epsL, epsR, _p(F))
        return F[:self.N]

    def flux_phys(self, U):
        U = np.zeros(4) + np.pad(np.asarray(U, float), (0, 4 - len(U)))
        fx, fy = np.zeros(4), np.zeros(4)
        self.lib.orc_flux_phys(self.o, _p(U), _p(fx), _p(fy))
        return fx[:self.N], fy[:self.N]

    def l1_error(self, t):
        err = np.empty(self.N)
        self.lib.orc_l1_error(self.o, t, _p(err))
        return err

    def eval_exact(self, n=0, t=0.0):
        u = [np.empty(self.K) for _ in range(3)]
        self.lib.orc_eval_exact(self.o, n, t, *[_p(x) for x in u])
        return u

    def eval_u(self, n=0):
        u = [np.empty(self.K) for _ in range(3)]
        self.lib.orc_eval_u(self.o, n, _p(u[0]), _p(u[1]), _p(u[2]))
        return u
