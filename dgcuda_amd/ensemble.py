"""Ensembles of independent meshes stepped together (include/dg.h, dg_ens_*).

No counterpart in the reference, which runs one mesh per process.  Each
instance has its own mesh, order p (0..5), problem parameters and CFL dt;
one RK stage of the whole ensemble is one kernel launch.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import DgConfig, DgEnsInfo, DgError, ptr


class Ensemble:
    def __init__(self, physics=_lib.SWE, integrator=_lib.SSPRK3, flux=_lib.FLUX_DEFAULT,
                 positivity=True, cfl=0.0, gamma=1.4, g=9.8, eps=1.0, mu=1.0, a=(1.0, 1.0),
                 problem=_lib.PROB_SWE_BUMP, problem_params=(), device=0):
        self.lib = _lib.load()
        cfg = DgConfig()
        self.lib.dg_config_default(ctypes.byref(cfg))
        cfg.physics, cfg.p, cfg.flux, cfg.integrator = physics, 1, flux, integrator
        cfg.limiter, cfg.positivity, cfg.cfl = 0, int(positivity), cfl
        cfg.gamma, cfg.g, cfg.eps, cfg.mu = gamma, g, eps, mu
        cfg.ax, cfg.ay = a
        cfg.problem = problem
        for i in range(8):
            cfg.problem_params[i] = float(problem_params[i]) if i < len(problem_params) else 0.0
        cfg.device = device
        e = ctypes.c_void_p()
        rc = self.lib.dg_ens_create(ctypes.byref(e), ctypes.byref(cfg))
        if rc != 0:
            raise DgError(rc, "dg_ens_create rejected the configuration")
        self.e = e
        self.physics = physics
        self.N = _lib.NVARS[physics]
        self.shapes = []

    def _chk(self, rc):
        if rc != 0:
            msg = self.lib.dg_ens_last_error(self.e)
            raise DgError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "e", None):
            self.lib.dg_ens_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, mesh, p, problem_params=None):
        """one instance; returns its index"""
        arrs = [np.ascontiguousarray(mesh.V, np.float64),
                np.ascontiguousarray(mesh.elem_sides, np.int32),
                np.ascontiguousarray(mesh.side_xy, np.float64),
                np.ascontiguousarray(mesh.left, np.int32),
                np.ascontiguousarray(mesh.right, np.int32),
                np.ascontiguousarray(mesh.lsn, np.int32),
                np.ascontiguousarray(mesh.rsn, np.int32)]
        V, es, sxy, L, R, ls, rs = arrs
        prm = None
        if problem_params is not None:
            prm = np.zeros(8)
            prm[:len(problem_params)] = problem_params
        idx = ctypes.c_int()
        self._chk(self.lib.dg_ens_add(self.e, int(p), ptr(prm), mesh.K, ptr(V), ptr(es), mesh.S,
                                      ptr(sxy), ptr(L), ptr(R), ptr(ls), ptr(rs),
                                      ctypes.byref(idx)))
        self.shapes.append((self.N, (p + 1) * (p + 2) // 2, mesh.K))
        return idx.value

    def build(self, block=0, split=False):
        """block 64 / 128 / 256 elements per block, one stage launch for all
        order groups or (split) one per group; block <= 0 times every
        candidate and keeps the fastest"""
        b = int(block)
        if b > 0 and split:
            b |= _lib.ENS_SPLIT
        self._chk(self.lib.dg_ens_build(self.e, b))
        return self

    def step(self, nsteps, dt=None):
        self._chk(self.lib.dg_ens_step(self.e, int(nsteps), float(dt or 0.0)))

    def coeffs(self, i):
        out = np.empty(self.shapes[i])
        self._chk(self.lib.dg_ens_get_coeffs(self.e, int(i), ptr(out), _lib.HOST))
        return out

    def set_coeffs(self, i, C):
        C = np.ascontiguousarray(C, np.float64).reshape(self.shapes[i])
        self._chk(self.lib.dg_ens_set_coeffs(self.e, int(i), ptr(C), _lib.HOST))

    def state(self, i):
        """(t, dt of the last step, steps taken)"""
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._chk(self.lib.dg_ens_state(self.e, int(i), ctypes.byref(t), ctypes.byref(dt),
                                        ctypes.byref(n)))
        return t.value, dt.value, n.value

    def l2_error(self, i, t):
        err = np.empty(self.N)
        self._chk(self.lib.dg_ens_l2_error(self.e, int(i), float(t), ptr(err)))
        return err

    def info(self):
        inf = DgEnsInfo()
        self._chk(self.lib.dg_ens_info(self.e, ctypes.byref(inf)))
        return inf
