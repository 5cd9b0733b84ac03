"""ctypes binding of libdgamd.so (the C ABI of include/dg.h).

The library is the product: nothing here falls back to another
implementation.  If libdgamd.so is missing, load() raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgamd.so")

DG_OK = 0
DG_ERR_ARG = -1
DG_ERR_STATE = -2
DG_ERR_HIP = -3
DG_ERR_IO = -4
DG_ERR_NAN = -5

ADVECTION, EULER, SWE, MAXWELL_TM, MAXWELL_TE = range(5)
FLUX_DEFAULT, FLUX_LLF, FLUX_UPWIND, FLUX_ROE = range(4)
RK4, RK2MID, SSPRK2, SSPRK3 = 1, 2, 3, 4
(PROB_ADV_SINSIN, PROB_ADV_LINEAR, PROB_EULER_VORTEX, PROB_EULER_SUPERSONIC, PROB_EULER_CONST,
 PROB_SWE_BUMP, PROB_MAXWELL_PULSE, PROB_CONST, PROB_ADV_ROTHILL) = range(9)
HOST, DEVICE = 0, 1
OBJ_LINEAR, OBJ_ENTROPY, OBJ_PATCH_SQ = range(3)

VOLUME_DEFAULT, VOLUME_VALU, VOLUME_MFMA = 0, 1, 2
ENS_SPLIT = 0x10000  # dg_ens_build: a stage launch per order group
NVARS = {ADVECTION: 1, EULER: 4, SWE: 3, MAXWELL_TM: 3, MAXWELL_TE: 3}

# every function the header declares
EXPORTS = [
    "dg_config_default", "dg_create", "dg_destroy", "dg_last_error", "dg_set_mesh",
    "dg_load_pmsh", "dg_set_eps", "dg_project_ic", "dg_set_coeffs", "dg_get_coeffs",
    "dg_eval_rhs", "dg_eval_rhs_noalloc", "dg_step", "dg_run", "dg_l2_error", "dg_l1_error", "dg_eval_exact", "dg_eval_u", "dg_lambda", "dg_limit", "dg_info",
    "dg_geometry", "dg_sync", "dg_stream", "dg_last_timing", "dg_adjoint_set_objective",
    "dg_adjoint_solve", "dg_adjoint_rhs", "dg_adjoint_timing", "dg_adjoint_objective",
    "dg_ens_create", "dg_ens_destroy", "dg_ens_last_error", "dg_ens_add", "dg_ens_build",
    "dg_ens_step", "dg_ens_get_coeffs", "dg_ens_set_coeffs", "dg_ens_state", "dg_ens_info",
    "dg_ens_l2_error", "dg_boundary_points", "dg_set_boundary_states", "dg_get_boundary_states",
]


class DgConfig(ctypes.Structure):
    _fields_ = [
        ("physics", ctypes.c_int), ("p", ctypes.c_int), ("flux", ctypes.c_int),
        ("integrator", ctypes.c_int), ("limiter", ctypes.c_int), ("positivity", ctypes.c_int),
        ("cfl", ctypes.c_double), ("gamma", ctypes.c_double), ("g", ctypes.c_double),
        ("eps", ctypes.c_double), ("mu", ctypes.c_double), ("ax", ctypes.c_double),
        ("ay", ctypes.c_double), ("problem", ctypes.c_int),
        ("problem_params", ctypes.c_double * 8), ("device", ctypes.c_int),
        ("volume", ctypes.c_int),
    ]


class DgInfo(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_int), ("S", ctypes.c_int), ("N", ctypes.c_int), ("n_p", ctypes.c_int),
        ("n_q", ctypes.c_int), ("n_q1", ctypes.c_int), ("dofs", ctypes.c_longlong),
        ("min_r", ctypes.c_double), ("vgprs_stage", ctypes.c_int),
        ("adj_fix_fired", ctypes.c_int), ("adj_resident", ctypes.c_int),
    ]


class DgEnsInfo(ctypes.Structure):
    _fields_ = [
        ("ninst", ctypes.c_int), ("dofs", ctypes.c_longlong), ("elements", ctypes.c_longlong),
        ("block", ctypes.c_int), ("tune_ms", ctypes.c_double * 3), ("vgprs", ctypes.c_int),
        ("launches_per_step", ctypes.c_int), ("ms_per_step", ctypes.c_double),
        ("tune_split_ms", ctypes.c_double * 3), ("split", ctypes.c_int),
        ("block_per_order", ctypes.c_int * 6), ("tune_best_ms", ctypes.c_double),
    ]


_lib = None


def load():
    """Load libdgamd.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libdgamd.so not built (run __graft_entry__.build()): %s" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D, VP = ctypes.POINTER, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    H = VP
    sig = {
        "dg_config_default": (None, [P(DgConfig)]),
        "dg_create": (I, [P(H), P(DgConfig)]),
        "dg_destroy": (None, [H]),
        "dg_last_error": (ctypes.c_char_p, [H]),
        "dg_set_mesh": (I, [H, I, VP, VP, I, VP, VP, VP, VP, VP]),
        "dg_load_pmsh": (I, [H, ctypes.c_char_p]),
        "dg_set_eps": (I, [H, VP]),
        "dg_project_ic": (I, [H]),
        "dg_set_coeffs": (I, [H, VP, I]),
        "dg_get_coeffs": (I, [H, VP, I]),
        "dg_eval_rhs": (I, [H, VP, D, D, VP, I]),
        "dg_eval_rhs_noalloc": (I, [H, VP, D, D, VP]),
        "dg_step": (I, [H, I, D, P(D), P(D)]),
        "dg_run": (I, [H, I, D, D, D, P(D), P(I), P(D)]),
        "dg_l2_error": (I, [H, D, VP]),
        "dg_l1_error": (I, [H, D, VP]),
        "dg_eval_exact": (I, [H, I, D, VP, VP, VP]),
        "dg_eval_u": (I, [H, I, VP, VP, VP]),
        "dg_lambda": (I, [H, VP]),
        "dg_limit": (I, [H]),
        "dg_info": (I, [H, P(DgInfo)]),
        "dg_geometry": (I, [H, VP, VP, VP, VP, VP]),
        "dg_sync": (I, [H]),
        "dg_stream": (VP, [H]),
        "dg_last_timing": (I, [H, P(D), P(I)]),
        "dg_adjoint_set_objective": (I, [H, I, VP, I, VP, I]),
        "dg_adjoint_solve": (I, [H, I, D, I, P(D), VP, VP, I]),
        "dg_adjoint_rhs": (I, [H, VP, D, D, VP, VP, I]),
        "dg_adjoint_timing": (I, [H, P(D), P(D)]),
        "dg_adjoint_objective": (I, [H, P(D)]),
        "dg_ens_create": (I, [P(H), P(DgConfig)]),
        "dg_ens_destroy": (None, [H]),
        "dg_ens_last_error": (ctypes.c_char_p, [H]),
        "dg_ens_add": (I, [H, I, VP, I, VP, VP, I, VP, VP, VP, VP, VP, P(I)]),
        "dg_ens_build": (I, [H, I]),
        "dg_ens_step": (I, [H, I, D]),
        "dg_ens_get_coeffs": (I, [H, I, VP, I]),
        "dg_ens_set_coeffs": (I, [H, I, VP, I]),
        "dg_ens_state": (I, [H, I, P(D), P(D), P(ctypes.c_longlong)]),
        "dg_ens_info": (I, [H, P(DgEnsInfo)]),
        "dg_ens_l2_error": (I, [H, I, D, VP]),
        "dg_boundary_points": (I, [H, P(I), VP, VP, VP, VP]),
        "dg_set_boundary_states": (I, [H, VP, I]),
        "dg_get_boundary_states": (I, [H, D, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class DgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dg error %d: %s" % (code, msg))
        self.code = code


def ptr(a):
    """Address of a contiguous numpy array (host) or torch tensor (device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be contiguous")
        return a.ctypes.data
    return a.data_ptr()
