"""MI355X-native DG right-hand side for 2D conservation laws (gfx950, HIP).

libdgamd.so (C ABI in include/dg.h) holds the kernels and the driver;
this package binds it (Solver), generates and reads meshes (mesh), and
wraps the discrete adjoint for torch (autograd).
"""
from . import _lib  # noqa: F401
from ._lib import (ADVECTION, EULER, SWE, MAXWELL_TM, MAXWELL_TE, FLUX_DEFAULT, FLUX_LLF,  # noqa: F401
                   FLUX_UPWIND, FLUX_ROE, RK4, RK2MID, SSPRK2, SSPRK3, PROB_ADV_SINSIN,
                   PROB_ADV_LINEAR, PROB_EULER_VORTEX, PROB_EULER_SUPERSONIC, PROB_EULER_CONST,
                   PROB_SWE_BUMP, PROB_MAXWELL_PULSE, PROB_CONST, DgError)
from .mesh import Mesh, structured, read_pmsh, write_pmsh  # noqa: F401
from .solver import Solver  # noqa: F401
