// dg_adjoint.hip -- discrete adjoint kernels (filled in below).
#include "dg_internal.h"
