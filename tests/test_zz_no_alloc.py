"""dg_eval_rhs_noalloc: the allocation-free variant of dg_eval_rhs (device
pointers, fixed per-GPU scratch).  Two calls in a row on two handles, the
second on a smaller mesh, each equal to dg_eval_rhs on the same input (the
same kernels: bitwise); the device's free memory does not change across the
calls; sizes and overlaps it cannot take are refused before any launch."""
import ctypes
import gc

import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.solver import DgError
from tests.helpers import make_pair, unit_mesh

pytestmark = pytest.mark.gpu


def _mem_free():
    """free device memory as the HIP runtime reports it (hipMemGetInfo)"""
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _case(physics, p, n, periodic):
    import torch
    s, o = make_pair(physics, p, unit_mesh(n, periodic=periodic))
    s.project_ic()
    u0 = s.coeffs()
    u = u0 * (1.0 + 0.05 * np.random.default_rng(n).standard_normal(u0.shape))
    if physics == L.EULER:
        u[:, 1:, ::7] *= 300.0  # every seventh element fails the positivity check
    return s, torch.from_numpy(u).cuda()


def _noalloc(s, u, k, t, dt):
    rc = s.lib.dg_eval_rhs_noalloc(s.h, u.data_ptr(), t, dt, k.data_ptr())
    s._chk(rc)
    s._chk(s.lib.dg_sync(s.h))


def test_noalloc_matches_eval_rhs_twice():
    import torch
    cases = [(L.EULER, 2, 16, True), (L.SWE, 1, 6, False)]  # second: smaller, walled
    handles = [_case(*c) for c in cases]
    ks = [torch.empty_like(u) for _, u in handles]
    for (s, u), k in zip(handles, ks):  # every kernel loaded once before the check
        _noalloc(s, u, k, 0.1, 0.01)
    refs = [s.eval_rhs(u, t=0.1, dt=0.01) for s, u in handles]  # the existing entry point
    for k in ks:
        k.fill_(float("nan"))
    torch.cuda.synchronize()
    # earlier tests' handles are destroyed (their device buffers freed) before
    # the window, and no collection runs inside it
    gc.collect()
    gc.disable()
    try:
        free0 = _mem_free()
        for (s, u), k in zip(handles, ks):  # two calls in a row, the second on a smaller mesh
            _noalloc(s, u, k, 0.1, 0.01)
        torch.cuda.synchronize()
        free1 = _mem_free()
    finally:
        gc.enable()
    for (s, u), k, ref in zip(handles, ks, refs):
        print("noalloc K", s.K, "max |diff|", float((k - ref).abs().max()))
        assert bool(torch.isfinite(ref).all())
        assert torch.equal(k, ref)
    assert free1 == free0, (free0, free1)
    for s, _ in handles:
        s.close()


def test_noalloc_refuses_overlapping_buffers():
    s, u = _case(L.EULER, 2, 6, True)
    with pytest.raises(DgError) as ei:
        s._chk(s.lib.dg_eval_rhs_noalloc(s.h, u.data_ptr(), 0.0, 0.01, u.data_ptr()))
    assert ei.value.code == L.DG_ERR_ARG and "overlap" in str(ei.value)
    s.close()
