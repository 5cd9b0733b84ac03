"""CUDA tensors cross the C ABI as device pointers (DG_DEVICE): the state,
the right-hand side, the adjoint's gradients and the torch.autograd
bindings keep device tensors in device memory.  Every C call that takes a
memory kind is recorded, so the tests see which path each argument took, and
every device-path result is compared with the host path's."""
import numpy as np
import pytest

from dgcuda_amd import _lib as L
from dgcuda_amd.solver import Solver
from tests.helpers import CASES, unit_mesh

pytestmark = pytest.mark.gpu

_WHERE = {"dg_set_coeffs", "dg_get_coeffs", "dg_eval_rhs", "dg_adjoint_solve", "dg_adjoint_rhs",
          "dg_adjoint_set_objective"}


class _Recorder:
    """forwards to the library and records (function, memory kind)"""

    def __init__(self, lib):
        self._lib = lib
        self.calls = []

    def __getattr__(self, name):
        f = getattr(self._lib, name)
        if name not in _WHERE:
            return f

        def call(*args):
            self.calls.append((name, args[-1]))
            return f(*args)
        return call

    def kinds(self, name):
        return [w for n, w in self.calls if n == name]


def _solver(physics=L.EULER, p=2, n=6, integrator=L.SSPRK3):
    prob, prm, fl = CASES[physics]
    s = Solver(physics, p, integrator=integrator, flux=fl, problem=prob, problem_params=prm)
    s.set_mesh(unit_mesh(n, periodic=True)).project_ic()
    rec = _Recorder(s.lib)
    s.lib = rec
    return s, rec


def test_state_round_trip_on_device():
    import torch
    s, rec = _solver()
    host = s.coeffs()
    x = torch.from_numpy(host).cuda() * 1.5
    s.set_coeffs(x)
    y = torch.empty_like(x)
    s.coeffs(out=y)
    assert y.device.type == "cuda"
    assert torch.equal(x, y)
    assert rec.kinds("dg_set_coeffs") == [L.DEVICE]
    assert rec.kinds("dg_get_coeffs") == [L.HOST, L.DEVICE]
    np.testing.assert_array_equal(s.coeffs(), host * 1.5)
    # a CPU tensor takes the host path
    s.set_coeffs(torch.from_numpy(host))
    assert rec.kinds("dg_set_coeffs")[-1] == L.HOST
    np.testing.assert_array_equal(s.coeffs(), host)
    s.close()


def test_eval_rhs_device_equals_host():
    import torch
    s, rec = _solver()
    u = s.coeffs()
    k_host = s.eval_rhs(u, t=0.0, dt=1e-2)
    k_dev = s.eval_rhs(torch.from_numpy(u).cuda(), t=0.0, dt=1e-2)
    assert torch.is_tensor(k_dev) and k_dev.device.type == "cuda"
    assert rec.kinds("dg_eval_rhs") == [L.HOST, L.DEVICE]
    np.testing.assert_array_equal(k_dev.cpu().numpy(), k_host)
    s.close()


def test_device_arguments_checked_before_the_call():
    import torch
    s, rec = _solver()
    n = s.size
    with pytest.raises(TypeError):
        s.set_coeffs(torch.zeros(n, dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        s.set_coeffs(torch.zeros(n - 1, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        s.coeffs(out=torch.zeros(2 * n, dtype=torch.float64, device="cuda")[::2])
    with pytest.raises(ValueError):
        s.eval_rhs(torch.zeros(n, dtype=torch.float64, device="cuda"), out=np.empty(n))
    with pytest.raises(ValueError):
        s.set_eps(np.ones(s.K - 1))
    assert rec.calls == []
    s.close()


def test_autograd_on_device_matches_host():
    """DGObjective and DGRhs with CUDA tensors: the same J and gradients as
    with CPU tensors, gradients on the device, every crossing DG_DEVICE"""
    import torch
    from dgcuda_amd import autograd as A
    s, rec = _solver(integrator=L.SSPRK2)
    s.set_objective(L.OBJ_ENTROPY)
    c = s.coeffs()
    grads = {}
    for dev in ("cpu", "cuda"):
        u0 = torch.from_numpy(c).to(dev).requires_grad_(True)
        J = A.objective(s, u0, nsteps=5, dt=1e-3, every=2)
        J.backward()
        assert u0.grad.device.type == dev
        grads[dev] = (J.item(), u0.grad.cpu().numpy())
        u = torch.from_numpy(c).to(dev).requires_grad_(True)
        w = torch.from_numpy(np.random.default_rng(3).standard_normal(c.shape)).to(dev)
        u1 = u + A.rhs(s, u, 0.0, 1e-3)
        (w * u1).sum().backward()
        assert u.grad.device.type == dev
        grads[dev + "_rhs"] = u.grad.cpu().numpy()
    assert grads["cpu"][0] == grads["cuda"][0]
    np.testing.assert_array_equal(grads["cpu"][1], grads["cuda"][1])
    np.testing.assert_array_equal(grads["cpu_rhs"], grads["cuda_rhs"])
    assert rec.kinds("dg_adjoint_solve") == [L.HOST, L.DEVICE]
    assert rec.kinds("dg_adjoint_rhs") == [L.HOST, L.DEVICE]
    assert rec.kinds("dg_eval_rhs") == [L.HOST, L.DEVICE]
    s.close()
