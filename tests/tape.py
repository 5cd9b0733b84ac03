"""An FP64 tape of the oracle's stage map, for checking the discrete adjoint.

TEST INFRASTRUCTURE ONLY.  The oracle's right-hand side (oracle/dg_oracle.c:
eval_surface, eval_volume and eval_rhs_assemble, which restate the
reference's conserv_kernels.cu:850-985, :992-1083 and time_integrator.cu:
26-101) is written again here in torch float64 on the CPU, so that torch's
reverse mode records every operation of a few Runge-Kutta steps and returns
the exact gradient of an objective of the final state with respect to the
initial coefficients and the per-element permittivity.  The geometry (J,
xr, yr, xs, ys, side lengths and normals) and the reference-element tables
(basis, weighted gradients, side traces, quadrature weights) are the
oracle's own, read through its C interface, so the tape evaluates the same
discrete map as the oracle.

Covered: Maxwell TE (exact-Riemann upwind flux with an impedance jump,
conductor walls), Euler (local Lax-Friedrichs flux, reflecting walls or a
periodic mesh), the integrators' Butcher tableaus, and the patch and
entropy objectives.  Nonsmooth points take the branch the oracle's C code
takes: |x| has derivative 0 at 0, and the LLF speed is the right state's
when the two are equal.
"""
import numpy as np
import torch

from dgcuda_amd import _lib as L

# Butcher tableaus (a, b) of the integrators, by the library's integrator ids
TABLEAUS = {
    L.RK4: ([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]],
            [1 / 6, 1 / 3, 1 / 3, 1 / 6]),
    L.RK2MID: ([[0, 0], [0.5, 0]], [0.0, 1.0]),
    L.SSPRK2: ([[0, 0], [1.0, 0]], [0.5, 0.5]),
    L.SSPRK3: ([[0, 0, 0], [1.0, 0, 0], [0.25, 0.25, 0]], [1 / 6, 1 / 6, 2 / 3]),
}


def _t(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype)


class Tape:
    """The stage map k = dt M^-1 R(u) of one oracle configuration as torch
    operations; coefficients are [N, n_p, K] (the oracle's SoA order)."""

    def __init__(self, orc, mesh, physics, gamma=1.4, mu=1.0):
        if physics not in (L.MAXWELL_TE, L.EULER):
            raise ValueError("the tape covers Maxwell TE and Euler")
        if np.any(mesh.right < -1):
            raise ValueError("the tape covers walls and interior sides only")
        self.phys, self.gamma, self.mu = physics, gamma, mu
        geo, tab = orc.geometry(), orc.tables()
        self.B, self.GX, self.GY = _t(tab["basis"]), _t(tab["gx"]), _t(tab["gy"])
        self.W, self.W1 = _t(tab["w"]), _t(tab["w1"])
        side = _t(tab["side"])  # [3, n_p, nq1]
        self.J = _t(geo["J"])
        self.xr, self.yr, self.xs, self.ys = (_t(geo[k]) for k in ("xr", "yr", "xs", "ys"))
        self.nx, self.ny, self.len = _t(geo["nx"]), _t(geo["ny"]), _t(geo["len"])
        self.left = _t(mesh.left, torch.long)
        right = np.asarray(mesh.right)
        self.inner = _t(np.nonzero(right >= 0)[0], torch.long)
        self.wall = _t(np.nonzero(right == -1)[0], torch.long)
        self.right_in = _t(right[right >= 0], torch.long)
        # the left element's traces at the side's points; the right element
        # walks them backwards (point j is its point nq1 - 1 - j)
        self.phiL = side[_t(mesh.lsn, torch.long)]                  # [S, n_p, nq1]
        self.phiR = side[_t(mesh.rsn, torch.long)].flip(-1)         # [S, n_p, nq1]
        cx, cy = mesh.centroids()
        self.cx, self.cy = _t(cx), _t(cy)

    # ------------------------------------------------------------ physics

    def _pressure(self, U):
        return (self.gamma - 1.) * (U[3] - (U[1] * U[1] + U[2] * U[2]) / 2. / U[0])

    def _flux(self, U):
        """physical flux (fx, fy) of states U [N, ...] (oracle eval_flux)"""
        if self.phys == L.MAXWELL_TE:
            z = torch.zeros_like(U[0])
            return torch.stack([z, U[2], U[1]]), torch.stack([-U[2], z, -U[0]])
        p = self._pressure(U)
        r, ru, rv, E = U[0], U[1], U[2], U[3]
        fx = torch.stack([ru, ru * ru / r + p, ru * rv / r, ru * (E + p) / r])
        fy = torch.stack([rv, ru * rv / r, rv * rv / r + p, rv * (E + p) / r])
        return fx, fy

    def _wall(self, UL, nx, ny):
        """reflected ghost state (oracle reflect_state)"""
        if self.phys == L.MAXWELL_TE:
            dot = UL[0] * nx + UL[1] * ny
            return torch.stack([-UL[0] + 2 * dot * nx, -UL[1] + 2 * dot * ny, UL[2]])
        dot = UL[1] * nx + UL[2] * ny
        return torch.stack([UL[0], UL[1] - 2 * dot * nx, UL[2] - 2 * dot * ny, UL[3]])

    def _riemann(self, UL, UR, nx, ny, epsL, epsR):
        """numerical flux (oracle riemann: te_upwind for TE, LLF for Euler)"""
        if self.phys == L.MAXWELL_TE:
            zl, zr = torch.sqrt(self.mu / epsL), torch.sqrt(self.mu / epsR)
            etl, etr = -UL[0] * ny + UL[1] * nx, -UR[0] * ny + UR[1] * nx
            hs = (zl * UL[2] + zr * UR[2] + (etl - etr)) / (zl + zr)
            es = (zr * etl + zl * etr + zl * zr * (UL[2] - UR[2])) / (zl + zr)
            return torch.stack([-ny * hs, nx * hs, es])
        fxl, fyl = self._flux(UL)
        fxr, fyr = self._flux(UR)
        g = self.gamma
        cl = torch.sqrt(g * self._pressure(UL) / UL[0])
        cr = torch.sqrt(g * self._pressure(UR) / UR[0])
        sl = nx * (UL[1] / UL[0]) + ny * (UL[2] / UL[0])
        sr = nx * (UR[1] / UR[0]) + ny * (UR[2] / UR[0])
        lm, rm = (sl.abs() + cl).abs(), (sr.abs() + cr).abs()
        lam = torch.where(lm > rm, lm, rm)
        return 0.5 * ((fxl + fxr) * nx + (fyl + fyr) * ny + lam * (UL - UR))

    # ------------------------------------------------------------ stage map

    def rhs(self, C, dt, eps=None):
        """k = dt / J (volume + side terms), TE rows scaled by 1/eps and 1/mu"""
        # volume (oracle eval_volume)
        U = torch.einsum("nke,kj->nje", C, self.B)
        fx, fy = self._flux(U)
        vol = (torch.einsum("nje,ij->nie", fx, self.GX) * self.ys
               - torch.einsum("nje,ij->nie", fx, self.GY) * self.yr
               - torch.einsum("nje,ij->nie", fy, self.GX) * self.xs
               + torch.einsum("nje,ij->nie", fy, self.GY) * self.xr)
        # sides (oracle eval_surface)
        UL = torch.einsum("nis,sij->nsj", C[:, :, self.left], self.phiL)
        nx, ny = self.nx[:, None], self.ny[:, None]
        S, nq1 = self.phiL.shape[0], self.phiL.shape[2]
        UR = torch.zeros_like(UL)
        UR = UR.index_copy(1, self.inner, torch.einsum(
            "nis,sij->nsj", C[:, :, self.right_in], self.phiR[self.inner]))
        UR = UR.index_copy(1, self.wall, self._wall(UL[:, self.wall], nx[self.wall], ny[self.wall]))
        if self.phys == L.MAXWELL_TE:
            epsL = eps[self.left][:, None].expand(S, nq1)
            epsR = epsL.clone()
            epsR = epsR.index_copy(0, self.inner, eps[self.right_in][:, None].expand(-1, nq1))
        else:
            epsL = epsR = None
        F = self._riemann(UL, UR, nx, ny, epsL, epsR)                   # [N, S, nq1]
        FW = F * self.W1 * (self.len[:, None] / 2)
        rl = -torch.einsum("nsj,sij->nis", FW, self.phiL)
        rr = torch.einsum("nsj,sij->nis", FW, self.phiR)
        # assembly (oracle eval_rhs_assemble)
        tot = vol.index_add(2, self.left, rl)
        tot = tot.index_add(2, self.right_in, rr[:, :, self.inner])
        k = tot * (dt / self.J)
        if self.phys == L.MAXWELL_TE:
            k = torch.cat([k[:2] / eps, k[2:] / self.mu])
        return k

    def run(self, C, nsteps, dt, integrator, eps=None):
        """nsteps explicit RK steps with the integrator's Butcher tableau"""
        a, b = TABLEAUS[integrator]
        for _ in range(nsteps):
            ks = []
            for r in range(len(b)):
                ur = C
                for c in range(r):
                    if a[r][c] != 0:
                        ur = ur + a[r][c] * ks[c]
                ks.append(self.rhs(ur, dt, eps))
            for r in range(len(b)):
                if b[r] != 0:
                    C = C + b[r] * ks[r]
        return C

    # ------------------------------------------------------------ objectives

    def objective(self, C, kind, params=()):
        """the library's objectives: entropy (integral of -rho log(p /
        rho^gamma)) or the patch integral of u_var^2 over elements whose
        centroid lies in [x0, x1] x [y0, y1]"""
        U = torch.einsum("nke,kj->nje", C, self.B)
        if kind == L.OBJ_ENTROPY:
            p = self._pressure(U)
            s = -U[0] * (torch.log(p) - self.gamma * torch.log(U[0]))
            return torch.sum(self.J * torch.einsum("je,j->e", s, self.W))
        if kind == L.OBJ_PATCH_SQ:
            x0, x1, y0, y1, var = params
            inside = ((self.cx >= x0) & (self.cx <= x1) & (self.cy >= y0) & (self.cy <= y1))
            q = torch.einsum("je,j->e", U[int(var)] ** 2, self.W)
            return torch.sum(torch.where(inside, self.J * q, torch.zeros_like(q)))
        raise ValueError("objective kind %d is not on the tape" % kind)

    def gradient(self, C0, nsteps, dt, integrator, kind, params=(), eps=None):
        """(J, dJ/dC0, dJ/deps or None) recorded on the tape"""
        C0 = _t(C0).requires_grad_(True)
        e = _t(eps).requires_grad_(True) if eps is not None else None
        Jv = self.objective(self.run(C0, nsteps, dt, integrator, e), kind, params)
        grads = torch.autograd.grad(Jv, [C0] + ([e] if e is not None else []))
        return (float(Jv.detach()), grads[0].numpy(), grads[1].numpy() if e is not None else None)
