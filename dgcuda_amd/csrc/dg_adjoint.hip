// dg_adjoint.hip -- discrete adjoint of the DG right-hand side (no counterpart
// in the reference; it differentiates the map that k_stage computes).
//
// For k(u) = dt * M^-1 R(u) (M^-1 = 1/J per element; Maxwell TE also 1/eps,
// 1/mu per row) one launch of k_vjp computes, per element e,
//   g_e = (dk/du_e)^T lam
// from lam on e and its three neighbours.  It has the structure of the
// forward kernel: the volume term is the transposed interpolation /
// flux-Jacobian / projection chain, and each side's numerical flux is
// differentiated in the element that needs it -- the adjoint of the side
// flux collects both elements' weighted test functions, so nothing is
// scattered and no atomics are used.  Nonsmooth points (the LLF max, |u.n|,
// the upwind switch) take the derivative of the branch the forward pass
// takes.  For Maxwell TE the same launch accumulates dJ/deps per element
// (impedances in the side fluxes and the 1/eps row scaling).
#include "dg_device.h"
#include "dg_internal.h"
#include "dg_physics.h"
#include "dg_tabvals.h"

namespace dgk {

// ------------------------------------------------------------ flux VJPs

// g_U += (d(F.n)/dU)^T g for the Euler flux along (nx, ny)
__device__ __forceinline__ void euler_fn_vjp(const PhysParams &ph, const double *U, double nx,
                                             double ny, const double *g, double *gU) {
  const double ir = 1.0 / U[0];
  const double u = U[1] * ir, v = U[2] * ir;
  const double gm1 = ph.gamma - 1.;
  const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * v));
  const double vn = u * nx + v * ny;
  const double a = g[1] * U[1] + g[2] * U[2] + g[3] * (U[3] + p);
  const double b = g[1] * nx + g[2] * ny + g[3] * vn;
  // d vn = (-vn, nx, ny, 0) / rho ; d p = (gamma-1) (|v|^2/2, -u, -v, 1)
  gU[0] += a * (-vn * ir) + b * gm1 * 0.5 * (u * u + v * v);
  gU[1] += a * (nx * ir) - b * gm1 * u + vn * g[1] + g[0] * nx;
  gU[2] += a * (ny * ir) - b * gm1 * v + vn * g[2] + g[0] * ny;
  gU[3] += b * gm1 + vn * g[3];
}

// g_U += c * d(|u.n| + c_sound)/dU (Euler)
__device__ __forceinline__ void euler_speed_vjp(const PhysParams &ph, const double *U, double nx,
                                                double ny, double c, double *gU) {
  const double ir = 1.0 / U[0];
  const double u = U[1] * ir, v = U[2] * ir;
  const double gm1 = ph.gamma - 1.;
  const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * v));
  const double vn = u * nx + v * ny;
  const double cs = sqrt(ph.gamma * p * ir);
  const double sg = vn > 0 ? 1. : (vn < 0 ? -1. : 0.);
  const double q = ph.gamma / (2. * cs) * ir;  // dc = q (dp - p/rho e0)
  gU[0] += c * (sg * (-vn * ir) + q * (gm1 * 0.5 * (u * u + v * v) - p * ir));
  gU[1] += c * (sg * nx * ir - q * gm1 * u);
  gU[2] += c * (sg * ny * ir - q * gm1 * v);
  gU[3] += c * (q * gm1);
}

__device__ __forceinline__ void swe_fn_vjp(const PhysParams &ph, const double *U, double nx,
                                           double ny, const double *g, double *gU) {
  const double ir = 1.0 / U[0];
  const double vn = (U[1] * nx + U[2] * ny) * ir;
  const double a = g[1] * U[1] + g[2] * U[2];
  const double b = g[1] * nx + g[2] * ny;
  gU[0] += a * (-vn * ir) + b * ph.G * U[0];
  gU[1] += a * nx * ir + vn * g[1] + g[0] * nx;
  gU[2] += a * ny * ir + vn * g[2] + g[0] * ny;
}

__device__ __forceinline__ void swe_speed_vjp(const PhysParams &ph, const double *U, double nx,
                                              double ny, double c, double *gU) {
  const double ir = 1.0 / U[0];
  const double vn = (U[1] * nx + U[2] * ny) * ir;
  const double sg = vn > 0 ? 1. : (vn < 0 ? -1. : 0.);
  const double cs = sqrt(U[0] * ph.G);
  gU[0] += c * (sg * (-vn * ir) + ph.G / (2. * cs));
  gU[1] += c * sg * nx * ir;
  gU[2] += c * sg * ny * ir;
}

// g_U += (d(F.n)/dU)^T g for the physical flux of each system
template <int PHYS>
__device__ __forceinline__ void fn_vjp(const PhysParams &ph, const double *U, double nx, double ny,
                                       const double *g, double *gU) {
  if constexpr (PHYS == PH_ADV) {
    gU[0] += (ph.ax * nx + ph.ay * ny) * g[0];
  } else if constexpr (PHYS == PH_EULER) {
    euler_fn_vjp(ph, U, nx, ny, g, gU);
  } else if constexpr (PHYS == PH_SWE) {
    swe_fn_vjp(ph, U, nx, ny, g, gU);
  } else if constexpr (PHYS == PH_TM) {
    // F.n = (-U2 ny / mu, U2 nx / mu, (U1 nx - U0 ny) / eps)
    gU[0] += -g[2] * ny / ph.eps;
    gU[1] += g[2] * nx / ph.eps;
    gU[2] += (-g[0] * ny + g[1] * nx) / ph.mu;
  } else {
    // TE, material-free: F.n = (-U2 ny, U2 nx, U1 nx - U0 ny)
    gU[0] += -g[2] * ny;
    gU[1] += g[2] * nx;
    gU[2] += -g[0] * ny + g[1] * nx;
  }
}

// (g_UL, g_UR) of the numerical flux F(UL, UR).n; for TE also the
// derivatives with respect to the two impedances (gzl, gzr)
template <int PHYS>
__device__ __forceinline__ void riemann_vjp(const PhysParams &ph, const double *UL,
                                            const double *UR, double nx, double ny, double epsL,
                                            double epsR, const double *g, double *gL, double *gR,
                                            double &gzl, double &gzr) {
  constexpr int N = nvars_of(PHYS);
  gzl = gzr = 0.;
  if constexpr (PHYS == PH_TE) {
    const double zl = sqrt(ph.mu / epsL), zr = sqrt(ph.mu / epsR);
    const double iz = 1.0 / (zl + zr);
    const double etl = -UL[0] * ny + UL[1] * nx, etr = -UR[0] * ny + UR[1] * nx;
    const double hs = (zl * UL[2] + zr * UR[2] + (etl - etr)) * iz;
    const double es = (zr * etl + zl * etr + zl * zr * (UL[2] - UR[2])) * iz;
    const double ghs = -ny * g[0] + nx * g[1], ges = g[2];
    const double getl = (ghs + ges * zr) * iz, getr = (-ghs + ges * zl) * iz;
    gL[0] += -ny * getl;
    gL[1] += nx * getl;
    gL[2] += (ghs * zl + ges * zl * zr) * iz;
    gR[0] += -ny * getr;
    gR[1] += nx * getr;
    gR[2] += (ghs * zr - ges * zl * zr) * iz;
    gzl = ghs * (UL[2] - hs) * iz + ges * (etr + zr * (UL[2] - UR[2]) - es) * iz;
    gzr = ghs * (UR[2] - hs) * iz + ges * (etl + zl * (UL[2] - UR[2]) - es) * iz;
    return;
  } else {
    int fl = ph.flux;
    if (fl == FL_DEFAULT) fl = (PHYS == PH_ADV) ? FL_UPWIND : FL_LLF;
    if constexpr (PHYS == PH_EULER || PHYS == PH_TM) fl = FL_LLF;
    if constexpr (PHYS == PH_ADV) {
      if (fl == FL_UPWIND) {
        const double an = nx * ph.ax + ny * ph.ay;
        if (an > 0)
          gL[0] += an * g[0];
        else
          gR[0] += an * g[0];
        return;
      }
    }
    // LLF: 1/2 [(F.n)(UL) + (F.n)(UR) + lam (UL - UR)], lam = max(sL, sR)
    const double sl = llf_speed<PHYS>(ph, UL, nx, ny), sr = llf_speed<PHYS>(ph, UR, nx, ny);
    const double lam = sl > sr ? sl : sr;
    double h[NMAX];
    double jump = 0.;
#pragma unroll
    for (int n = 0; n < N; n++) {
      h[n] = 0.5 * g[n];
      jump += g[n] * (UL[n] - UR[n]);
    }
    fn_vjp<PHYS>(ph, UL, nx, ny, h, gL);
    fn_vjp<PHYS>(ph, UR, nx, ny, h, gR);
#pragma unroll
    for (int n = 0; n < N; n++) {
      gL[n] += lam * h[n];
      gR[n] -= lam * h[n];
    }
    if constexpr (PHYS == PH_EULER) {
      if (sl > sr)
        euler_speed_vjp(ph, UL, nx, ny, 0.5 * jump, gL);
      else
        euler_speed_vjp(ph, UR, nx, ny, 0.5 * jump, gR);
    } else if constexpr (PHYS == PH_SWE) {
      if (sl > sr)
        swe_speed_vjp(ph, UL, nx, ny, 0.5 * jump, gL);
      else
        swe_speed_vjp(ph, UR, nx, ny, 0.5 * jump, gR);
    }
  }
}

// g_UL += (d UR / d UL)^T g_UR for the wall ghost state of wall_state()
template <int PHYS>
__device__ __forceinline__ void wall_vjp(const PhysParams &ph, double wx, double wy,
                                         const double *gR, double *gL) {
  if constexpr (PHYS == PH_ADV) {
    gL[0] += gR[0];
  } else if constexpr (PHYS == PH_EULER || PHYS == PH_SWE) {
    const int E = (PHYS == PH_EULER) ? 3 : -1;
    gL[0] += gR[0];
    if (E > 0) gL[3] += gR[3];
    if (PHYS == PH_EULER && ph.prob == PB_EULER_SUPERSONIC) {
      // (m.t) t with t = (wy, -wx): symmetric projector
      gL[1] += wy * wy * gR[1] - wx * wy * gR[2];
      gL[2] += -wx * wy * gR[1] + wx * wx * gR[2];
    } else {
      // I - 2 w w^T (symmetric)
      gL[1] += (1 - 2 * wx * wx) * gR[1] - 2 * wx * wy * gR[2];
      gL[2] += -2 * wx * wy * gR[1] + (1 - 2 * wy * wy) * gR[2];
    }
  } else {
    // -I + 2 w w^T on the first two components, identity on the third
    gL[0] += (-1 + 2 * wx * wx) * gR[0] + 2 * wx * wy * gR[1];
    gL[1] += 2 * wx * wy * gR[0] + (-1 + 2 * wy * wy) * gR[1];
    gL[2] += gR[2];
  }
}

// ------------------------------------------------------------ VJP kernel

// the adjoint weight of entry q (row offset) of element e
__device__ __forceinline__ double lam_at(const AdjArgs &a, size_t q, int e) {
  return a.lc * gld(a.lam + q, e);
}

template <int PHYS>
__host__ __device__ constexpr bool linear_flux() {
  return PHYS == PH_ADV || PHYS == PH_TM || PHYS == PH_TE;
}

// MODE 1: (dk/du)^T lam (and, for TE with a.geps, dJ/deps) in one pass.  For
// the linear systems the transpose does not read the state: MODE 0 computes
// only (dk/du)^T lam without it, MODE 2 only the TE eps sensitivity.
// The element's own weights (and state, where read) wait in LDS while its
// sides are evaluated, when they fit in 48 KB per block: 256-thread blocks
// up to 24 parked doubles per element, 128-thread blocks up to 48.
template <int PHYS, int P, int MODE>
__host__ __device__ constexpr int vjp_parked() {
  return nvars_of(PHYS) * np_of(P) * ((MODE != 0 || !linear_flux<PHYS>()) ? 2 : 1);
}
template <int PHYS, int P, int MODE>
__host__ __device__ constexpr bool vjp_park() {
  return vjp_parked<PHYS, P, MODE>() <= 48;
}
template <int PHYS, int P, int MODE>
__host__ __device__ constexpr int vjp_blk() {
  return vjp_parked<PHYS, P, MODE>() <= 24 || !vjp_park<PHYS, P, MODE>() ? 256 : 128;
}

template <int PHYS, int P, int MODE>
__global__ __launch_bounds__((vjp_blk<PHYS, P, MODE>())) void k_vjp(const AdjArgs a) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, NQ1 = TC::NQ1, N = nvars_of(PHYS);
  constexpr bool TE = PHYS == PH_TE;
  constexpr bool kU = MODE != 0 || !linear_flux<PHYS>();  // the state is read
  constexpr bool kG = MODE != 2;                           // (dk/du)^T lam is formed
  constexpr bool kE = TE && MODE != 0;                     // dJ/deps is formed
  constexpr int NU = kU ? N : 1, NPU = kU ? NP : 1;
  constexpr int NG = kG ? N : 1, NPG = kG ? NP : 1;
  constexpr int B = vjp_blk<PHYS, P, MODE>();
  constexpr bool kPark = vjp_park<PHYS, P, MODE>();
  __shared__ double s_side[3 * NP * NQ1];
  __shared__ double s_w[kPark ? N * NP : 1][B];
  __shared__ double s_u[kPark && kU ? N * NP : 1][B];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3 * NP * NQ1; q++) s_side[q] = TC::SIDE[q];
  }
  __syncthreads();
  const int K = a.m.K;
  const int nblk = (K + B - 1) / B;
  const int e = remap_block(blockIdx.x, nblk) * B + threadIdx.x;
  const bool act = e < K;
  const int ee = act ? e : K - 1;

  double c[NU][NPU], w[N][NP], g[NG][NPG];
  double eps_me = 1.;
  if constexpr (TE) eps_me = gld(a.m.eps, ee);
  double sc[NMAX];
  {
    const double s = a.dt / gld(a.m.J, ee);
#pragma unroll
    for (int n = 0; n < N; n++) sc[n] = s;
    if constexpr (TE) {
      sc[0] = s * (1. / eps_me);
      sc[1] = sc[0];
      sc[2] = s * (1. / a.ph.mu);
    }
  }
#pragma unroll
  for (int n = 0; n < N; n++)
#pragma unroll
    for (int i = 0; i < NP; i++) {
      if constexpr (kU) c[n][i] = gld(a.us + (size_t)(n * NP + i) * K, ee);
      w[n][i] = sc[n] * lam_at(a, (size_t)(n * NP + i) * K, ee);
      if constexpr (kG) g[n][i] = 0.;
    }
  // TE eps sensitivity: geps collects the impedance terms; ew is the E rows'
  // forward residual dotted with the weights, sum_i w_i R_E,i, formed from the
  // weight interpolations the transpose computes anyway
  double geps = 0., ew = 0.;

  // volume: transpose of interpolation -> flux -> projection
  {
    const double xr = gld(a.m.xr, ee), yr = gld(a.m.yr, ee);
    const double xs = gld(a.m.xs, ee), ys = gld(a.m.ys, ee);
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      double U[NMAX] = {0., 0., 0., 0.}, gfx[NMAX], gfy[NMAX], gU[NMAX] = {0., 0., 0., 0.};
      double wav[NMAX], wbv[NMAX];
#pragma unroll
      for (int n = 0; n < N; n++) {
        if constexpr (kU) {
          double s = c[n][0] * TC::B[j];
#pragma unroll
          for (int i = 1; i < NP; i++) s = fma(c[n][i], TC::B[i * NQ + j], s);
          U[n] = s;
        }
        double wa = w[n][0] * TC::GX[j], wb = w[n][0] * TC::GY[j];
#pragma unroll
        for (int i = 1; i < NP; i++) {
          wa = fma(w[n][i], TC::GX[i * NQ + j], wa);
          wb = fma(w[n][i], TC::GY[i * NQ + j], wb);
        }
        wav[n] = wa;
        wbv[n] = wb;
        gfx[n] = wa * ys - wb * yr;
        gfy[n] = wb * xr - wa * xs;
      }
      if constexpr (kG) {
        fn_vjp<PHYS>(a.ph, U, 1.0, 0.0, gfx, gU);
        fn_vjp<PHYS>(a.ph, U, 0.0, 1.0, gfy, gU);
#pragma unroll
        for (int n = 0; n < N; n++)
#pragma unroll
          for (int m = 0; m < NP; m++) g[n][m] = fma(gU[n], TC::B[m * NQ + j], g[n][m]);
      }
      if constexpr (kE) {
        double fx[NMAX], fy[NMAX];
        flux<PHYS>(a.ph, U, fx, fy);
#pragma unroll
        for (int n = 0; n < 2; n++) {
          const double ga = fx[n] * ys - fy[n] * xs, gb = fy[n] * xr - fx[n] * yr;
          ew = fma(ga, wav[n], fma(gb, wbv[n], ew));
        }
      }
    }
  }
  if constexpr (kG) pin(g);
  if constexpr (kPark) {
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) {
        s_w[n * NP + i][threadIdx.x] = w[n][i];
        if constexpr (kU) s_u[n * NP + i][threadIdx.x] = c[n][i];
      }
    asm volatile("" ::: "memory");  // the copies in registers are dead from here
  }

  // sides: adjoint of each side flux from both elements' weights
#pragma unroll
  for (int k = 0; k < 3; k++) {
    __builtin_amdgcn_sched_barrier(0);
    const int2 sl = a.m.slot[(size_t)k * K + ee];
    const int nb = sl.x, nbls = sl.y & 3, role = (sl.y >> 2) & 1;
    const double nx = gld(a.m.snx + (size_t)k * K, ee), ny = gld(a.m.sny + (size_t)k * K, ee);
    const double len = gld(a.m.slen + (size_t)k * K, ee);
    double Um[NQ1][NMAX], Un[NQ1][NMAX], Wm[NQ1][NMAX], Wn[NQ1][NMAX];
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
      for (int n = 0; n < NMAX; n++) {
        Um[jj][n] = 0.;
        Un[jj][n] = 0.;
        Wn[jj][n] = 0.;
      }
    double eps_nb = eps_me;
    if (nb >= 0) {
      double snb[NMAX];
      {
        const double s = a.dt / gld(a.m.J, nb);
        if constexpr (TE) eps_nb = gld(a.m.eps, nb);
#pragma unroll
        for (int n = 0; n < N; n++) snb[n] = s;
        if constexpr (TE) {
          snb[0] = s * (1. / eps_nb);
          snb[1] = snb[0];
          snb[2] = s * (1. / a.ph.mu);
        }
      }
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) {
          const double lv = snb[n] * lam_at(a, (size_t)(n * NP + i) * K, nb);
          double cv = 0.;
          if constexpr (kU) cv = gld(a.us + (size_t)(n * NP + i) * K, nb);
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++) {
            const double ph = s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj];
            if constexpr (kU) Un[jj][n] = fma(cv, ph, Un[jj][n]);
            Wn[jj][n] = fma(lv, ph, Wn[jj][n]);
          }
        }
    }
#pragma unroll
    for (int n = 0; n < N; n++) {
      double wr[NP], ur[NPU];
#pragma unroll
      for (int i = 0; i < NP; i++) {
        wr[i] = kPark ? s_w[n * NP + i][threadIdx.x] : w[n][i];
        if constexpr (kU) ur[i] = kPark ? s_u[n * NP + i][threadIdx.x] : c[n][i];
      }
#pragma unroll
      for (int jj = 0; jj < NQ1; jj++) {
        double q = wr[0] * TC::SIDE[(k * NP) * NQ1 + jj];
#pragma unroll
        for (int i = 1; i < NP; i++) q = fma(wr[i], TC::SIDE[(k * NP + i) * NQ1 + jj], q);
        Wm[jj][n] = q;
        if constexpr (kU) {
          double s = ur[0] * TC::SIDE[(k * NP) * NQ1 + jj];
#pragma unroll
          for (int i = 1; i < NP; i++) s = fma(ur[i], TC::SIDE[(k * NP + i) * NQ1 + jj], s);
          Um[jj][n] = s;
        }
      }
    }
    const double *bv = nullptr;
    if (nb < 0) {
      bv = a.m.bnd_val + (size_t)(sl.y >> 3) * (NQ1 * NMAX);
#pragma unroll
      for (int jj = 0; jj < NQ1; jj++) {
#pragma unroll
        for (int n = 0; n < N; n++) Wn[jj][n] = 0.;
        if constexpr (kU) {
          if (nb == -1) {
            wall_state<PHYS>(a.ph, Um[jj], Un[jj], bv[jj * NMAX], bv[jj * NMAX + 1]);
          } else {
#pragma unroll
            for (int n = 0; n < N; n++) Un[jj][n] = bv[jj * NMAX + n];
          }
        }
      }
    }
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      double UL[NMAX], UR[NMAX], gF[NMAX] = {0., 0., 0., 0.}, gL[NMAX] = {0., 0., 0., 0.},
                                 gR[NMAX] = {0., 0., 0., 0.};
      const double f = (role ? 0.5 : -0.5) * len * TC::W1[jj];
#pragma unroll
      for (int n = 0; n < NMAX; n++) {
        UL[n] = role ? Un[jj][n] : Um[jj][n];
        UR[n] = role ? Um[jj][n] : Un[jj][n];
      }
#pragma unroll
      for (int n = 0; n < N; n++) gF[n] = f * (Wm[jj][n] - Wn[jj][n]);
      const double eL = role ? eps_nb : eps_me, eR = role ? eps_me : eps_nb;
      double gzl, gzr;
      riemann_vjp<PHYS>(a.ph, UL, UR, nx, ny, eL, eR, gF, gL, gR, gzl, gzr);
      if constexpr (kG) {
        double gme[NMAX];
#pragma unroll
        for (int n = 0; n < N; n++) gme[n] = role ? gR[n] : gL[n];
        if (nb == -1) wall_vjp<PHYS>(a.ph, bv[jj * NMAX], bv[jj * NMAX + 1], gR, gme);
#pragma unroll
        for (int n = 0; n < N; n++)
#pragma unroll
          for (int m = 0; m < NP; m++)
            g[n][m] = fma(gme[n], TC::SIDE[(k * NP + m) * NQ1 + jj], g[n][m]);
      }
      if constexpr (kE) {
        // impedance of this element: z = sqrt(mu / eps), dz/deps = -z / (2 eps)
        const double zme = sqrt(a.ph.mu / eps_me);
        double gz = role ? gzr : gzl;
        if (nb < 0) gz = gzl + gzr;  // the conductor ghost carries this element's impedance
        geps += gz * (-zme / (2. * eps_me));
        double F[NMAX];
        riemann<PHYS>(a.ph, UL, UR, nx, ny, eL, eR, F);
#pragma unroll
        for (int n = 0; n < 2; n++) ew = fma(f * F[n], Wm[jj][n], ew);
      }
    }
    if constexpr (kG) pin(g);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (!act) return;
  if constexpr (kG) {
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const size_t q = (size_t)(n * NP + i) * K;
        double v = a.scale * g[n][i];
#pragma unroll
        for (int t = 0; t < 4; t++)
          if (t < a.nout) v += gld(a.ot[t] + q, e);
        gst(a.gout + q, e, v);
      }
    if (a.gout2) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) {
          const size_t q = (size_t)(n * NP + i) * K;
          double v = a.c2g * g[n][i];
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (t < a.n2) v = fma(a.o2c[t], gld(a.o2t[t] + q, e), v);
          gst(a.gout2 + q, e, v);
        }
    }
  }
  if constexpr (kE) {
    // k_E = dt / (J eps) R_E  ->  d(k_E)/deps = -k_E / eps; with the weights
    // w = dt / (J eps) lam that is lam . dk_E/deps = -(1/eps) sum_i w_i R_E,i
    a.geps[e] += a.scale * (geps - ew / eps_me);
  }
}

void launch_adjoint_rhs(int phys, int p, const AdjArgs &a, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    auto go = [&](auto M) {
      constexpr int md = decltype(M)::value, B = vjp_blk<ph, pp, md>();
      hipLaunchKernelGGL((k_vjp<ph, pp, md>), dim3(nblocks(a.m.K, B)), dim3(B), 0, s, a);
    };
    // TE: the lean transpose (state not read), then, for the eps gradient,
    // the eps sensitivity on its own (one pass doing both needs more
    // registers than a wave has)
    if constexpr (ph == PH_TE) {
      go(std::integral_constant<int, 0>{});
      if (a.geps) go(std::integral_constant<int, 2>{});
    } else {
      go(std::integral_constant<int, 1>{});
    }
  });
}

// ------------------------------------------------------------ objectives

// per-element objective value and gradient with respect to the element's
// coefficients.  kind 1 entropy: integral of -rho log(p / rho^gamma) (Euler);
// kind 2: integral of u_var^2 over elements whose centroid lies in
// [x0, x1] x [y0, y1] (prm = x0, x1, y0, y1, var)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_objective(int kind, const MeshDev m, const PhysParams ph,
                                                   const double *c, const double *prm,
                                                   double *val, double *grad) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, N = nvars_of(PHYS);
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  const double J = m.J[e];
  double gr[N][NP];
  for (int n = 0; n < N; n++)
    for (int i = 0; i < NP; i++) gr[n][i] = 0.;
  double v = 0.;
  if (kind == 2) {
    const double cx = (m.V[e] + m.V[2 * (size_t)K + e] + m.V[4 * (size_t)K + e]) / 3.;
    const double cy = (m.V[(size_t)K + e] + m.V[3 * (size_t)K + e] + m.V[5 * (size_t)K + e]) / 3.;
    const int var = (int)prm[4];
    if (cx >= prm[0] && cx <= prm[1] && cy >= prm[2] && cy <= prm[3] && var >= 0 && var < N) {
      for (int j = 0; j < NQ; j++) {
        double U = 0.;
        for (int i = 0; i < NP; i++) U += c[(size_t)(var * NP + i) * K + e] * TC::B[i * NQ + j];
        v += J * TC::W[j] * U * U;
        for (int i = 0; i < NP; i++) gr[var][i] += 2. * J * TC::W[j] * U * TC::B[i * NQ + j];
      }
    }
  } else if (kind == 1 && PHYS == PH_EULER) {
    const double gm = ph.gamma, gm1 = gm - 1.;
    for (int j = 0; j < NQ; j++) {
      double U[NMAX];
      for (int n = 0; n < N; n++) {
        double s = 0.;
        for (int i = 0; i < NP; i++) s += c[(size_t)(n * NP + i) * K + e] * TC::B[i * NQ + j];
        U[n] = s;
      }
      const double rho = U[0], ir = 1. / rho, u = U[1] * ir, vv = U[2] * ir;
      const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * vv));
      const double S = -rho * (log(p) - gm * log(rho));
      v += J * TC::W[j] * S;
      // dS/dU = -(log p - gamma log rho) - rho (dp/p - gamma/rho e0)
      const double lg = log(p) - gm * log(rho);
      const double rp = rho / p;
      double dS[NMAX];
      dS[0] = -lg - rp * gm1 * 0.5 * (u * u + vv * vv) + gm;
      dS[1] = rp * gm1 * u;
      dS[2] = rp * gm1 * vv;
      dS[3] = -rp * gm1;
      for (int n = 0; n < N; n++)
        for (int i = 0; i < NP; i++) gr[n][i] += J * TC::W[j] * dS[n] * TC::B[i * NQ + j];
    }
  }
  val[e] = v;
  if (grad)
    for (int n = 0; n < N; n++)
      for (int i = 0; i < NP; i++) grad[(size_t)(n * NP + i) * K + e] = gr[n][i];
}

void launch_objective(int phys, int p, int kind, const MeshDev &m, const PhysParams &ph,
                      const double *c, const double *prm, double *obj_elem, double *grad,
                      hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_objective<ph_, pp>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, kind, m,
                       ph, c, prm, obj_elem, grad);
  });
}

// ------------------------------------------------------------ elementwise

}  // namespace dgk
