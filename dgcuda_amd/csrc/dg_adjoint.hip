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
// takes.  For Maxwell TE, dJ/deps per element has two parts: the impedances
// in the side fluxes (k_eps_side, one thread per side, both elements' terms)
// and the 1/eps row scaling, -lam . k_E / eps, which k_vjp adds in its
// epilogue from the forward stage's stored k.
#include "dg_device.h"
#include "dg_internal.h"
#include "dg_physics.h"
#include "dg_tabvals.h"
#include "dg_stage_q.h"

namespace dgk {

// ------------------------------------------------------------ flux VJPs
// (positive densities and depths: the quotients and roots take the refined
// hardware estimates frcp / fsqrt, as the forward fluxes do)

// g_U += (d(F.n)/dU)^T g for the Euler flux along (nx, ny)
__device__ __forceinline__ void euler_fn_vjp(const PhysParams &ph, const double *U, double nx,
                                             double ny, const double *g, double *gU) {
  const double ir = frcp(U[0]);
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
  const double ir = frcp(U[0]);
  const double u = U[1] * ir, v = U[2] * ir;
  const double gm1 = ph.gamma - 1.;
  const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * v));
  const double vn = u * nx + v * ny;
  const double cs = fsqrt(ph.gamma * p * ir);
  const double sg = vn > 0 ? 1. : (vn < 0 ? -1. : 0.);
  const double q = 0.5 * ph.gamma * frcp(cs) * ir;  // dc = q (dp - p/rho e0)
  gU[0] += c * (sg * (-vn * ir) + q * (gm1 * 0.5 * (u * u + v * v) - p * ir));
  gU[1] += c * (sg * nx * ir - q * gm1 * u);
  gU[2] += c * (sg * ny * ir - q * gm1 * v);
  gU[3] += c * (q * gm1);
}

__device__ __forceinline__ void swe_fn_vjp(const PhysParams &ph, const double *U, double nx,
                                           double ny, const double *g, double *gU) {
  const double ir = frcp(U[0]);
  const double vn = (U[1] * nx + U[2] * ny) * ir;
  const double a = g[1] * U[1] + g[2] * U[2];
  const double b = g[1] * nx + g[2] * ny;
  gU[0] += a * (-vn * ir) + b * ph.G * U[0];
  gU[1] += a * nx * ir + vn * g[1] + g[0] * nx;
  gU[2] += a * ny * ir + vn * g[2] + g[0] * ny;
}

__device__ __forceinline__ void swe_speed_vjp(const PhysParams &ph, const double *U, double nx,
                                              double ny, double c, double *gU) {
  const double ir = frcp(U[0]);
  const double vn = (U[1] * nx + U[2] * ny) * ir;
  const double sg = vn > 0 ? 1. : (vn < 0 ? -1. : 0.);
  const double cs = fsqrt(U[0] * ph.G);
  gU[0] += c * (sg * (-vn * ir) + 0.5 * ph.G * frcp(cs));
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

// d|x|/dx, 0 at 0 (the branch the FP64 tape's |x| takes)
__device__ __forceinline__ double sgn0(double x) { return x > 0. ? 1. : (x < 0. ? -1. : 0.); }

// (g_UL, g_UR) += the transpose of the shallow-water Roe flux's derivative
// applied to g: reverse-mode differentiation of swe_roe (dg_physics.h), line
// by line backwards through its Roe averages, wave strengths and speeds.
// The central part is fn_vjp's; the dissipation F -= d / 2 is what follows.
__device__ inline void swe_roe_vjp(const PhysParams &ph, const double *UL, const double *UR,
                                   double nx, double ny, const double *g, double *gL, double *gR) {
  double h[3] = {0.5 * g[0], 0.5 * g[1], 0.5 * g[2]};
  fn_vjp<PH_SWE>(ph, UL, nx, ny, h, gL);
  fn_vjp<PH_SWE>(ph, UR, nx, ny, h, gR);
  // the forward pass of the dissipation
  const double G = ph.G, hl = UL[0], hr = UR[0];
  const double ul = UL[1] / hl, vl = UL[2] / hl, ur = UR[1] / hr, vr = UR[2] / hr;
  const double sl = sqrt(hl), sr = sqrt(hr);
  const double ht = 0.5 * (hl + hr);
  const double isum = 1.0 / (sl + sr);
  const double nu = sl * ul + sr * ur, nv = sl * vl + sr * vr;
  const double ut = nu * isum, vt = nv * isum;
  const double ct = sqrt(G * ht);
  const double unt = ut * nx + vt * ny, utt = -ut * ny + vt * nx;
  const double dh = hr - hl;
  const double dqn = (UR[1] - UL[1]) * nx + (UR[2] - UL[2]) * ny;
  const double dqt = -(UR[1] - UL[1]) * ny + (UR[2] - UL[2]) * nx;
  const double wn = (dqn - unt * dh) / ct;
  const double a1 = 0.5 * (dh - wn), a3 = 0.5 * (dh + wn), a2 = dqt - utt * dh;
  const double l1 = fabs(unt - ct), l2 = fabs(unt), l3 = fabs(unt + ct);
  const double p1 = l1 * a1, p2 = l2 * a2, p3 = l3 * a3;
  // backwards from d = (d0, d1, d2), whose adjoint is -g / 2
  const double bd0 = -0.5 * g[0], bd1 = -0.5 * g[1], bd2 = -0.5 * g[2];
  const double bp1 = bd0 + bd1 * (ut - ct * nx) + bd2 * (vt - ct * ny);
  const double bp2 = -bd1 * ny + bd2 * nx;
  const double bp3 = bd0 + bd1 * (ut + ct * nx) + bd2 * (vt + ct * ny);
  double but = bd1 * (p1 + p3), bvt = bd2 * (p1 + p3);
  double bct = nx * bd1 * (p3 - p1) + ny * bd2 * (p3 - p1);
  const double ba1 = bp1 * l1, ba2 = bp2 * l2, ba3 = bp3 * l3;
  const double bl1 = bp1 * a1, bl2 = bp2 * a2, bl3 = bp3 * a3;
  const double s1 = sgn0(unt - ct), s2 = sgn0(unt), s3 = sgn0(unt + ct);
  double bunt = bl1 * s1 + bl2 * s2 + bl3 * s3;
  bct += -bl1 * s1 + bl3 * s3;
  const double bwn = 0.5 * (ba3 - ba1);
  double bdh = 0.5 * (ba1 + ba3) - ba2 * utt;
  const double butt = -ba2 * dh, bdqt = ba2;
  const double bdqn = bwn / ct;
  bunt -= bwn * dh / ct;
  bdh -= bwn * unt / ct;
  bct -= bwn * wn / ct;
  // the jumps
  gL[1] += -bdqn * nx + bdqt * ny;
  gR[1] += bdqn * nx - bdqt * ny;
  gL[2] += -bdqn * ny - bdqt * nx;
  gR[2] += bdqn * ny + bdqt * nx;
  double bhl = -bdh, bhr = bdh;
  // the Roe averages
  but += bunt * nx - butt * ny;
  bvt += bunt * ny + butt * nx;
  const double bht = bct * G / (2. * ct);
  bhl += 0.5 * bht;
  bhr += 0.5 * bht;
  const double bisum = but * nu + bvt * nv;
  const double bsum = -bisum * isum * isum;
  const double bsl = bsum + (but * ul + bvt * vl) * isum;
  const double bsr = bsum + (but * ur + bvt * vr) * isum;
  const double bul = but * sl * isum, bvl = bvt * sl * isum;
  const double bur = but * sr * isum, bvr = bvt * sr * isum;
  bhl += bsl / (2. * sl) - (bul * ul + bvl * vl) / hl;
  bhr += bsr / (2. * sr) - (bur * ur + bvr * vr) / hr;
  gL[0] += bhl;
  gL[1] += bul / hl;
  gL[2] += bvl / hl;
  gR[0] += bhr;
  gR[1] += bur / hr;
  gR[2] += bvr / hr;
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
    if constexpr (PHYS == PH_SWE) {
      if (fl == FL_ROE) {
        swe_roe_vjp(ph, UL, UR, nx, ny, g, gL, gR);
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
__device__ __forceinline__ void wall_vjp(const PhysParams &ph, double wx, double wy, double rule,
                                         const double *gR, double *gL) {
  if constexpr (PHYS == PH_ADV) {
    gL[0] += gR[0];
  } else if constexpr (PHYS == PH_EULER || PHYS == PH_SWE) {
    const int E = (PHYS == PH_EULER) ? 3 : -1;
    gL[0] += gR[0];
    if (E > 0) gL[3] += gR[3];
    if (PHYS == PH_EULER && rule != 0.) {
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

// out[j K] = c0 g[j] + sum_{t < nt} oc[t] ot[t][j K] at element e, j < NE.
// The branch on nt is per array, so each array's NE loads are requested
// together (a branch per entry makes each load a round trip of its own: the
// RK4 stage-0 launch, which combines four arrays, took twice as long as the
// others).  The terms are added in the order t = 0, 1, ...
template <int NE>
__device__ __forceinline__ void combine_store(double *out, const double (&g)[NE], double c0,
                                              const double *const (&ot)[4], const double (&oc)[4],
                                              int nt, int K, int e) {
  double v[NE];
#pragma unroll
  for (int j = 0; j < NE; j++) v[j] = c0 * g[j];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (t < nt) {
      const double *src = ot[t];
      const double c = oc[t];
#pragma unroll
      for (int j = 0; j < NE; j++) v[j] = fma(c, gld(src + (size_t)j * K, e), v[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < NE; j++) gst_nt(out + (size_t)j * K, e, v[j]);
}

// the adjoint weight of entry q (row offset) of element e
__device__ __forceinline__ double lam_at(const AdjArgs &a, size_t q, int e) {
  return a.lc * gld(a.lam + q, e);
}

template <int PHYS>
__host__ __device__ constexpr bool linear_flux() {
  return PHYS == PH_ADV || PHYS == PH_TM || PHYS == PH_TE;
}

// MODE 1: (dk/du)^T lam, reading the state.  For the linear systems the
// transpose does not read the state: MODE 0 computes it without.
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
  constexpr int NU = kU ? N : 1, NPU = kU ? NP : 1;
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

  double c[NU][NPU], w[N][NP], g[N][NP];
  double eps_me = 1.;
  if constexpr (TE) eps_me = gld(a.m.eps, ee);
  // advection whose velocity varies in space (the rotating hill): the
  // element's vertices give the points where the velocity is evaluated, as
  // in the forward stage
  double Vv[6] = {0., 0., 0., 0., 0., 0.};
  const bool vary = PHYS == PH_ADV && a.ph.prob == PB_ADV_ROTHILL;
  if constexpr (PHYS == PH_ADV) {
    if (vary) {
#pragma unroll
      for (int q = 0; q < 6; q++) Vv[q] = gld(a.m.V + (size_t)q * K, ee);
    }
  }
  double sc[NMAX];
  const double Jme = gld(a.m.J, ee);
  {
    const double s = a.dt / Jme;
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
      g[n][i] = 0.;
    }

  // volume: transpose of interpolation -> flux -> projection
  {
    const double xr = gld(a.m.xr, ee), yr = gld(a.m.yr, ee);
    const double xs = gld(a.m.xs, ee), ys = gld(a.m.ys, ee);
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      double U[NMAX] = {0., 0., 0., 0.}, gfx[NMAX], gfy[NMAX], gU[NMAX] = {0., 0., 0., 0.};
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
        gfx[n] = wa * ys - wb * yr;
        gfy[n] = wb * xr - wa * xs;
      }
      PhysParams pv = a.ph;
      if constexpr (PHYS == PH_ADV) {
        if (vary) {
          double x, y;
          elem_point(Vv, TC::R1[j], TC::R2[j], x, y);
          rothill_velocity(x, y, pv.ax, pv.ay);
        }
      }
      fn_vjp<PHYS>(pv, U, 1.0, 0.0, gfx, gU);
      fn_vjp<PHYS>(pv, U, 0.0, 1.0, gfy, gU);
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int m = 0; m < NP; m++) g[n][m] = fma(gU[n], TC::B[m * NQ + j], g[n][m]);
    }
  }
  pin(g);
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
    if constexpr (kPark) asm volatile("" ::: "memory");  // re-read the parked values per side
    const int2 sl = a.m.slot[(size_t)k * K + ee];
    const int nb = sl.x, nbls = sl.y & 3, role = (sl.y >> 2) & 1;
    const int sid = sl.y >> 3;
    const double nx = gld(a.m.snx, sid), ny = gld(a.m.sny, sid), len = gld(a.m.slen, sid);
    double Un[NQ1][NMAX], Wn[NQ1][NMAX];
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
      for (int n = 0; n < NMAX; n++) {
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
      // the neighbour's weights, then its state, each consumed into its
      // traces before the next requests (a single set in flight)
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) {
          const double lv = snb[n] * lam_at(a, (size_t)(n * NP + i) * K, nb);
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++)
            Wn[jj][n] = fma(lv, s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj], Wn[jj][n]);
        }
      if constexpr (kU) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int n = 0; n < N; n++)
#pragma unroll
          for (int i = 0; i < NP; i++) {
            const double cv = gld(a.us + (size_t)(n * NP + i) * K, nb);
#pragma unroll
            for (int jj = 0; jj < NQ1; jj++)
              Un[jj][n] = fma(cv, s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj], Un[jj][n]);
          }
      }
    }
    const double *bv =
        nb < 0 ? a.m.bnd_val + (size_t)a.m.side[sid].w * (NQ1 * NMAX) : nullptr;
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      // own traces at this point
      double Um[NMAX] = {0., 0., 0., 0.}, Wm[NMAX] = {0., 0., 0., 0.};
#pragma unroll
      for (int n = 0; n < N; n++) {
        double q = 0., u = 0.;
#pragma unroll
        for (int i = 0; i < NP; i++) {
          const double ph = TC::SIDE[(k * NP + i) * NQ1 + jj];
          q = fma(kPark ? s_w[n * NP + i][threadIdx.x] : w[n][i], ph, q);
          if constexpr (kU) u = fma(kPark ? s_u[n * NP + i][threadIdx.x] : c[n][i], ph, u);
        }
        Wm[n] = q;
        Um[n] = u;
      }
      if (nb < 0) {
        // boundary side (this element is its left): ghost state, no weight
        if constexpr (kU) {
          if (nb == -1) {
            wall_state<PHYS>(a.ph, Um, Un[jj], bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2]);
          } else {
#pragma unroll
            for (int n = 0; n < N; n++) Un[jj][n] = bv[jj * NMAX + n];
          }
        }
      }
      double UL[NMAX], UR[NMAX], gF[NMAX] = {0., 0., 0., 0.}, gL[NMAX] = {0., 0., 0., 0.},
                                 gR[NMAX] = {0., 0., 0., 0.};
      const double f = (role ? 0.5 : -0.5) * len * TC::W1[jj];
#pragma unroll
      for (int n = 0; n < NMAX; n++) {
        UL[n] = role ? Un[jj][n] : Um[n];
        UR[n] = role ? Um[n] : Un[jj][n];
      }
#pragma unroll
      for (int n = 0; n < N; n++) gF[n] = f * (Wm[n] - Wn[jj][n]);
      const double eL = role ? eps_nb : eps_me, eR = role ? eps_me : eps_nb;
      PhysParams pv = a.ph;
      if constexpr (PHYS == PH_ADV) {
        // the velocity at this element's side point jj, the forward's point
        if (vary) {
          double x, y;
          side_point<NQ1>(Vv, TC::X1, k, jj, x, y);
          rothill_velocity(x, y, pv.ax, pv.ay);
        }
      }
      double gzl, gzr;
      riemann_vjp<PHYS>(pv, UL, UR, nx, ny, eL, eR, gF, gL, gR, gzl, gzr);
      double gme[NMAX];
#pragma unroll
      for (int n = 0; n < N; n++) gme[n] = role ? gR[n] : gL[n];
      if (nb == -1) wall_vjp<PHYS>(a.ph, bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2], gR, gme);
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int m = 0; m < NP; m++)
          g[n][m] = fma(gme[n], TC::SIDE[(k * NP + m) * NQ1 + jj], g[n][m]);
    }
    pin(g);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (!act) return;
  {
    double gf[N * NP];
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) gf[n * NP + i] = g[n][i];
    if (a.gout) combine_store<N * NP>(a.gout, gf, a.scale, a.ot, a.oc, a.nout, K, e);
    if (a.gout2) combine_store<N * NP>(a.gout2, gf, a.c2g, a.o2t, a.o2c, a.n2, K, e);
  }
  if constexpr (TE) {
    if (a.geps) {
      // the impedance terms of the element's three sides, then the row
      // scaling: k_E = dt / (J eps) R_E, so lc lam . dk_E/deps = -(lc/eps) lam . k_E
      double gz = 0.;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int info = a.m.slot[(size_t)k * K + e].y;
        gz += gld(a.gzs + (size_t)((info >> 2) & 1) * a.m.S, info >> 3);
      }
      // the parked weights are w = dt / (J eps) lc lam on the E rows, so
      // lc lam . k_E / eps = (J / dt) w . k_E
      double ew = 0.;
#pragma unroll
      for (int n = 0; n < 2; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) {
          const size_t q = (size_t)(n * NP + i) * K;
          const double wv = kPark ? s_w[n * NP + i][threadIdx.x] : sc[n] * lam_at(a, q, e);
          ew = fma(wv, gld(a.kst + q, e), ew);
        }
      a.geps[e] += a.scale * (gz - ew * (Jme / a.dt));
    }
  }
}

// Maxwell TE: the impedance part of dJ/deps, one thread per side.  The
// side's flux depends on z = sqrt(mu / eps) of its two elements; the
// adjoint weight of the flux is the jump of the two elements' weight traces
// (both elements test the same flux with opposite signs), so one evaluation
// per side gives both elements' terms: gzs[s] for the left element and
// gzs[S + s] for the right (0 at a boundary, whose ghost state carries the
// left element's impedance).
template <int P>
__global__ __launch_bounds__(BLK) void k_eps_side(const AdjArgs a) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ1 = TC::NQ1, N = 3;
  __shared__ double s_side[3 * NP * NQ1];
  for (int q = threadIdx.x; q < 3 * NP * NQ1; q += BLK) s_side[q] = TC::SIDE[q];
  __syncthreads();
  const int S = a.m.S, K = a.m.K;
  const int s = remap_block(blockIdx.x, (S + BLK - 1) / BLK) * BLK + threadIdx.x;
  if (s >= S) return;
  const int4 si = a.m.side[s];
  const int L = si.x, R = si.y, lsn = si.z & 3, rsn = (si.z >> 2) & 3;
  const int Rg = R >= 0 ? R : L;
  const double nx = a.m.snx[s], ny = a.m.sny[s], len = a.m.slen[s];
  const double epsL = a.m.eps[L], epsR = R >= 0 ? a.m.eps[Rg] : epsL;
  // weight scales dt / (J eps) (E rows) and dt / (J mu) (H row), times lc
  const double tL = a.lc * a.dt / a.m.J[L], tR = a.lc * a.dt / a.m.J[Rg];
  const double scL[N] = {tL / epsL, tL / epsL, tL / a.ph.mu};
  const double scR[N] = {tR / epsR, tR / epsR, tR / a.ph.mu};
  // traces at the side's points in the left element's order (the right
  // element walks the side the other way); of the weights only the jump
  // D = W_left - W_right enters.  Each array's loads are consumed before the
  // next array's go out, which bounds the registers in flight.
  double UL[NQ1][N], UR[NQ1][N], D[NQ1][N];
#pragma unroll
  for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
    for (int n = 0; n < N; n++) UL[jj][n] = UR[jj][n] = D[jj][n] = 0.;
  auto trace = [&](const double *src, int el, int ls, bool rev, double sgn, const double *sc,
                   double(&T)[NQ1][N]) {
    double v[N][NP];
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) v[n][i] = gld(src + (size_t)(n * NP + i) * K, el);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      double ph[NQ1];
#pragma unroll
      for (int jj = 0; jj < NQ1; jj++) ph[jj] = s_side[(ls * NP + i) * NQ1 + (rev ? NQ1 - 1 - jj : jj)];
#pragma unroll
      for (int n = 0; n < N; n++) {
        const double x = sc ? v[n][i] * (sgn * sc[n]) : v[n][i];
#pragma unroll
        for (int jj = 0; jj < NQ1; jj++) T[jj][n] = fma(x, ph[jj], T[jj][n]);
      }
    }
  };
  trace(a.us, L, lsn, false, 1., nullptr, UL);
  __builtin_amdgcn_sched_barrier(0);
  trace(a.lam, L, lsn, false, 1., scL, D);
  __builtin_amdgcn_sched_barrier(0);
  if (R >= 0) {
    trace(a.us, R, rsn, true, 1., nullptr, UR);
    __builtin_amdgcn_sched_barrier(0);
    trace(a.lam, R, rsn, true, -1., scR, D);
  } else {
    // ghost state of the boundary condition; the ghost has no weight
    const double *bv = a.m.bnd_val + (size_t)si.w * (NQ1 * NMAX);
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      if (R == -1) {
        wall_state<PH_TE>(a.ph, UL[jj], UR[jj], bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2]);
      } else {
#pragma unroll
        for (int n = 0; n < N; n++) UR[jj][n] = bv[jj * NMAX + n];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  double gl = 0., gr = 0.;
#pragma unroll
  for (int jj = 0; jj < NQ1; jj++) {
    const double f = -0.5 * len * TC::W1[jj];
    double gF[NMAX] = {0., 0., 0., 0.}, gLd[NMAX] = {0., 0., 0., 0.}, gRd[NMAX] = {0., 0., 0., 0.};
#pragma unroll
    for (int n = 0; n < N; n++) gF[n] = f * D[jj][n];
    double gzl, gzr;
    riemann_vjp<PH_TE>(a.ph, UL[jj], UR[jj], nx, ny, epsL, epsR, gF, gLd, gRd, gzl, gzr);
    gl += gzl;
    gr += gzr;
  }
  // dz/deps = -z / (2 eps)
  const double zl = sqrt(a.ph.mu / epsL), zr = sqrt(a.ph.mu / epsR);
  double outL, outR = 0.;
  if (R >= 0) {
    outL = gl * (-zl / (2. * epsL));
    outR = gr * (-zr / (2. * epsR));
  } else {
    outL = (gl + gr) * (-zl / (2. * epsL));
  }
  gst_nt(a.gzs, s, outL);
  gst_nt(a.gzs + S, s, outR);
}

// (dk/du)^T lam for large Euler elements (p >= 4), four lanes per element as
// the stage kernel k_stage_q: lane n holds variable n's state, weights and
// result.  On the sides, lane l evaluates the whole flux adjoint at point l
// of each group of four points (inputs and results transposed within the
// quad through LDS); at leftover points every lane evaluates it and keeps
// its component.
template <int PHYS, int P>
__global__ __launch_bounds__(kQuadBlk, (P == 4 ? 2 : 1)) void k_vjp_q(const AdjArgs a) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, NQ1 = TC::NQ1;
  constexpr int EB = kQuadBlk / 4;
  __shared__ double s_side[3 * NP * NQ1];
  __shared__ double s_tr[kQuadBlk][5];  // transpose rows (padded against bank conflicts)
  // with one leftover point per side (NQ1 = 4 k + 1) the three sides' last
  // points form one more group after the side loop: each lane's traces
  // there wait in its row, and the results come back through row side 0
  constexpr bool kLast = TC::NQ1 % 4 == 1;
  __shared__ double s_lo[kLast ? kQuadBlk : 1][3][4];
  for (int q = threadIdx.x; q < 3 * NP * NQ1; q += kQuadBlk) s_side[q] = TC::SIDE[q];
  __syncthreads();
  const int K = a.m.K;
  const int nblk = (K + EB - 1) / EB;
  const int n = threadIdx.x & 3;
  const int e = remap_block(blockIdx.x, nblk) * EB + (threadIdx.x >> 2);
  const bool act = e < K;
  const int ee = act ? e : K - 1;

  double c[NP], w[NP], g[NP];
  const double sc = a.dt / gld(a.m.J, ee);
#pragma unroll
  for (int i = 0; i < NP; i++) {
    const size_t q = (size_t)(n * NP + i) * K;
    c[i] = gld(a.us + q, ee);
    w[i] = sc * lam_at(a, q, ee);
    g[i] = 0.;
  }

  // volume: transpose of interpolation -> flux -> projection
  {
    const double xr = gld(a.m.xr, ee), yr = gld(a.m.yr, ee);
    const double xs = gld(a.m.xs, ee), ys = gld(a.m.ys, ee);
    // groups of four points: lane n evaluates the flux adjoint at point j0 + n
    constexpr int NGV = NQ / 4;
#pragma unroll
    for (int g4 = 0; g4 < NGV; g4++) {
      const int j0 = 4 * g4;
      double xu[4], xa[4], xb[4];
#pragma unroll
      for (int jl = 0; jl < 4; jl++) {
        const int j = j0 + jl;
        double u = c[0] * TC::B[j], wa = w[0] * TC::GX[j], wb = w[0] * TC::GY[j];
#pragma unroll
        for (int i = 1; i < NP; i++) {
          u = fma(c[i], TC::B[i * NQ + j], u);
          wa = fma(w[i], TC::GX[i * NQ + j], wa);
          wb = fma(w[i], TC::GY[i * NQ + j], wb);
        }
        xu[jl] = u;
        xa[jl] = wa * ys - wb * yr;
        xb[jl] = wb * xr - wa * xs;
      }
      double U[4], gfx[4], gfy[4], gU[4] = {0., 0., 0., 0.}, gn[4];
      quad_transpose(s_tr, xu, U);
      quad_transpose(s_tr, xa, gfx);
      quad_transpose(s_tr, xb, gfy);
      fn_vjp<PHYS>(a.ph, U, 1.0, 0.0, gfx, gU);
      fn_vjp<PHYS>(a.ph, U, 0.0, 1.0, gfy, gU);
      quad_transpose(s_tr, gU, gn);
#pragma unroll
      for (int v = 0; v < 4; v++)
#pragma unroll
        for (int m = 0; m < NP; m++) g[m] = fma(gn[v], TC::B[m * NQ + j0 + v], g[m]);
    }
    // leftover points: every lane evaluates the flux adjoint
#pragma unroll
    for (int j = 4 * NGV; j < NQ; j++) {
      double u = c[0] * TC::B[j], wa = w[0] * TC::GX[j], wb = w[0] * TC::GY[j];
#pragma unroll
      for (int i = 1; i < NP; i++) {
        u = fma(c[i], TC::B[i * NQ + j], u);
        wa = fma(w[i], TC::GX[i * NQ + j], wa);
        wb = fma(w[i], TC::GY[i * NQ + j], wb);
      }
      double U[NMAX], gfx[NMAX], gfy[NMAX], gU[NMAX] = {0., 0., 0., 0.};
      quad_gather(u, U);
      quad_gather(wa * ys - wb * yr, gfx);
      quad_gather(wb * xr - wa * xs, gfy);
      fn_vjp<PHYS>(a.ph, U, 1.0, 0.0, gfx, gU);
      fn_vjp<PHYS>(a.ph, U, 0.0, 1.0, gfy, gU);
      const double gn = pick4(gU, n);
#pragma unroll
      for (int m = 0; m < NP; m++) g[m] = fma(gn, TC::B[m * NQ + j], g[m]);
    }
  }
  pin(g);

  // sides: adjoint of each side flux from both elements' weights
#pragma unroll
  for (int k = 0; k < 3; k++) {
    __builtin_amdgcn_sched_barrier(0);
    const int2 sl = a.m.slot[(size_t)k * K + ee];
    const int nb = sl.x, nbls = sl.y & 3, role = (sl.y >> 2) & 1;
    const int sid = sl.y >> 3;
    const double nx = gld(a.m.snx, sid), ny = gld(a.m.sny, sid), len = gld(a.m.slen, sid);
    double Un[NQ1], Wn[NQ1], Um[NQ1], Wm[NQ1];
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      double u = 0., q = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const double ph = TC::SIDE[(k * NP + i) * NQ1 + jj];
        u = fma(c[i], ph, u);
        q = fma(w[i], ph, q);
      }
      Um[jj] = u;
      Wm[jj] = q;
      Un[jj] = 0.;
      Wn[jj] = 0.;
    }
    {
      // the neighbour's state and weights, five modes per round trip, each
      // at a 32-bit lane offset (row + element; one scalar base per array,
      // as per-mode bases would spill the scalar registers), the offset
      // passed through an opaque asm and the traces pinned after each batch,
      // so the compiler neither issues the loads one at a time nor hoists
      // every side's to the kernel's start (a boundary side reads the
      // element's own row as a stand-in and folds in nothing)
      const bool has = nb >= 0;
      int off = n * NP * K + (has ? nb : ee);
      asm volatile("" : "+v"(off));
      const double snb = a.dt / gld(a.m.J, off - n * NP * K);
#pragma unroll
      for (int i0 = 0; i0 < NP; i0 += 5) {
        int o2 = off;
        asm volatile("" : "+v"(o2));
        double lb[5], cb[5];
#pragma unroll
        for (int b = 0; b < 5; b++)
          if (i0 + b < NP) {
            lb[b] = gld(a.lam, o2 + (i0 + b) * K);
            cb[b] = gld(a.us, o2 + (i0 + b) * K);
          }
#pragma unroll
        for (int b = 0; b < 5; b++) {
          const int i = i0 + b;
          if (i < NP) {
            const double lv = has ? snb * (a.lc * lb[b]) : 0., cv = has ? cb[b] : 0.;
#pragma unroll
            for (int jj = 0; jj < NQ1; jj++) {
              const double ph = s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj];
              Wn[jj] = fma(lv, ph, Wn[jj]);
              Un[jj] = fma(cv, ph, Un[jj]);
            }
          }
        }
        pin(Wn);
        pin(Un);
      }
    }
    const double *bv =
        nb < 0 ? a.m.bnd_val + (size_t)a.m.side[sid].w * (NQ1 * NMAX) : nullptr;
    // groups of four points: lane n evaluates point jj0 + n
    constexpr int NG = NQ1 / 4;
#pragma unroll
    for (int g4 = 0; g4 < NG; g4++) {
      const int jj0 = 4 * g4, pt = jj0 + n;
      double Uo[NMAX], Ux[NMAX], Wo[NMAX], Wx[NMAX];
      {
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = Um[jj0 + j];
        quad_transpose(s_tr, x, Uo);
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = Wm[jj0 + j];
        quad_transpose(s_tr, x, Wo);
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = Un[jj0 + j];
        quad_transpose(s_tr, x, Ux);
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = Wn[jj0 + j];
        quad_transpose(s_tr, x, Wx);
      }
      if (nb < 0) {
        // boundary side (this element is its left): ghost state, no weight
        if (nb == -1) {
          wall_state<PHYS>(a.ph, Uo, Ux, bv[pt * NMAX], bv[pt * NMAX + 1], bv[pt * NMAX + 2]);
        } else {
#pragma unroll
          for (int m = 0; m < NMAX; m++) Ux[m] = bv[pt * NMAX + m];
        }
      }
      double UL[NMAX], UR[NMAX], gF[NMAX] = {0., 0., 0., 0.}, gL[NMAX] = {0., 0., 0., 0.},
                                 gR[NMAX] = {0., 0., 0., 0.};
      double w1[4];
#pragma unroll
      for (int j = 0; j < 4; j++) w1[j] = TC::W1[jj0 + j];
      const double f = (role ? 0.5 : -0.5) * len * pick4(w1, n);
#pragma unroll
      for (int m = 0; m < NMAX; m++) {
        UL[m] = role ? Ux[m] : Uo[m];
        UR[m] = role ? Uo[m] : Ux[m];
        gF[m] = f * (Wo[m] - Wx[m]);
      }
      double gzl, gzr;
      riemann_vjp<PHYS>(a.ph, UL, UR, nx, ny, 1., 1., gF, gL, gR, gzl, gzr);
      double gme[4];
#pragma unroll
      for (int m = 0; m < 4; m++) gme[m] = role ? gR[m] : gL[m];
      if (nb == -1) wall_vjp<PHYS>(a.ph, bv[pt * NMAX], bv[pt * NMAX + 1], bv[pt * NMAX + 2], gR, gme);
      // this lane's component at each point of the group
      double gn[4];
      quad_transpose(s_tr, gme, gn);
#pragma unroll
      for (int v = 0; v < 4; v++)
#pragma unroll
        for (int m = 0; m < NP; m++)
          g[m] = fma(gn[v], TC::SIDE[(k * NP + m) * NQ1 + jj0 + v], g[m]);
    }
    if constexpr (kLast) {
      s_lo[threadIdx.x][k][0] = Um[NQ1 - 1];
      s_lo[threadIdx.x][k][1] = Wm[NQ1 - 1];
      s_lo[threadIdx.x][k][2] = Un[NQ1 - 1];
      s_lo[threadIdx.x][k][3] = Wn[NQ1 - 1];
    }
    // leftover points: every lane evaluates the whole flux adjoint
#pragma unroll
    for (int jj = 4 * NG; jj < (kLast ? 4 * NG : NQ1); jj++) {
      double Uo[NMAX], Ux[NMAX], Wo[NMAX], Wx[NMAX];
      quad_gather(Um[jj], Uo);
      quad_gather(Wm[jj], Wo);
      quad_gather(Un[jj], Ux);
      quad_gather(Wn[jj], Wx);
      if (nb < 0) {
        // boundary side (this element is its left): ghost state, no weight
        if (nb == -1) {
          wall_state<PHYS>(a.ph, Uo, Ux, bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2]);
        } else {
#pragma unroll
          for (int m = 0; m < NMAX; m++) Ux[m] = bv[jj * NMAX + m];
        }
      }
      double UL[NMAX], UR[NMAX], gF[NMAX] = {0., 0., 0., 0.}, gL[NMAX] = {0., 0., 0., 0.},
                                 gR[NMAX] = {0., 0., 0., 0.};
      const double f = (role ? 0.5 : -0.5) * len * TC::W1[jj];
#pragma unroll
      for (int m = 0; m < NMAX; m++) {
        UL[m] = role ? Ux[m] : Uo[m];
        UR[m] = role ? Uo[m] : Ux[m];
        gF[m] = f * (Wo[m] - Wx[m]);
      }
      double gzl, gzr;
      riemann_vjp<PHYS>(a.ph, UL, UR, nx, ny, 1., 1., gF, gL, gR, gzl, gzr);
      double gme[NMAX];
#pragma unroll
      for (int m = 0; m < NMAX; m++) gme[m] = role ? gR[m] : gL[m];
      if (nb == -1) wall_vjp<PHYS>(a.ph, bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2], gR, gme);
      const double gn = pick4(gme, n);
#pragma unroll
      for (int m = 0; m < NP; m++) g[m] = fma(gn, TC::SIDE[(k * NP + m) * NQ1 + jj], g[m]);
    }
    pin(g);
  }
  if constexpr (kLast) {
    // the sides' last points: lane ks < 3 evaluates side ks's (lane 3
    // repeats side 2 and its result is not used)
    constexpr int jl = NQ1 - 1;
    const int ks = n < 3 ? n : 2, base = threadIdx.x & ~3;
    wave_sync();
    double Uo[NMAX], Ux[NMAX], Wo[NMAX], Wx[NMAX];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      Uo[v] = s_lo[base + v][ks][0];
      Wo[v] = s_lo[base + v][ks][1];
      Ux[v] = s_lo[base + v][ks][2];
      Wx[v] = s_lo[base + v][ks][3];
    }
    const int2 sl = a.m.slot[(size_t)ks * K + ee];
    const int nb = sl.x, role = (sl.y >> 2) & 1, sid = sl.y >> 3;
    const double nx = gld(a.m.snx, sid), ny = gld(a.m.sny, sid), len = gld(a.m.slen, sid);
    const double *bv =
        nb < 0 ? a.m.bnd_val + (size_t)a.m.side[sid].w * (NQ1 * NMAX) : nullptr;
    if (nb < 0) {
      if (nb == -1) {
        wall_state<PHYS>(a.ph, Uo, Ux, bv[jl * NMAX], bv[jl * NMAX + 1], bv[jl * NMAX + 2]);
      } else {
#pragma unroll
        for (int m = 0; m < NMAX; m++) Ux[m] = bv[jl * NMAX + m];
      }
    }
    double UL[NMAX], UR[NMAX], gF[NMAX] = {0., 0., 0., 0.}, gL[NMAX] = {0., 0., 0., 0.},
                               gR[NMAX] = {0., 0., 0., 0.};
    const double f = (role ? 0.5 : -0.5) * len * TC::W1[jl];
#pragma unroll
    for (int m = 0; m < NMAX; m++) {
      UL[m] = role ? Ux[m] : Uo[m];
      UR[m] = role ? Uo[m] : Ux[m];
      gF[m] = f * (Wo[m] - Wx[m]);
    }
    double gzl, gzr;
    riemann_vjp<PHYS>(a.ph, UL, UR, nx, ny, 1., 1., gF, gL, gR, gzl, gzr);
    double gme[NMAX];
#pragma unroll
    for (int m = 0; m < NMAX; m++) gme[m] = role ? gR[m] : gL[m];
    if (nb == -1) wall_vjp<PHYS>(a.ph, bv[jl * NMAX], bv[jl * NMAX + 1], bv[jl * NMAX + 2], gR, gme);
    wave_sync();
#pragma unroll
    for (int m = 0; m < 4; m++) s_lo[threadIdx.x][0][m] = gme[m];
    wave_sync();
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
      const double gk = s_lo[base + kk][0][n];
#pragma unroll
      for (int m = 0; m < NP; m++) g[m] = fma(gk, TC::SIDE[(kk * NP + m) * NQ1 + jl], g[m]);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  if (!act) return;
  // (the launcher passes no combination terms: k_combine adds them after,
  // see launch_adjoint_rhs)
  if (a.gout) {
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const size_t q = (size_t)(n * NP + i) * K;
      double v = a.scale * g[i];
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (t < a.nout) v = fma(a.oc[t], gld(a.ot[t] + q, e), v);
      gst_nt(a.gout + q, e, v);
    }
  }
  if (a.gout2) {
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const size_t q = (size_t)(n * NP + i) * K;
      double v = a.c2g * g[i];
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (t < a.n2) v = fma(a.o2c[t], gld(a.o2t[t] + q, e), v);
      gst_nt(a.gout2 + q, e, v);
    }
  }
}

// The RK combination added to a stored c0 g: out = fma(oc[nt-1],
// ot[nt-1], ... fma(oc[0], ot[0], out)), in the order the stage adjoint
// kernels use, so the result is the same to the bit.  With `in` set the
// start value is c0 * in (another output that holds g itself) instead of
// out.  Block row y is coefficient row y (K entries), one thread per
// element; an entry's loads are all requested before the first is used.
struct CombArgs {
  double *out;
  const double *in;
  double c0;
  const double *ot[4];
  double oc[4];
  int nt, K;
};
__global__ __launch_bounds__(BLK) void k_combine(const CombArgs c) {
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= c.K) return;
  const size_t q = (size_t)blockIdx.y * c.K;
  double x[4];
#pragma unroll
  for (int t = 0; t < 4; t++) x[t] = t < c.nt ? gld(c.ot[t] + q, e) : 0.;
  double v = c.in ? c.c0 * gld(c.in + q, e) : gld(c.out + q, e);
#pragma unroll
  for (int t = 0; t < 4; t++)
    if (t < c.nt) v = fma(c.oc[t], x[t], v);
  gst_nt(c.out + q, e, v);
}

void launch_adjoint_rhs(int phys, int p, const AdjArgs &a, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    auto go = [&](auto M) {
      constexpr int md = decltype(M)::value, B = vjp_blk<ph, pp, md>();
      hipLaunchKernelGGL((k_vjp<ph, pp, md>), dim3(nblocks(a.m.K, B)), dim3(B), 0, s, a);
    };
    // the linear systems take the transpose that does not read the state;
    // for the TE eps gradient the side terms go first (k_vjp sums them)
    if constexpr (stage_quad<ph, pp>()) {
      // the four-lane kernel stores c0 g only and k_combine adds the RK
      // terms: in its epilogue each term would be a round trip per entry
      // (batching them there spills, the kernel being at its register cap)
      // When both outputs are wanted and the first is g itself (scale 1, no
      // terms), the second starts from c2g times the first: the kernel
      // writes one array fewer.
      const bool from_first = a.gout && a.gout2 && a.nout == 0 && a.scale == 1.0;
      AdjArgs b = a;
      b.nout = 0;
      b.n2 = 0;
      if (from_first) b.gout2 = nullptr;
      hipLaunchKernelGGL((k_vjp_q<ph, pp>), dim3(nblocks(a.m.K, kQuadBlk / 4)), dim3(kQuadBlk), 0,
                         s, b);
      auto comb = [&](double *out, const double *in, double c0, const double *const(&ot)[4],
                      const double(&oc)[4], int nt) {
        if (!out || (nt <= 0 && !in)) return;
        CombArgs c;
        c.out = out;
        c.in = in;
        c.c0 = c0;
        for (int t = 0; t < 4; t++) {
          c.ot[t] = ot[t];
          c.oc[t] = oc[t];
        }
        c.nt = nt;
        c.K = a.m.K;
        hipLaunchKernelGGL(k_combine, dim3(nblocks(a.m.K, BLK), nvars_of(ph) * np_of(pp)),
                           dim3(BLK), 0, s, c);
      };
      comb(a.gout, nullptr, 0., a.ot, a.oc, a.nout);
      comb(a.gout2, from_first ? a.gout : nullptr, a.c2g, a.o2t, a.o2c, a.n2);
    } else if constexpr (ph == PH_TE) {
      if (a.geps)
        hipLaunchKernelGGL((k_eps_side<pp>), dim3(nblocks(a.m.S, BLK)), dim3(BLK), 0, s, a);
      go(std::integral_constant<int, 0>{});
    } else if constexpr (linear_flux<ph>()) {
      go(std::integral_constant<int, 0>{});
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
