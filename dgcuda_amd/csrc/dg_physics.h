// dg_physics.h -- device physics: fluxes, numerical fluxes, wall / inflow /
// outflow states and problem states for the five systems.
//
// Reference plugins (paths relative to src/2d/conservation/):
//   Euler      euler/euler.cu:61-160 (pressure, flux), 170-237 (LLF), 248-277 (CFL speed)
//   advection  advection/advection.cu:51-71 (flux), 94-116 (upwind), 157-183 (CFL speed)
//   shallow w. shallowwater/shallowwater.cu:44-109 (flux), 118-185 (LLF), 196-224
//   Maxwell TM maxwells/maxwells.cu:67-84 (flux), 93-163 (LLF), 174-185
//   walls      euler/supersonic/supersonic.cu:65-87, shallowwater/test/test.cu:63-77,
//              maxwells/cmirror/cmirror.cu:71-84
// Arithmetic is FP64 throughout.  Divisions by the density are done once as a
// reciprocal; the results agree with the reference expressions to rounding.
#pragma once

#include "dg_internal.h"

namespace dgk {

enum { PH_ADV = 0, PH_EULER = 1, PH_SWE = 2, PH_TM = 3, PH_TE = 4 };
enum { FL_DEFAULT = 0, FL_LLF = 1, FL_UPWIND = 2, FL_ROE = 3 };
enum {
  PB_ADV_SINSIN = 0, PB_ADV_LINEAR = 1, PB_EULER_VORTEX = 2, PB_EULER_SUPERSONIC = 3,
  PB_EULER_CONST = 4, PB_SWE_BUMP = 5, PB_MAXWELL_PULSE = 6, PB_CONST = 7, PB_ADV_ROTHILL = 8
};

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- problems

// rotating hill (advection/rotatinghill/rotatinghill.cu): a Gaussian hill
// carried by the solid-body rotation a = 2 pi (-y, x) (:46-49), period 1
__device__ __forceinline__ double rothill_u0(double x, double y) {  // :25-39
  const double x0 = 0.2, y0 = 0.0, r = 0.15;
  const double dx = x - x0, dy = y - y0;
  return 5 * exp(-(dx * dx + dy * dy) / (2 * r * r));
}
// the hill rotated to time t, the inflow state (:51-61)
__device__ __forceinline__ double rothill_inflow(double x, double y, double t) {
  const double c = cospi(2 * t), s = sinpi(2 * t);
  return rothill_u0(x * c + y * s, -x * s + y * c);
}
// advection velocity at (x, y) for a problem whose velocity varies in space
__device__ __forceinline__ void rothill_velocity(double x, double y, double &ax, double &ay) {
  ax = -2 * kPi * y;
  ay = 2 * kPi * x;
}

// physical point of 2D quadrature point (r, s) of an element with vertices V
// (get_coordinates_2d conserv_kernels.cu:231-235) and of side point j of
// local side k (get_coordinates_1d :238-261), with the 1D Gauss points X1
__device__ __forceinline__ void elem_point(const double *V, double r, double s, double &x,
                                           double &y) {
  x = V[2] * r + V[4] * s + V[0] * (1 - r - s);
  y = V[3] * r + V[5] * s + V[1] * (1 - r - s);
}
template <int NQ1>
__device__ __forceinline__ void side_point(const double *V, const double *X1, int k, int j,
                                           double &x, double &y) {
  double r, s;
  if (k == 0) {
    r = 0.5 + 0.5 * X1[j];
    s = 0.;
  } else if (k == 1) {
    r = (1. - X1[j]) / 2.;
    s = (1. + X1[j]) / 2.;
  } else {
    r = 0.;
    s = 0.5 + 0.5 * X1[NQ1 - 1 - j];
  }
  elem_point(V, r, s, x, y);
}

__device__ inline void vortex_state(const PhysParams &ph, double x, double y, double t,
                                    double *U) {
  const double *pp = ph.pp;
  double g = ph.gamma, L = pp[5];
  double dx = x - (pp[1] + pp[3] * t), dy = y - (pp[2] + pp[4] * t);
  if (L > 0) {
    dx -= L * floor(dx / L + 0.5);
    dy -= L * floor(dy / L + 0.5);
  }
  double r2 = dx * dx + dy * dy;
  double f = pp[0] / (2. * kPi) * exp(0.5 * (1. - r2));
  double u = pp[3] - f * dy, v = pp[4] + f * dx;
  double T = 1. - (g - 1.) * pp[0] * pp[0] / (8. * g * kPi * kPi) * exp(1. - r2);
  double rho = pow(T, 1. / (g - 1.));
  double p = pow(rho, g);
  U[0] = rho;
  U[1] = rho * u;
  U[2] = rho * v;
  U[3] = p / (g - 1.) + 0.5 * rho * (u * u + v * v);
}

// euler/supersonic/supersonic.cu:27-37
__device__ inline void supersonic_state(double x, double y, double *U) {
  const double GAMMA = 1.4, MACH = 2.25;
  double r = sqrt(x * x + y * y);
  U[0] = pow(1 + 1.0125 * (1. - 1. / (r * r)), 2.5);
  U[1] = U[0] * sin(atan(y / x)) * MACH / r;
  U[2] = U[0] * -cos(atan(y / x)) * MACH / r;
  double p = (1.0 / GAMMA) * pow(U[0], GAMMA);
  U[3] = 0.5 * U[0] * (MACH * MACH / (r * r)) + p * (1. / (GAMMA - 1.));
}

// exact / initial / inflow state of the problem (all NMAX entries written)
template <int N>
__device__ inline void exact_state(const PhysParams &ph, double x, double y, double t,
                                   double *U) {
  const double *pp = ph.pp;
  for (int n = 0; n < NMAX; n++) U[n] = 0.;
  switch (ph.prob) {
    case PB_ADV_SINSIN: {
      double X = x - ph.ax * t, Y = y - ph.ay * t;
      U[0] = sin(2. * kPi * X) * sin(2. * kPi * Y);
      break;
    }
    case PB_ADV_LINEAR:  // advection/exact/exact.cu:25-27
      U[0] = x - y;
      break;
    case PB_EULER_VORTEX:
      vortex_state(ph, x, y, t, U);
      break;
    case PB_EULER_SUPERSONIC:
      supersonic_state(x, y, U);
      break;
    case PB_EULER_CONST:
      U[0] = pp[0];
      U[1] = pp[0] * pp[1];
      U[2] = pp[0] * pp[2];
      U[3] = pp[3] / (ph.gamma - 1.) + 0.5 * pp[0] * (pp[1] * pp[1] + pp[2] * pp[2]);
      break;
    case PB_SWE_BUMP: {
      double d2 = (x - pp[2]) * (x - pp[2]) + (y - pp[3]) * (y - pp[3]);
      U[0] = pp[0] + pp[1] * exp(-d2 / (pp[5] * pp[4] * pp[4]));
      break;
    }
    case PB_MAXWELL_PULSE: {
      double d2 = (x - pp[1]) * (x - pp[1]) + (y - pp[2]) * (y - pp[2]);
      U[N - 1] = pp[0] * exp(-d2 / (pp[3] * pp[3]));
      break;
    }
    case PB_CONST:
      for (int n = 0; n < N; n++) U[n] = pp[n];
      break;
    case PB_ADV_ROTHILL:  // U_exact is U0 (rotatinghill.cu:91-98): one period later
      U[0] = rothill_u0(x, y);
      break;
  }
}

// wall ghost state (U_reflection of the problem plugins) from the interior
// state UL, a unit wall vector (wx, wy) and, for Euler, a rule: 0 mirrors
// the momentum about the wall (its normal part flips); 1 keeps only its
// tangential part (euler/supersonic's "algorithm 2", supersonic.cu:65-87,
// with the radial direction aligned with the normal as the wall vector).
// The vector and rule are boundary data (the side's outward normal and 0
// unless the problem or the caller says otherwise).
// (shallowwater/test/test.cu:63-77, euler/lshock/lshock.cu:57-71,
// maxwells/cmirror/cmirror.cu:71-84; TE conductor: tangential E flips)
template <int PHYS>
__device__ __forceinline__ void wall_state(const PhysParams &ph, const double *UL, double *UR,
                                           double wx, double wy, double rule) {
  if constexpr (PHYS == PH_ADV) {
    UR[0] = UL[0];
  } else if constexpr (PHYS == PH_EULER) {
    UR[0] = UL[0];
    UR[3] = UL[3];
    if (rule != 0.) {
      double q = UL[1] * wy - UL[2] * wx;
      UR[1] = q * wy;
      UR[2] = -q * wx;
    } else {
      double d = UL[1] * wx + UL[2] * wy;
      UR[1] = UL[1] - 2 * d * wx;
      UR[2] = UL[2] - 2 * d * wy;
    }
  } else if constexpr (PHYS == PH_SWE) {
    UR[0] = UL[0];
    double d = UL[1] * wx + UL[2] * wy;
    UR[1] = UL[1] - 2 * d * wx;
    UR[2] = UL[2] - 2 * d * wy;
  } else {
    UR[2] = UL[2];
    double d = UL[0] * wx + UL[1] * wy;
    UR[0] = -UL[0] + 2 * d * wx;
    UR[1] = -UL[1] + 2 * d * wy;
  }
}

// boundary data of one point (eval_boundary conserv_kernels.cu:820-843 with
// the point of get_coordinates_1d :238-261): code 1 wall -> wall vector,
// 2 outflow / 3 inflow -> the problem's state (the outflow of advection/exact
// and advection/rotatinghill leaves the ghost state at zero, exact.cu:50-52,
// rotatinghill.cu:69-71; the rotating hill's inflow is its rotated hill)
template <int PHYS>
__device__ inline void boundary_data(const PhysParams &ph, double x, double y, double t,
                                     double nx, double ny, int code, double *out) {
  constexpr int N = nvars_of(PHYS);
  double W[NMAX];
  for (int n = 0; n < NMAX; n++) out[n] = 0.;
  if (code == 1) {
    double wx = nx, wy = ny;
    if (PHYS == PH_EULER && ph.prob == PB_EULER_SUPERSONIC) {
      double d = sqrt(x * x + y * y);
      wx = x / d;
      wy = y / d;
      if (wx * nx + wy * ny < 0) {
        wx = -wx;
        wy = -wy;
      }
    }
    out[0] = wx;
    out[1] = wy;
    out[2] = (PHYS == PH_EULER && ph.prob == PB_EULER_SUPERSONIC) ? 1. : 0.;
  } else if (code == 2 && (ph.prob == PB_ADV_LINEAR || ph.prob == PB_ADV_ROTHILL)) {
    // zero state
  } else if (code == 3 && ph.prob == PB_ADV_ROTHILL) {
    out[0] = rothill_inflow(x, y, t);
  } else {
    exact_state<N>(ph, x, y, t, W);
    for (int n = 0; n < N; n++) out[n] = W[n];
  }
}

// ---------------------------------------------------------------- physics

// 1/x for a positive normal x: the hardware estimate refined by two Newton
// steps (as the IEEE division sequence refines it), without the scaling and
// fix-up steps that only matter for denormal or infinite operands.  The
// result is within about an ulp of the correctly rounded quotient.
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) for a positive normal x: the reciprocal square root estimate, one
// Goldschmidt step and two residual corrections (the refinement of the
// library square root, without its range scaling)
__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

template <int PHYS>
__device__ inline bool is_physical(const PhysParams &ph, const double *U) {
  if constexpr (PHYS == PH_EULER) {
    if (U[0] < 0) return false;
    double p = (ph.gamma - 1.) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) * (1.0 / U[0]));
    return !(p < 0);
  } else if constexpr (PHYS == PH_SWE) {
    return U[0] >= 0.;
  }
  return true;
}

template <int PHYS>
__host__ __device__ constexpr bool has_positivity() {
  return PHYS == PH_EULER || PHYS == PH_SWE;
}

// physical flux (fx, fy) of state U
template <int PHYS>
__device__ inline void flux(const PhysParams &ph, const double *U, double *fx, double *fy) {
  if constexpr (PHYS == PH_ADV) {
    fx[0] = ph.ax * U[0];
    fy[0] = ph.ay * U[0];
  } else if constexpr (PHYS == PH_EULER) {
    double ir = frcp(U[0]);
    double u = U[1] * ir, v = U[2] * ir;
    double p = (ph.gamma - 1.) * (U[3] - 0.5 * (U[1] * u + U[2] * v));
    double H = U[3] + p;
    fx[0] = U[1];
    fx[1] = U[1] * u + p;
    fx[2] = U[1] * v;
    fx[3] = u * H;
    fy[0] = U[2];
    fy[1] = U[2] * u;
    fy[2] = U[2] * v + p;
    fy[3] = v * H;
  } else if constexpr (PHYS == PH_SWE) {
    double ir = frcp(U[0]);
    double u = U[1] * ir, v = U[2] * ir;
    double q = 0.5 * ph.G * U[0] * U[0];
    fx[0] = U[1];
    fx[1] = U[1] * u + q;
    fx[2] = U[1] * v;
    fy[0] = U[2];
    fy[1] = U[2] * u;
    fy[2] = U[2] * v + q;
  } else if constexpr (PHYS == PH_TM) {
    double im = 1. / ph.mu, ie = 1. / ph.eps;
    fx[0] = 0.;
    fx[1] = im * U[2];
    fx[2] = ie * U[1];
    fy[0] = -im * U[2];
    fy[1] = 0.;
    fy[2] = -ie * U[0];
  } else {  // TE: material-free; 1/eps, 1/mu scale the element residual
    fx[0] = 0.;
    fx[1] = U[2];
    fx[2] = U[1];
    fy[0] = -U[2];
    fy[1] = 0.;
    fy[2] = -U[0];
  }
}

// Contravariant fluxes of state U for the volume integral:
//   ga = fx * ys - fy * xs,  gb = fy * xr - fx * yr
// so that fx (gx ys - gy yr) + fy (gy xr - gx xs) = gx ga + gy gb
// (eval_volume conserv_kernels.cu:1062-1071).  Euler and shallow water form
// them from the contravariant velocities (u ys - v xs, v xr - u yr) directly,
// which takes about half the operations of the flux followed by the
// transform; the other physics transform their (linear) fluxes.
template <int PHYS>
__device__ __forceinline__ void flux_contra(const PhysParams &ph, const double *U, double xr,
                                            double yr, double xs, double ys, double *ga,
                                            double *gb) {
  if constexpr (PHYS == PH_EULER) {
    const double ir = frcp(U[0]);
    const double u = U[1] * ir, v = U[2] * ir;
    const double p = (ph.gamma - 1.) * (U[3] - 0.5 * (U[1] * u + U[2] * v));
    const double H = U[3] + p;
    const double ua = u * ys - v * xs, ub = v * xr - u * yr;
    ga[0] = U[0] * ua;
    ga[1] = fma(U[1], ua, p * ys);
    ga[2] = fma(U[2], ua, -(p * xs));
    ga[3] = H * ua;
    gb[0] = U[0] * ub;
    gb[1] = fma(U[1], ub, -(p * yr));
    gb[2] = fma(U[2], ub, p * xr);
    gb[3] = H * ub;
  } else if constexpr (PHYS == PH_SWE) {
    const double ir = frcp(U[0]);
    const double u = U[1] * ir, v = U[2] * ir;
    const double q = 0.5 * ph.G * U[0] * U[0];
    const double ua = u * ys - v * xs, ub = v * xr - u * yr;
    ga[0] = U[0] * ua;
    ga[1] = fma(U[1], ua, q * ys);
    ga[2] = fma(U[2], ua, -(q * xs));
    gb[0] = U[0] * ub;
    gb[1] = fma(U[1], ub, -(q * yr));
    gb[2] = fma(U[2], ub, q * xr);
  } else {
    constexpr int N = nvars_of(PHYS);
    double fx[NMAX], fy[NMAX];
    flux<PHYS>(ph, U, fx, fy);
#pragma unroll
    for (int n = 0; n < N; n++) {
      ga[n] = fx[n] * ys - fy[n] * xs;
      gb[n] = fy[n] * xr - fx[n] * yr;
    }
  }
}

// LLF wave-speed bound of one state along n: |u.n| + c (euler.cu:170-213)
template <int PHYS>
__device__ inline double llf_speed(const PhysParams &ph, const double *U, double nx, double ny) {
  if constexpr (PHYS == PH_EULER) {
    double ir = frcp(U[0]);
    double u = U[1] * ir, v = U[2] * ir;
    double p = (ph.gamma - 1.) * (U[3] - 0.5 * (U[1] * u + U[2] * v));
    double c = fsqrt(ph.gamma * p * ir);
    double s = nx * u + ny * v;
    return fabs(fabs(s) + c);
  } else if constexpr (PHYS == PH_SWE) {
    double ir = frcp(U[0]);
    double s = nx * U[1] * ir + ny * U[2] * ir;
    double c = fsqrt(U[0] * ph.G);
    return fabs(fabs(s) + c);
  } else if constexpr (PHYS == PH_ADV) {
    return sqrt(ph.ax * ph.ax + ph.ay * ph.ay);
  } else {
    return fabs(1. / sqrt(ph.eps * ph.mu));
  }
}

// normal flux (F.n) of one state and its LLF speed bound |u.n| + c, the
// pair the LLF flux needs, computed together (one reciprocal, one sqrt)
template <int PHYS>
__device__ __forceinline__ double flux_n_speed(const PhysParams &ph, const double *U, double nx,
                                               double ny, double *Fn) {
  if constexpr (PHYS == PH_EULER) {
    const double ir = frcp(U[0]);
    const double vn = (U[1] * nx + U[2] * ny) * ir;
    const double p = (ph.gamma - 1.) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) * ir);
    Fn[0] = U[0] * vn;
    Fn[1] = U[1] * vn + p * nx;
    Fn[2] = U[2] * vn + p * ny;
    Fn[3] = (U[3] + p) * vn;
    return fabs(fabs(vn) + fsqrt(ph.gamma * p * ir));
  } else if constexpr (PHYS == PH_SWE) {
    const double ir = frcp(U[0]);
    const double vn = (U[1] * nx + U[2] * ny) * ir;
    const double q = 0.5 * ph.G * U[0] * U[0];
    Fn[0] = U[0] * vn;
    Fn[1] = U[1] * vn + q * nx;
    Fn[2] = U[2] * vn + q * ny;
    return fabs(fabs(vn) + fsqrt(U[0] * ph.G));
  } else if constexpr (PHYS == PH_ADV) {
    Fn[0] = (ph.ax * nx + ph.ay * ny) * U[0];
    return sqrt(ph.ax * ph.ax + ph.ay * ph.ay);
  } else if constexpr (PHYS == PH_TM) {
    Fn[0] = -U[2] * ny / ph.mu;
    Fn[1] = U[2] * nx / ph.mu;
    Fn[2] = (U[1] * nx - U[0] * ny) / ph.eps;
    return fabs(1. / sqrt(ph.eps * ph.mu));
  } else {
    Fn[0] = -U[2] * ny;
    Fn[1] = U[2] * nx;
    Fn[2] = U[1] * nx - U[0] * ny;
    return 0.;
  }
}

// Roe flux for shallow water (extension; the reference has none): Roe state
// h~ = (hl + hr) / 2, sqrt(h)-weighted velocity, c~ = sqrt(G h~); the jump
// dU is split exactly on the eigenvectors (1, u~ -+ c~ n), (0, t), so that
// sum_k lambda_k a_k r_k = (F(UR) - F(UL)).n, and
// F = central flux - sum_k |lambda_k| a_k r_k / 2
__device__ inline void swe_roe(const PhysParams &ph, const double *UL, const double *UR, double nx,
                               double ny, double *F) {
  double fxl[3], fyl[3], fxr[3], fyr[3];
  flux<PH_SWE>(ph, UL, fxl, fyl);
  flux<PH_SWE>(ph, UR, fxr, fyr);
  // the Roe averages need positive depths, so the quotients and roots take
  // the refined hardware estimates (1/h is the one the two physical fluxes
  // above formed)
  double G = ph.G, hl = UL[0], hr = UR[0];
  const double irl = frcp(hl), irr = frcp(hr);
  double ul = UL[1] * irl, vl = UL[2] * irl, ur = UR[1] * irr, vr = UR[2] * irr;
  double sl = fsqrt(hl), sr = fsqrt(hr);
  double ht = 0.5 * (hl + hr);
  double isum = frcp(sl + sr);
  double ut = (sl * ul + sr * ur) * isum, vt = (sl * vl + sr * vr) * isum;
  double ct = fsqrt(G * ht);
  double unt = ut * nx + vt * ny, utt = -ut * ny + vt * nx;
  double dh = hr - hl;
  double dqn = (UR[1] - UL[1]) * nx + (UR[2] - UL[2]) * ny;
  double dqt = -(UR[1] - UL[1]) * ny + (UR[2] - UL[2]) * nx;
  double wn = (dqn - unt * dh) * frcp(ct);
  double a1 = 0.5 * (dh - wn), a3 = 0.5 * (dh + wn), a2 = dqt - utt * dh;
  double l1 = fabs(unt - ct), l2 = fabs(unt), l3 = fabs(unt + ct);
  double d0 = l1 * a1 + l3 * a3;
  double d1 = l1 * a1 * (ut - ct * nx) + l2 * a2 * (-ny) + l3 * a3 * (ut + ct * nx);
  double d2 = l1 * a1 * (vt - ct * ny) + l2 * a2 * nx + l3 * a3 * (vt + ct * ny);
  F[0] = 0.5 * ((fxl[0] + fxr[0]) * nx + (fyl[0] + fyr[0]) * ny) - 0.5 * d0;
  F[1] = 0.5 * ((fxl[1] + fxr[1]) * nx + (fyl[1] + fyr[1]) * ny) - 0.5 * d1;
  F[2] = 0.5 * ((fxl[2] + fxr[2]) * nx + (fyl[2] + fyr[2]) * ny) - 0.5 * d2;
}

// exact Riemann (upwind) flux of TE Maxwell with an impedance jump (extension)
__device__ inline void te_upwind(const PhysParams &ph, const double *UL, const double *UR,
                                 double nx, double ny, double epsL, double epsR, double *F) {
  double zl = sqrt(ph.mu / epsL), zr = sqrt(ph.mu / epsR);
  double etl = -UL[0] * ny + UL[1] * nx, etr = -UR[0] * ny + UR[1] * nx;
  double iz = 1.0 / (zl + zr);
  double hs = (zl * UL[2] + zr * UR[2] + (etl - etr)) * iz;
  double es = (zr * etl + zl * etr + zl * zr * (UL[2] - UR[2])) * iz;
  F[0] = -ny * hs;
  F[1] = nx * hs;
  F[2] = es;
}

// physics whose numerical flux is always LLF (Euler, TM Maxwell)
template <int PHYS>
__host__ __device__ constexpr bool llf_only() {
  return PHYS == PH_EULER || PHYS == PH_TM;
}

// numerical flux F(UL, UR) . n (riemann_solver of the plugins)
template <int PHYS>
__device__ inline void riemann(const PhysParams &ph, const double *UL, const double *UR, double nx,
                               double ny, double epsL, double epsR, double *F) {
  constexpr int N = nvars_of(PHYS);
  if constexpr (PHYS == PH_TE) {
    te_upwind(ph, UL, UR, nx, ny, epsL, epsR, F);
    return;
  } else {
    int fl = ph.flux;
    if (fl == FL_DEFAULT) fl = (PHYS == PH_ADV) ? FL_UPWIND : FL_LLF;
    if constexpr (PHYS == PH_EULER || PHYS == PH_TM) fl = FL_LLF;
    if constexpr (PHYS == PH_SWE) {
      if (fl == FL_ROE) {
        swe_roe(ph, UL, UR, nx, ny, F);
        return;
      }
    }
    if constexpr (PHYS == PH_ADV) {
      if (fl == FL_UPWIND) {  // advection.cu:94-116
        double an = nx * ph.ax + ny * ph.ay;
        double u = (an > 0) ? UL[0] : UR[0];
        F[0] = ph.ax * u * nx + ph.ay * u * ny;
        return;
      }
    }
    // LLF: 1/2 [(F.n)(UL) + (F.n)(UR) + lam (UL - UR)], lam = max of the two
    // speed bounds (euler.cu:170-237)
    double Fl[NMAX], Fr[NMAX];
    const double sl = flux_n_speed<PHYS>(ph, UL, nx, ny, Fl);
    const double sr = flux_n_speed<PHYS>(ph, UR, nx, ny, Fr);
    const double lam = sl > sr ? sl : sr;
#pragma unroll
    for (int n = 0; n < N; n++) F[n] = 0.5 * ((Fl[n] + Fr[n]) + lam * (UL[n] - UR[n]));
  }
}

// CFL speed of an element from its cell averages (eval_global_lambda)
template <int PHYS>
__device__ inline double cfl_speed(const PhysParams &ph, const double *U, double eps_e) {
  if constexpr (PHYS == PH_EULER) {
    double p = (ph.gamma - 1.) * (U[3] - (U[1] * U[1] + U[2] * U[2]) / 2. / U[0]);
    double c = sqrt(ph.gamma * p / U[0]);
    double s = sqrt(U[1] * U[1] + U[2] * U[2]) / U[0];
    return (s > 0) ? s + c : -s + c;
  } else if constexpr (PHYS == PH_SWE) {
    double c = sqrt(U[0] * ph.G);
    double s = sqrt(U[1] * U[1] + U[2] * U[2]) / U[0];
    return (s > 0) ? s + c : -s + c;
  } else if constexpr (PHYS == PH_ADV) {
    return sqrt(ph.ax * ph.ax + ph.ay * ph.ay);
  } else if constexpr (PHYS == PH_TE) {
    return fabs(1. / sqrt(eps_e * ph.mu));
  } else {
    return fabs(1. / sqrt(ph.eps * ph.mu));
  }
}

}  // namespace dgk
