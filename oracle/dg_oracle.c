/* dg_oracle.c -- CPU restatement of DGCUDA's DG right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library.
 *
 * Paths below are relative to the reference's src/2d/conservation/.  Every
 * function restates the GPU tree's semantics in its loop and summation
 * order and keeps its SoA layout C[n*K*n_p + i*K + e]; element loops that the
 * GPU runs one thread per element become OpenMP loops whose iterations write
 * disjoint outputs, so the result does not depend on the thread count.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off (oracle/Makefile).  Contraction
 * is off so every product and sum rounds exactly as written.
 *
 * Pinning (tests/): quadrature moments and basis Gram matrix (closed forms and
 * an exact Gram-Schmidt), geometry closed forms, polynomial preservation of
 * advection/exact, free-stream preservation, discrete conservation, LLF and
 * upwind consistency, brute-force gather, RK stability polynomials, the
 * convergence order p+1, and the supersonic-vortex meshes of the reference.
 * Parity unpinned: check_physical under the reference's races (racy by
 * design on the GPU; the sequential order here is one legal interleaving).
 */
#include "dg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ref_tables.h"

#define NMAX 4
#define NPMAX DGT_NPMAX
#define NQMAX 25
#define NQ1MAX 6

struct orc {
    int phys, p, n_p, N, nq, nq1, flux, integ, limiter, posfix;
    double cfl, gamma, G, eps, mu, ax, ay;
    int prob;
    double pp[8];
    /* reference-element tables (basis.cu:168-244) */
    double basis[NPMAX * NQMAX], gx[NPMAX * NQMAX], gy[NPMAX * NQMAX];
    double side[3 * NPMAX * NQ1MAX], vertex[3 * NPMAX];
    double w[NQMAX], r1[NQMAX], r2[NQMAX], w1[NQ1MAX], r1d[NQ1MAX];
    /* mesh (main.cu:205-277) and geometry (conserv_kernels.cu:412-585) */
    int K, S;
    double *V1x, *V1y, *V2x, *V2y, *V3x, *V3y;
    int *es1, *es2, *es3;
    double *sx1, *sy1, *sx2, *sy2;
    int *left, *right, *lsn, *rsn;
    double *J, *xr, *yr, *xs, *ys, *len, *Nx, *Ny, *diam;
    double min_r;
    double *eps_e;
    /* state (conserv_kernels.cu:25-36) */
    double *c, *k1, *k2, *k3, *k4, *rhs_v, *rhs_l, *rhs_r;
};

static int g_par = 1; /* OpenMP on */

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    g_par = (n != 1);
}

/* ---------------------------------------------------------------- basis */

/* phi, phi_x, phi_y: basis.cu:68-143 (graded monomials, host pow) */
static double phi(double r, double s, int n) {
    double R[64];
    int k = 0, power = 0, j, i;
    double sum = 0.;
    while (k <= n) {
        for (j = 0; j <= power; j++) {
            R[k] = pow(r, power - j) * pow(s, j);
            k++;
        }
        power++;
    }
    for (i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
    return sum;
}

static double phi_x(double r, double s, int n) {
    double R[64];
    int k = 1, power = 1, j, i;
    double sum = 0.;
    R[0] = 0;
    while (k <= n) {
        for (j = 0; j <= power; j++) {
            R[k] = (power - j) * pow(r, power - j - 1) * pow(s, j);
            if (j == power) R[k] = 0;
            k++;
        }
        power++;
    }
    for (i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
    return sum;
}

static double phi_y(double r, double s, int n) {
    double R[64];
    int k = 1, power = 1, j, i;
    double sum = 0.;
    R[0] = 0;
    while (k <= n) {
        for (j = 0; j <= power; j++) {
            R[k] = pow(r, power - j) * j * pow(s, j - 1);
            if (j == 0) R[k] = 0;
            k++;
        }
        power++;
    }
    for (i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
    return sum;
}

/* set_quadrature main.cu:124-194 + preval_basis basis.cu:168-244 */
static void build_tables(orc *o) {
    int p = o->p, i, j;
    const double *q2 = dgt_q2[p], *q1 = dgt_q1[p];
    o->nq = dgt_nq2[p];
    o->nq1 = dgt_nq1[p];
    for (i = 0; i < o->nq; i++) {
        o->r1[i] = q2[3 * i];
        o->r2[i] = q2[3 * i + 1];
        o->w[i] = q2[3 * i + 2] / 2.;
    }
    for (i = 0; i < o->nq1; i++) {
        o->r1d[i] = q1[2 * i];
        o->w1[i] = q1[2 * i + 1];
    }
    for (i = 0; i < o->n_p; i++) {
        o->vertex[i * 3 + 0] = phi(0., 0., i);
        o->vertex[i * 3 + 1] = phi(1., 0., i);
        o->vertex[i * 3 + 2] = phi(0., 1., i);
        for (j = 0; j < o->nq; j++) {
            o->basis[i * o->nq + j] = phi(o->r1[j], o->r2[j], i);
            o->gx[i * o->nq + j] = o->w[j] * phi_x(o->r1[j], o->r2[j], i);
            o->gy[i * o->nq + j] = o->w[j] * phi_y(o->r1[j], o->r2[j], i);
        }
        for (j = 0; j < o->nq1; j++) {
            const double *sr = o->r1d;
            int m = o->nq1;
            o->side[0 * (m * o->n_p) + i * m + j] = phi(0.5 + 0.5 * sr[j], 0., i);
            o->side[1 * (m * o->n_p) + i * m + j] = phi((1. - sr[j]) / 2., (1. + sr[j]) / 2., i);
            o->side[2 * (m * o->n_p) + i * m + j] = phi(0., 0.5 + 0.5 * sr[m - 1 - j], i);
        }
    }
}

/* ------------------------------------------------------------- problems */

/* get_coordinates_2d conserv_kernels.cu:231-235 */
static void coords2d(const orc *o, double *X, const double *V, int j) {
    X[0] = V[2] * o->r1[j] + V[4] * o->r2[j] + V[0] * (1 - o->r1[j] - o->r2[j]);
    X[1] = V[3] * o->r1[j] + V[5] * o->r2[j] + V[1] * (1 - o->r1[j] - o->r2[j]);
}

/* get_coordinates_1d conserv_kernels.cu:238-261 */
static void coords1d(const orc *o, double *X, const double *V, int j, int left_side) {
    double a = 0., b = 0.;
    switch (left_side) {
    case 0: a = 0.5 + 0.5 * o->r1d[j]; b = 0.; break;
    case 1: a = (1. - o->r1d[j]) / 2.; b = (1. + o->r1d[j]) / 2.; break;
    case 2: a = 0.; b = 0.5 + 0.5 * o->r1d[o->nq1 - 1 - j]; break;
    }
    X[0] = V[2] * a + V[4] * b + V[0] * (1 - a - b);
    X[1] = V[3] * a + V[5] * b + V[1] * (1 - a - b);
}

static double euler_E(double g, double rho, double u, double v, double p) {
    return p / (g - 1.) + 0.5 * rho * (u * u + v * v);
}

/* Shu's isentropic vortex (extension; the configs' manufactured solution):
   pp = beta, x0, y0, u_inf, v_inf, period L (L <= 0: no wrap). */
static void vortex_state(const orc *o, double *U, double x, double y, double t) {
    const double *pp = o->pp;
    double g = o->gamma, L = pp[5];
    double dx = x - (pp[1] + pp[3] * t), dy = y - (pp[2] + pp[4] * t);
    if (L > 0) {
        dx -= L * floor(dx / L + 0.5);
        dy -= L * floor(dy / L + 0.5);
    }
    double r2 = dx * dx + dy * dy;
    double f = pp[0] / (2. * M_PI) * exp(0.5 * (1. - r2));
    double u = pp[3] - f * dy, v = pp[4] + f * dx;
    double T = 1. - (g - 1.) * pp[0] * pp[0] / (8. * g * M_PI * M_PI) * exp(1. - r2);
    double rho = pow(T, 1. / (g - 1.));
    double p = pow(rho, g);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = euler_E(g, rho, u, v, p);
}

/* euler/supersonic/supersonic.cu:27-37 (GAMMA 1.4, MACH 2.25 as #defined) */
static void supersonic_U0(double *U, double x, double y) {
    const double GAMMA = 1.4, MACH = 2.25;
    double r = sqrt(x * x + y * y);
    U[0] = pow(1 + 1.0125 * (1. - 1. / (r * r)), 2.5);
    U[1] = U[0] * sin(atan(y / x)) * MACH / r;
    U[2] = U[0] * -cos(atan(y / x)) * MACH / r;
    double p = (1.0 / GAMMA) * pow(U[0], GAMMA);
    U[3] = 0.5 * U[0] * (MACH * MACH / (r * r)) + p * (1. / (GAMMA - 1.));
}

/* rotating hill, advection/rotatinghill/rotatinghill.cu: U0 :25-39, the
   inflow (the hill rotated to time t) :51-61 with cospi/sinpi written as
   cos/sin of pi * (2t), the velocity 2 pi (-y, x) :46-49 */
static double rothill_u0(double x, double y) {
    double x0 = 0.2, y0 = 0.0, r = 0.15;
    return 5 * exp(-(pow(x - x0, 2) + pow(y - y0, 2)) / (2 * r * r));
}

static double rothill_inflow(double x, double y, double t) {
    double c = cos(M_PI * (2 * t)), s = sin(M_PI * (2 * t));
    return rothill_u0(x * c + y * s, -x * s + y * c);
}

/* the advection velocity at (x, y): get_velocity of the plugin */
static void velocity(const orc *o, const double *X, double *A) {
    if (o->prob == ORC_PROB_ADV_ROTHILL) {
        A[0] = -2 * M_PI * X[1];
        A[1] = 2 * M_PI * X[0];
    } else {
        A[0] = o->ax;
        A[1] = o->ay;
    }
}

/* the exact / initial state of the problem at (x, y, t) */
static void exact_state(const orc *o, double *U, double x, double y, double t) {
    const double *pp = o->pp;
    int n;
    for (n = 0; n < NMAX; n++) U[n] = 0.;
    switch (o->prob) {
    case ORC_PROB_ADV_SINSIN: {
        double X = x - o->ax * t, Y = y - o->ay * t;
        U[0] = sin(2. * M_PI * X) * sin(2. * M_PI * Y);
        break;
    }
    case ORC_PROB_ADV_LINEAR: /* advection/exact/exact.cu:25-27 */
        U[0] = x - y;
        break;
    case ORC_PROB_EULER_VORTEX:
        vortex_state(o, U, x, y, t);
        break;
    case ORC_PROB_EULER_SUPERSONIC:
        supersonic_U0(U, x, y);
        break;
    case ORC_PROB_EULER_CONST:
        U[0] = pp[0];
        U[1] = pp[0] * pp[1];
        U[2] = pp[0] * pp[2];
        U[3] = euler_E(o->gamma, pp[0], pp[1], pp[2], pp[3]);
        break;
    case ORC_PROB_SWE_BUMP: { /* pp = h0, A, cx, cy, sigma, k: h0 + A exp(-d^2/(k s^2)) */
        double d2 = (x - pp[2]) * (x - pp[2]) + (y - pp[3]) * (y - pp[3]);
        U[0] = pp[0] + pp[1] * exp(-d2 / (pp[5] * pp[4] * pp[4]));
        break;
    }
    case ORC_PROB_MAXWELL_PULSE: { /* pp = A, cx, cy, sigma: last variable */
        double d2 = (x - pp[1]) * (x - pp[1]) + (y - pp[2]) * (y - pp[2]);
        U[o->N - 1] = pp[0] * exp(-d2 / (pp[3] * pp[3]));
        break;
    }
    case ORC_PROB_CONST:
        for (n = 0; n < o->N; n++) U[n] = pp[n];
        break;
    case ORC_PROB_ADV_ROTHILL: /* U_exact ends in U0(U, x, y), rotatinghill.cu:91-98 */
        U[0] = rothill_u0(x, y);
        break;
    }
}

/* U_reflection of the problem plugins:
   supersonic.cu:65-87 (radial wall), shallowwater/test/test.cu:63-77 and
   euler/lshock/lshock.cu:57-71 (slip wall), maxwells/cmirror/cmirror.cu:71-84
   (TM conductor).  TE conductor (extension): E_t flips, E_n and Hz kept. */
static void reflect_state(const orc *o, const double *UL, double *UR, double x, double y,
                          double nx, double ny) {
    double dot;
    switch (o->phys) {
    case ORC_ADVECTION:
        UR[0] = UL[0];
        break;
    case ORC_EULER:
        if (o->prob == ORC_PROB_EULER_SUPERSONIC) {
            double Nx, Ny;
            UR[0] = UL[0];
            UR[3] = UL[3];
            dot = sqrt(x * x + y * y);
            Nx = x / dot;
            Ny = y / dot;
            if (Nx * nx + Ny * ny < 0) {
                Nx *= -1;
                Ny *= -1;
            }
            UR[1] = (UL[1] * Ny - UL[2] * Nx) * Ny;
            UR[2] = -(UL[1] * Ny - UL[2] * Nx) * Nx;
        } else {
            UR[0] = UL[0];
            UR[3] = UL[3];
            dot = UL[1] * nx + UL[2] * ny;
            UR[1] = UL[1] - 2 * dot * nx;
            UR[2] = UL[2] - 2 * dot * ny;
        }
        break;
    case ORC_SWE:
        UR[0] = UL[0];
        dot = UL[1] * nx + UL[2] * ny;
        UR[1] = UL[1] - 2 * dot * nx;
        UR[2] = UL[2] - 2 * dot * ny;
        break;
    case ORC_MAXWELL_TM:
        UR[2] = UL[2];
        dot = UL[0] * nx + UL[1] * ny;
        UR[0] = -UL[0] + 2 * dot * nx;
        UR[1] = -UL[1] + 2 * dot * ny;
        break;
    case ORC_MAXWELL_TE:
        UR[2] = UL[2];
        dot = UL[0] * nx + UL[1] * ny;
        UR[0] = -UL[0] + 2 * dot * nx;
        UR[1] = -UL[1] + 2 * dot * ny;
        break;
    }
}

/* eval_boundary + inflow/outflow/reflecting_boundary, conserv_kernels.cu:594-626,820-843.
   The outflow of advection/exact and advection/rotatinghill leaves U_right at
   zero (exact.cu:50-52, rotatinghill.cu:69-71); the rotating hill's inflow is
   the rotated hill. */
static void boundary_state(const orc *o, const double *UL, double *UR, const double *V,
                           double nx, double ny, int j, int left_side, double t, int right_idx) {
    double X[2];
    coords1d(o, X, V, j, left_side);
    switch (right_idx) {
    case -1: reflect_state(o, UL, UR, X[0], X[1], nx, ny); break;
    case -2:
        if (o->prob != ORC_PROB_ADV_LINEAR && o->prob != ORC_PROB_ADV_ROTHILL)
            exact_state(o, UR, X[0], X[1], t);
        break;
    case -3:
        if (o->prob == ORC_PROB_ADV_ROTHILL)
            UR[0] = rothill_inflow(X[0], X[1], t);
        else
            exact_state(o, UR, X[0], X[1], t);
        break;
    }
}

/* -------------------------------------------------------------- physics */

/* euler.cu:61-70 */
static double euler_p(const orc *o, const double *U) {
    return (o->gamma - 1.) * (U[3] - (U[1] * U[1] + U[2] * U[2]) / 2. / U[0]);
}

/* is_physical euler.cu:76-86; shallowwater.cu:48-50 */
static int is_physical(const orc *o, const double *U) {
    if (o->phys == ORC_EULER) {
        if (U[0] < 0) return 0;
        if (euler_p(o, U) < 0) return 0;
        return 1;
    }
    if (o->phys == ORC_SWE) return U[0] >= 0.;
    return 1;
}

static int has_positivity(const orc *o) { return o->phys == ORC_EULER || o->phys == ORC_SWE; }

/* check_physical euler.cu:92-116 (SWE: shallowwater.cu:56-80 without its
   out-of-bounds write of variable 3) */
static void check_physical(const orc *o, double *Cg, double *Cl, double *U, int idx) {
    int i, n, K = o->K, np = o->n_p;
    if (o->posfix || !has_positivity(o)) return;
    if (!is_physical(o, U)) {
        for (i = 1; i < np; i++)
            for (n = 0; n < o->N; n++) {
                Cg[K * np * n + i * K + idx] = 0.;
                Cl[np * n + i] = 0.;
            }
        for (n = 0; n < o->N; n++) U[n] = Cl[np * n + 0] * o->basis[0];
    }
}

/* eval_flux: advection.cu:51-71, euler.cu:138-160, shallowwater.cu:91-109,
   maxwells.cu:67-84.  TE (extension): material-free fluxes; 1/eps and 1/mu
   are applied to the element residual in eval_rhs. */
static void eval_flux(const orc *o, const double *U, double *fx, double *fy, const double *A) {
    switch (o->phys) {
    case ORC_ADVECTION: /* A: the velocity at the point */
        fx[0] = A[0] * U[0];
        fy[0] = A[1] * U[0];
        break;
    case ORC_EULER: {
        double p = euler_p(o, U);
        double rho = U[0], rhou = U[1], rhov = U[2], E = U[3];
        fx[0] = rhou;
        fx[1] = rhou * rhou / rho + p;
        fx[2] = rhou * rhov / rho;
        fx[3] = rhou * (E + p) / rho;
        fy[0] = rhov;
        fy[1] = rhou * rhov / rho;
        fy[2] = rhov * rhov / rho + p;
        fy[3] = rhov * (E + p) / rho;
        break;
    }
    case ORC_SWE: {
        double h = U[0], uh = U[1], vh = U[2], G = o->G;
        fx[0] = uh;
        fx[1] = uh * uh / h + 0.5 * G * h * h;
        fx[2] = uh * vh / h;
        fy[0] = vh;
        fy[1] = uh * vh / h;
        fy[2] = vh * vh / h + 0.5 * G * h * h;
        break;
    }
    case ORC_MAXWELL_TM:
        fx[0] = 0.;
        fx[1] = 1. / o->mu * U[2];
        fx[2] = 1. / o->eps * U[1];
        fy[0] = -1. / o->mu * U[2];
        fy[1] = 0.;
        fy[2] = -1. / o->eps * U[0];
        break;
    case ORC_MAXWELL_TE:
        fx[0] = 0.;
        fx[1] = U[2];
        fx[2] = U[1];
        fy[0] = -U[2];
        fy[1] = 0.;
        fy[2] = -U[0];
        break;
    }
}

/* wave speed c: euler.cu:122-127, shallowwater.cu:44-46, maxwells.cu:50-52 */
static double eval_c(const orc *o, const double *U) {
    if (o->phys == ORC_EULER) {
        double p = euler_p(o, U);
        return sqrt(o->gamma * p / U[0]);
    }
    if (o->phys == ORC_SWE) return sqrt(U[0] * o->G);
    return 1. / sqrt(o->eps * o->mu);
}

/* eval_lambda: euler.cu:170-213, shallowwater.cu:118-161 (abs() on doubles is
   fabs in device code), maxwells.cu:93-105, advection.cu:80-92 */
static double eval_lambda(const orc *o, const double *UL, const double *UR, double nx,
                          double ny, const double *A) {
    if (o->phys == ORC_EULER || o->phys == ORC_SWE) {
        double cl = eval_c(o, UL), cr = eval_c(o, UR);
        double ul = UL[1] / UL[0], vl = UL[2] / UL[0];
        double ur = UR[1] / UR[0], vr = UR[2] / UR[0];
        double sl = nx * ul + ny * vl, sr = nx * ur + ny * vr;
        double lm = (sl > 0.) ? sl + cl : -sl + cl;
        double rm = (sr > 0.) ? sr + cr : -sr + cr;
        return (fabs(lm) > fabs(rm)) ? fabs(lm) : fabs(rm);
    }
    if (o->phys == ORC_ADVECTION) return sqrt(A[0] * A[0] + A[1] * A[1]);
    {
        double cl = fabs(eval_c(o, UL)), cr = fabs(eval_c(o, UR));
        return (cl > cr) ? cl : cr;
    }
}

/* Roe flux for shallow water (extension, the ensemble config's flux; the
   reference has none).  Roe state: h~ = (hl + hr) / 2, velocity averaged
   with sqrt(h) weights, c~ = sqrt(G h~).  The conserved jump dU is split
   exactly on the eigenvectors r1 = (1, u~ - c~ n), r2 = (0, t), r3 = (1, u~ +
   c~ n) (t = (-ny, nx)): a1 + a3 = dh, c~ (a3 - a1) = d(hu.n) - u~.n dh,
   a2 = d(hu.t) - u~.t dh, so that sum_k lambda_k a_k r_k = (F(UR) - F(UL)).n
   (the Roe property) and F = central flux - sum_k |lambda_k| a_k r_k / 2. */
static void swe_roe(const orc *o, double *F, const double *UL, const double *UR, double nx,
                    double ny) {
    double fxl[NMAX], fyl[NMAX], fxr[NMAX], fyr[NMAX];
    double G = o->G;
    eval_flux(o, UL, fxl, fyl, NULL);
    eval_flux(o, UR, fxr, fyr, NULL);
    double hl = UL[0], hr = UR[0];
    double ul = UL[1] / hl, vl = UL[2] / hl, ur = UR[1] / hr, vr = UR[2] / hr;
    double sl = sqrt(hl), sr = sqrt(hr);
    double ht = 0.5 * (hl + hr);
    double ut = (sl * ul + sr * ur) / (sl + sr);
    double vt = (sl * vl + sr * vr) / (sl + sr);
    double ct = sqrt(G * ht);
    double unt = ut * nx + vt * ny, utt = -ut * ny + vt * nx;
    double dh = hr - hl;
    double dqn = (UR[1] - UL[1]) * nx + (UR[2] - UL[2]) * ny;
    double dqt = -(UR[1] - UL[1]) * ny + (UR[2] - UL[2]) * nx;
    double wn = (dqn - unt * dh) / ct;
    double a1 = 0.5 * (dh - wn);
    double a3 = 0.5 * (dh + wn);
    double a2 = dqt - utt * dh;
    double l1 = fabs(unt - ct), l2 = fabs(unt), l3 = fabs(unt + ct);
    double d0 = l1 * a1 + l3 * a3;
    double d1 = l1 * a1 * (ut - ct * nx) + l2 * a2 * (-ny) + l3 * a3 * (ut + ct * nx);
    double d2 = l1 * a1 * (vt - ct * ny) + l2 * a2 * nx + l3 * a3 * (vt + ct * ny);
    F[0] = 0.5 * ((fxl[0] + fxr[0]) * nx + (fyl[0] + fyr[0]) * ny) - 0.5 * d0;
    F[1] = 0.5 * ((fxl[1] + fxr[1]) * nx + (fyl[1] + fyr[1]) * ny) - 0.5 * d1;
    F[2] = 0.5 * ((fxl[2] + fxr[2]) * nx + (fyl[2] + fyr[2]) * ny) - 0.5 * d2;
}

/* Upwind (exact Riemann) flux for TE Maxwell with an impedance jump
   (extension): w+ = E_t + Z H travels along n, w- = E_t - Z H against it. */
static void te_upwind(const orc *o, double *F, const double *UL, const double *UR, double nx,
                      double ny, double epsL, double epsR) {
    double zl = sqrt(o->mu / epsL), zr = sqrt(o->mu / epsR);
    double etl = -UL[0] * ny + UL[1] * nx, etr = -UR[0] * ny + UR[1] * nx;
    double hs = (zl * UL[2] + zr * UR[2] + (etl - etr)) / (zl + zr);
    double es = (zr * etl + zl * etr + zl * zr * (UL[2] - UR[2])) / (zl + zr);
    F[0] = -ny * hs;
    F[1] = nx * hs;
    F[2] = es;
}

/* riemann_solver: euler.cu:217-237 (LLF), advection.cu:94-116 (upwind),
   shallowwater.cu:165-185 (LLF), maxwells.cu:108-163 (LLF) */
static void riemann(const orc *o, double *F, const double *UL, const double *UR, double nx,
                    double ny, double epsL, double epsR, const double *A) {
    int n, fl = o->flux;
    double fxl[NMAX], fyl[NMAX], fxr[NMAX], fyr[NMAX];
    if (fl == ORC_FLUX_DEFAULT) fl = (o->phys == ORC_ADVECTION || o->phys == ORC_MAXWELL_TE)
                                         ? ORC_FLUX_UPWIND : ORC_FLUX_LLF;
    if (o->phys == ORC_SWE && fl == ORC_FLUX_ROE) {
        swe_roe(o, F, UL, UR, nx, ny);
        return;
    }
    if (o->phys == ORC_MAXWELL_TE) {
        te_upwind(o, F, UL, UR, nx, ny, epsL, epsR);
        return;
    }
    eval_flux(o, UL, fxl, fyl, A);
    eval_flux(o, UR, fxr, fyr, A);
    if (o->phys == ORC_ADVECTION && fl == ORC_FLUX_UPWIND) {
        if (nx * A[0] + ny * A[1] > 0)
            F[0] = fxl[0] * nx + fyl[0] * ny;
        else
            F[0] = fxr[0] * nx + fyr[0] * ny;
        return;
    }
    double lambda = eval_lambda(o, UL, UR, nx, ny, A);
    for (n = 0; n < o->N; n++)
        F[n] = 0.5 * ((fxl[n] + fxr[n]) * nx + (fyl[n] + fyr[n]) * ny + lambda * (UL[n] - UR[n]));
}

static double elem_eps(const orc *o, int e) {
    return (o->phys == ORC_MAXWELL_TE && o->eps_e) ? o->eps_e[e] : o->eps;
}

/* ---------------------------------------------------------------- state */

static void *xalloc(size_t n) {
    void *p = calloc(n ? n : 1, 1);
    return p;
}

orc *orc_new(int phys, int p, int flux, int integ, double cfl, int limiter, int posfix) {
    if (p < 0 || p > DGT_PMAX || phys < 0 || phys > 4) return NULL;
    orc *o = (orc *)xalloc(sizeof(orc));
    o->phys = phys;
    o->p = p;
    o->n_p = (p + 1) * (p + 2) / 2;
    o->N = (phys == ORC_ADVECTION) ? 1 : (phys == ORC_EULER ? 4 : 3);
    o->flux = flux;
    o->integ = integ;
    o->cfl = cfl;
    o->limiter = limiter;
    o->posfix = posfix;
    o->gamma = 1.4;
    o->G = 9.8;
    o->eps = 1.;
    o->mu = 1.;
    o->ax = 1.;
    o->ay = 1.;
    build_tables(o);
    return o;
}

static void free_mesh(orc *o) {
    void **ptrs[] = {(void **)&o->V1x, (void **)&o->V1y, (void **)&o->V2x, (void **)&o->V2y,
                     (void **)&o->V3x, (void **)&o->V3y, (void **)&o->es1, (void **)&o->es2,
                     (void **)&o->es3, (void **)&o->sx1, (void **)&o->sy1, (void **)&o->sx2,
                     (void **)&o->sy2, (void **)&o->left, (void **)&o->right, (void **)&o->lsn,
                     (void **)&o->rsn, (void **)&o->J, (void **)&o->xr, (void **)&o->yr,
                     (void **)&o->xs, (void **)&o->ys, (void **)&o->len, (void **)&o->Nx,
                     (void **)&o->Ny, (void **)&o->diam, (void **)&o->c, (void **)&o->k1,
                     (void **)&o->k2, (void **)&o->k3, (void **)&o->k4, (void **)&o->rhs_v,
                     (void **)&o->rhs_l, (void **)&o->rhs_r, (void **)&o->eps_e};
    size_t i;
    for (i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) {
        free(*ptrs[i]);
        *ptrs[i] = NULL;
    }
}

void orc_free(orc *o) {
    if (!o) return;
    free_mesh(o);
    free(o);
}

int orc_set_params(orc *o, double gamma, double G, double eps, double mu, double ax, double ay) {
    o->gamma = gamma;
    o->G = G;
    o->eps = eps;
    o->mu = mu;
    o->ax = ax;
    o->ay = ay;
    return 0;
}

int orc_set_problem(orc *o, int prob, const double *pp, int npp) {
    int i;
    o->prob = prob;
    for (i = 0; i < 8; i++) o->pp[i] = (pp && i < npp) ? pp[i] : 0.;
    return 0;
}

/* the precompute order of run_dgcuda, main.cu:698-769 */
static void geometry(orc *o) {
    int K = o->K, S = o->S, idx;
    double min_r;
    /* preval_inscribed_circles conserv_kernels.cu:428-445 (FP32 sqrtf) */
    for (idx = 0; idx < K; idx++) {
        double a, b, c, k;
        a = sqrtf((float)((o->V1x[idx] - o->V2x[idx]) * (o->V1x[idx] - o->V2x[idx]) +
                          (o->V1y[idx] - o->V2y[idx]) * (o->V1y[idx] - o->V2y[idx])));
        b = sqrtf((float)((o->V2x[idx] - o->V3x[idx]) * (o->V2x[idx] - o->V3x[idx]) +
                          (o->V2y[idx] - o->V3y[idx]) * (o->V2y[idx] - o->V3y[idx])));
        c = sqrtf((float)((o->V1x[idx] - o->V3x[idx]) * (o->V1x[idx] - o->V3x[idx]) +
                          (o->V1y[idx] - o->V3y[idx]) * (o->V1y[idx] - o->V3y[idx])));
        k = 0.5 * (a + b + c);
        o->diam[idx] = 2 * sqrtf((float)(k * (k - a) * (k - b) * (k - c))) / k;
    }
    /* host min, main.cu:722-736 */
    min_r = o->diam[0];
    for (idx = 1; idx < K; idx++) min_r = (o->diam[idx] < min_r) ? o->diam[idx] : min_r;
    o->min_r = min_r;
    /* preval_jacobian conserv_kernels.cu:452-473 */
    for (idx = 0; idx < K; idx++) {
        double x1 = o->V1x[idx], y1 = o->V1y[idx], x2 = o->V2x[idx], y2 = o->V2y[idx];
        double x3 = o->V3x[idx], y3 = o->V3y[idx];
        o->J[idx] = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    }
    /* preval_side_length conserv_kernels.cu:412-421 */
    for (idx = 0; idx < S; idx++) {
        double dx = o->sx1[idx] - o->sx2[idx], dy = o->sy1[idx] - o->sy2[idx];
        o->len[idx] = sqrtf((float)(dx * dx + dy * dy));
    }
    /* preval_normals conserv_kernels.cu:481-511 */
    for (idx = 0; idx < S; idx++) {
        double x = o->sx2[idx] - o->sx1[idx], y = o->sy2[idx] - o->sy1[idx];
        double length = sqrtf((float)(x * x + y * y));
        o->Nx[idx] = -y / length;
        o->Ny[idx] = x / length;
    }
    /* preval_normals_direction conserv_kernels.cu:513-569 */
    for (idx = 0; idx < S; idx++) {
        int l = o->left[idx], side = o->lsn[idx];
        double tx = 0, ty = 0, ix = 0, iy = 0;
        switch (side) {
        case 0:
            tx = o->V3x[l]; ty = o->V3y[l];
            ix = (o->V1x[l] + o->V2x[l]) / 2.; iy = (o->V1y[l] + o->V2y[l]) / 2.;
            break;
        case 1:
            tx = o->V1x[l]; ty = o->V1y[l];
            ix = (o->V2x[l] + o->V3x[l]) / 2.; iy = (o->V2y[l] + o->V3y[l]) / 2.;
            break;
        case 2:
            tx = o->V2x[l]; ty = o->V2y[l];
            ix = (o->V1x[l] + o->V3x[l]) / 2.; iy = (o->V1y[l] + o->V3y[l]) / 2.;
            break;
        }
        double dot = o->Nx[idx] * (tx - ix) + o->Ny[idx] * (ty - iy);
        if (dot > 0) {
            o->Nx[idx] *= -1;
            o->Ny[idx] *= -1;
        }
    }
    /* preval_partials conserv_kernels.cu:571-585 */
    for (idx = 0; idx < K; idx++) {
        o->xr[idx] = o->V2x[idx] - o->V1x[idx];
        o->yr[idx] = o->V2y[idx] - o->V1y[idx];
        o->xs[idx] = o->V3x[idx] - o->V1x[idx];
        o->ys[idx] = o->V3y[idx] - o->V1y[idx];
    }
}

#define DUPD(dst, src, n)                                          \
    do {                                                           \
        dst = (double *)xalloc(sizeof(double) * (size_t)(n));      \
        memcpy(dst, src, sizeof(double) * (size_t)(n));            \
    } while (0)
#define DUPI(dst, src, n)                                          \
    do {                                                           \
        dst = (int *)xalloc(sizeof(int) * (size_t)(n));            \
        memcpy(dst, src, sizeof(int) * (size_t)(n));               \
    } while (0)

int orc_set_mesh(orc *o, int K, const double *V, const int *es, int S, const double *sxy,
                 const int *left, const int *right, const int *lsn, const int *rsn) {
    size_t dofs;
    free_mesh(o);
    o->K = K;
    o->S = S;
    DUPD(o->V1x, V + 0 * (size_t)K, K);
    DUPD(o->V1y, V + 1 * (size_t)K, K);
    DUPD(o->V2x, V + 2 * (size_t)K, K);
    DUPD(o->V2y, V + 3 * (size_t)K, K);
    DUPD(o->V3x, V + 4 * (size_t)K, K);
    DUPD(o->V3y, V + 5 * (size_t)K, K);
    DUPI(o->es1, es + 0 * (size_t)K, K);
    DUPI(o->es2, es + 1 * (size_t)K, K);
    DUPI(o->es3, es + 2 * (size_t)K, K);
    DUPD(o->sx1, sxy + 0 * (size_t)S, S);
    DUPD(o->sy1, sxy + 1 * (size_t)S, S);
    DUPD(o->sx2, sxy + 2 * (size_t)S, S);
    DUPD(o->sy2, sxy + 3 * (size_t)S, S);
    DUPI(o->left, left, S);
    DUPI(o->right, right, S);
    DUPI(o->lsn, lsn, S);
    DUPI(o->rsn, rsn, S);
    o->J = (double *)xalloc(sizeof(double) * K);
    o->xr = (double *)xalloc(sizeof(double) * K);
    o->yr = (double *)xalloc(sizeof(double) * K);
    o->xs = (double *)xalloc(sizeof(double) * K);
    o->ys = (double *)xalloc(sizeof(double) * K);
    o->diam = (double *)xalloc(sizeof(double) * K);
    o->len = (double *)xalloc(sizeof(double) * S);
    o->Nx = (double *)xalloc(sizeof(double) * S);
    o->Ny = (double *)xalloc(sizeof(double) * S);
    dofs = (size_t)K * o->n_p * o->N;
    o->c = (double *)xalloc(sizeof(double) * dofs);
    o->k1 = (double *)xalloc(sizeof(double) * dofs);
    o->k2 = (double *)xalloc(sizeof(double) * dofs);
    o->k3 = (double *)xalloc(sizeof(double) * dofs);
    o->k4 = (double *)xalloc(sizeof(double) * dofs);
    o->rhs_v = (double *)xalloc(sizeof(double) * dofs);
    o->rhs_l = (double *)xalloc(sizeof(double) * (size_t)S * o->n_p * o->N);
    o->rhs_r = (double *)xalloc(sizeof(double) * (size_t)S * o->n_p * o->N);
    geometry(o);
    return 0;
}

int orc_set_eps(orc *o, const double *eps_elem) {
    free(o->eps_e);
    o->eps_e = NULL;
    if (eps_elem) DUPD(o->eps_e, eps_elem, o->K);
    return 0;
}

static void elem_V(const orc *o, int e, double *V) {
    V[0] = o->V1x[e];
    V[1] = o->V1y[e];
    V[2] = o->V2x[e];
    V[3] = o->V2y[e];
    V[4] = o->V3x[e];
    V[5] = o->V3y[e];
}

/* init_conditions conserv_kernels.cu:201-228 with evalU0 (euler.cu:31-55,
   advection.cu:23-41, ...); only the N variables of the physics are written */
int orc_project_ic(orc *o) {
    int K = o->K, np = o->n_p, idx;
#pragma omp parallel for if (g_par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        double V[6], X[2], u0[NMAX], U[NMAX];
        int i, j, n;
        elem_V(o, idx, V);
        for (i = 0; i < np; i++) {
            for (n = 0; n < o->N; n++) U[n] = 0.;
            for (j = 0; j < o->nq; j++) {
                coords2d(o, X, V, j);
                exact_state(o, u0, X[0], X[1], 0.);
                for (n = 0; n < o->N; n++) U[n] += o->w[j] * u0[n] * o->basis[i * o->nq + j];
            }
            for (n = 0; n < o->N; n++) o->c[K * np * n + i * K + idx] = U[n];
        }
    }
    return 0;
}

int orc_dofs(const orc *o) { return o->K * o->n_p * o->N; }
int orc_nvars(const orc *o) { return o->N; }
int orc_np(const orc *o) { return o->n_p; }
int orc_nq(const orc *o) { return o->nq; }
int orc_nq1(const orc *o) { return o->nq1; }
double orc_min_r(const orc *o) { return o->min_r; }

int orc_set_coeffs(orc *o, const double *C) {
    memcpy(o->c, C, sizeof(double) * (size_t)orc_dofs(o));
    return 0;
}
int orc_get_coeffs(const orc *o, double *C) {
    memcpy(C, o->c, sizeof(double) * (size_t)orc_dofs(o));
    return 0;
}

/* Element-local positivity fix (posfix = 1): an element whose state is
   unphysical at any volume or side quadrature point keeps only its means.
   This is the race-free form of check_physical (euler.cu:92-116), applied
   to each stage input before it is evaluated. */
static void positivity_prepass(orc *o, double *C) {
    int K = o->K, np = o->n_p, idx;
    if (!o->posfix || !has_positivity(o)) return;
#pragma omp parallel for if (g_par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        double U[NMAX];
        int i, j, n, s, bad = 0;
        for (j = 0; j < o->nq && !bad; j++) {
            for (n = 0; n < o->N; n++) U[n] = 0.;
            for (i = 0; i < np; i++)
                for (n = 0; n < o->N; n++) U[n] += C[K * np * n + i * K + idx] * o->basis[o->nq * i + j];
            if (!is_physical(o, U)) bad = 1;
        }
        for (s = 0; s < 3 && !bad; s++)
            for (j = 0; j < o->nq1 && !bad; j++) {
                for (n = 0; n < o->N; n++) U[n] = 0.;
                for (i = 0; i < np; i++)
                    for (n = 0; n < o->N; n++)
                        U[n] += C[K * np * n + i * K + idx] * o->side[s * np * o->nq1 + i * o->nq1 + j];
                if (!is_physical(o, U)) bad = 1;
            }
        if (bad)
            for (i = 1; i < np; i++)
                for (n = 0; n < o->N; n++) C[K * np * n + i * K + idx] = 0.;
    }
}

/* eval_surface + eval_left_right, conserv_kernels.cu:850-985 */
static void eval_surface(orc *o, double *C, double t) {
    int S = o->S, K = o->K, np = o->n_p, nq1 = o->nq1, N = o->N, idx;
    int par = g_par && (o->posfix || !has_positivity(o));
#pragma omp parallel for if (par) schedule(static)
    for (idx = 0; idx < S; idx++) {
        int i, j, n;
        int left_idx = o->left[idx], right_idx = o->right[idx];
        int left_side = o->lsn[idx], right_side = o->rsn[idx];
        double len = o->len[idx], nx = o->Nx[idx], ny = o->Ny[idx];
        double Cl[NMAX * NPMAX], Cr[NMAX * NPMAX];
        double UL[NMAX], UR[NMAX], F[NMAX], V[6];
        double epsL = elem_eps(o, left_idx);
        double epsR = right_idx > -1 ? elem_eps(o, right_idx) : epsL;
        elem_V(o, left_idx, V);
        for (n = 0; n < N; n++)
            for (i = 0; i < np; i++) {
                Cl[n * np + i] = C[K * np * n + i * K + left_idx];
                if (right_idx > -1) Cr[n * np + i] = C[K * np * n + i * K + right_idx];
            }
        for (i = 0; i < np; i++)
            for (n = 0; n < N; n++) {
                o->rhs_l[S * np * n + i * S + idx] = 0.;
                o->rhs_r[S * np * n + i * S + idx] = 0.;
            }
        for (j = 0; j < nq1; j++) {
            for (n = 0; n < N; n++) {
                UL[n] = 0.;
                UR[n] = 0.;
            }
            for (i = 0; i < np; i++)
                for (n = 0; n < N; n++)
                    UL[n] += Cl[n * np + i] * o->side[left_side * np * nq1 + i * nq1 + j];
            check_physical(o, C, Cl, UL, left_idx);
            if (right_idx < 0) {
                boundary_state(o, UL, UR, V, nx, ny, j, left_side, t, right_idx);
            } else {
                for (i = 0; i < np; i++)
                    for (n = 0; n < N; n++)
                        UR[n] += Cr[n * np + i] * o->side[right_side * np * nq1 + i * nq1 + nq1 - j - 1];
                check_physical(o, C, Cr, UR, right_idx);
            }
            {
                /* the velocity at the left element's point (advection.cu:103-108) */
                double X[2], A[2];
                coords1d(o, X, V, j, left_side);
                velocity(o, X, A);
                riemann(o, F, UL, UR, nx, ny, epsL, epsR, A);
            }
            for (i = 0; i < np; i++)
                for (n = 0; n < N; n++) {
                    o->rhs_l[S * np * n + i * S + idx] +=
                        -len / 2 * (o->w1[j] * F[n] * o->side[left_side * np * nq1 + i * nq1 + j]);
                    o->rhs_r[S * np * n + i * S + idx] +=
                        len / 2 * (o->w1[j] * F[n] * o->side[right_side * np * nq1 + i * nq1 + nq1 - 1 - j]);
                }
        }
    }
}

/* eval_volume conserv_kernels.cu:992-1083 (source_term is zero for every
   physics here, euler.cu:282-287) */
static void eval_volume(orc *o, double *C, double t) {
    int K = o->K, np = o->n_p, nq = o->nq, N = o->N, idx;
    int par = g_par && (o->posfix || !has_positivity(o));
    (void)t;
#pragma omp parallel for if (par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        double Cl[NMAX * NPMAX], U[NMAX], fx[NMAX], fy[NMAX], V[6], X[2], A[2];
        int i, j, k, n;
        elem_V(o, idx, V);
        for (i = 0; i < np; i++)
            for (n = 0; n < N; n++) Cl[n * np + i] = C[K * np * n + i * K + idx];
        double xr = o->xr[idx], yr = o->yr[idx], xs = o->xs[idx], ys = o->ys[idx];
        for (i = 0; i < np; i++)
            for (n = 0; n < N; n++) o->rhs_v[K * np * n + i * K + idx] = 0.;
        for (j = 0; j < nq; j++) {
            for (n = 0; n < N; n++) U[n] = 0.;
            for (k = 0; k < np; k++)
                for (n = 0; n < N; n++) U[n] += Cl[n * np + k] * o->basis[nq * k + j];
            check_physical(o, C, Cl, U, idx);
            coords2d(o, X, V, j); /* the velocity at the point (advection.cu:57-64) */
            velocity(o, X, A);
            eval_flux(o, U, fx, fy, A);
            for (i = 0; i < np; i++)
                for (n = 0; n < N; n++)
                    o->rhs_v[K * np * n + i * K + idx] +=
                        fx[n] * (o->gx[nq * i + j] * ys - o->gy[nq * i + j] * yr) +
                        fy[n] * (-o->gx[nq * i + j] * xs + o->gy[nq * i + j] * xr);
        }
    }
}

/* eval_rhs time_integrator.cu:26-101: k = dt/J (volume + side terms), each
   side term taken from the left or right array.  TE (extension): the E rows
   are then divided by the element's permittivity and the H row by mu. */
static void eval_rhs_assemble(orc *o, double *c, double dt) {
    int K = o->K, S = o->S, np = o->n_p, N = o->N, idx;
#pragma omp parallel for if (g_par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        int i, n, q;
        int sid[3] = {o->es1[idx], o->es2[idx], o->es3[idx]};
        double j = o->J[idx];
        for (i = 0; i < np; i++)
            for (n = 0; n < N; n++) c[K * np * n + i * K + idx] = 0.;
        for (i = 0; i < np; i++)
            for (n = 0; n < N; n++)
                c[K * np * n + i * K + idx] += dt / j * o->rhs_v[K * np * n + i * K + idx];
        for (q = 0; q < 3; q++) {
            int s = sid[q];
            const double *src = (idx == o->left[s]) ? o->rhs_l : o->rhs_r;
            for (i = 0; i < np; i++)
                for (n = 0; n < N; n++)
                    c[K * np * n + i * K + idx] += dt / j * src[S * np * n + i * S + s];
        }
        if (o->phys == ORC_MAXWELL_TE) {
            double se = 1. / elem_eps(o, idx), sm = 1. / o->mu;
            for (i = 0; i < np; i++)
                for (n = 0; n < N; n++) c[K * np * n + i * K + idx] *= (n < 2) ? se : sm;
        }
    }
}

static void eval_stage(orc *o, double *C, double t, double dt, double *k) {
    positivity_prepass(o, C);
    eval_surface(o, C, t);
    eval_volume(o, C, t);
    eval_rhs_assemble(o, k, dt);
}

int orc_eval_rhs(orc *o, double *C, double t, double dt, double *k) {
    eval_stage(o, C, t, dt, k);
    return 0;
}

/* limit_c conserv_kernels.cu:640-720 (p = 1; min/max roles as in the reference) */
int orc_limit(orc *o, double *C) {
    int K = o->K, np = o->n_p, nq1 = o->nq1, idx;
    if (o->p != 1) return -1;
#pragma omp parallel for if (g_par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        int s1 = o->es1[idx], s2 = o->es2[idx], s3 = o->es3[idx];
        int nb[3], n, i, j, side;
        nb[0] = (o->left[s1] == idx) ? o->right[s1] : o->left[s1];
        nb[1] = (o->left[s2] == idx) ? o->right[s2] : o->left[s2];
        nb[2] = (o->left[s3] == idx) ? o->right[s3] : o->left[s3];
        for (i = 0; i < 3; i++) nb[i] = (nb[i] < 0) ? idx : nb[i];
        for (n = 0; n < o->N; n++) {
            double Uc = C[K * np * n + idx] * o->basis[0], Umin = Uc, Umax = Uc, Ui, alpha;
            double min_alpha = 1.;
            for (i = 0; i < 3; i++) {
                Ui = C[K * np * n + nb[i]] * o->basis[0];
                Umin = (Ui < Umin) ? Ui : Umin;
                Umax = (Ui > Umax) ? Ui : Umax;
            }
            for (j = 0; j < nq1; j++)
                for (side = 0; side < 3; side++) {
                    Ui = 0.;
                    for (i = 0; i < np; i++)
                        Ui += C[K * np * n + i * K + idx] * o->side[side * np * nq1 + i * nq1 + j];
                    if (Ui > Uc)
                        alpha = fmin(1., (Uc - Umin) / (Ui - Uc));
                    else if (Ui < Uc)
                        alpha = fmin(1., (Uc - Umax) / (Ui - Uc));
                    else
                        alpha = 1;
                    if (alpha < min_alpha) min_alpha = alpha;
                }
            if (min_alpha < 0) min_alpha = 0.;
            C[K * np * n + 1 * K + idx] *= min_alpha;
            C[K * np * n + 2 * K + idx] *= min_alpha;
        }
    }
    return 0;
}

/* eval_global_lambda euler.cu:248-277, shallowwater.cu:196-224,
   advection.cu:157-183, maxwells.cu:174-185 */
int orc_lambda(orc *o, double *lam) {
    int K = o->K, np = o->n_p, idx;
#pragma omp parallel for if (g_par) schedule(static)
    for (idx = 0; idx < K; idx++) {
        double U[NMAX];
        int n;
        if (o->phys == ORC_EULER || o->phys == ORC_SWE) {
            for (n = 0; n < o->N; n++) U[n] = o->c[K * np * n + idx] * o->basis[0];
            double c = eval_c(o, U);
            double s;
            if (o->phys == ORC_EULER)
                s = sqrt(U[1] * U[1] + U[2] * U[2]) / U[0];
            else
                s = sqrt(U[1] * U[1] + U[2] * U[2]) / U[0];
            lam[idx] = (s > 0) ? s + c : -s + c;
        } else if (o->phys == ORC_ADVECTION) {
            /* the velocity at quadrature point n_quad / 2 (advection.cu:176-181) */
            double V[6], X[2], A[2];
            elem_V(o, idx, V);
            coords2d(o, X, V, o->nq / 2);
            velocity(o, X, A);
            lam[idx] = sqrt(A[0] * A[0] + A[1] * A[1]);
        } else if (o->phys == ORC_MAXWELL_TE) {
            lam[idx] = fabs(1. / sqrt(elem_eps(o, idx) * o->mu));
        } else {
            lam[idx] = fabs(eval_c(o, U));
        }
    }
    return 0;
}

static void tempstorage(const orc *o, const double *c, double *kstar, const double *k,
                        double alpha) {
    int n = orc_dofs(o), i;
#pragma omp parallel for if (g_par) schedule(static)
    for (i = 0; i < n; i++) kstar[i] = c[i] + alpha * k[i];
}

/* SSP stage (extension): out = a*un + b*us + c*k, terms with a zero
   coefficient skipped, evaluated left to right */
static void ssp_combine(const orc *o, double *out, const double *un, double a, const double *us,
                        double b, const double *k, double cc) {
    int n = orc_dofs(o), i;
#pragma omp parallel for if (g_par) schedule(static)
    for (i = 0; i < n; i++) {
        double v;
        if (a != 0.)
            v = a * un[i] + b * us[i];
        else
            v = b * us[i];
        out[i] = v + cc * k[i];
    }
}

static void maybe_limit(orc *o, double *C) {
    if (o->limiter) orc_limit(o, C);
}

/* time_integrate_rk4 time_integrator.cu:139-479, time_integrate_rk2 :511-717,
   and the SSP-RK2/3 extensions in Shu-Osher form.  dt is recomputed every
   step from cell averages (:200-217) unless fixed_dt > 0; t is advanced
   before the stages (:226-231). */
/* the time loop of time_integrate_rk4 / rk2 (time_integrator.cu:193-433,
 * 550-717): while (t < endtime || steps left || (convergence && conv > tol)),
 * with the step that would pass endtime shortened to land on it (:225-231)
 * and conv = max |c - c_prev| over the first K*N state entries (:403-417).
 * c_prev of the first step is the initial state (the reference compares with
 * a buffer it never initialised). */
int orc_run_until(orc *o, int max_steps, double fixed_dt, double t0, double t_end,
                  double conv_tol, double *t_out, int *steps_out, double *conv_out,
                  double *dt_last) {
    int step = 0, K = o->K, i;
    double t = t0, dt = 0., conv = 1.;
    double *lam = (double *)xalloc(sizeof(double) * K);
    const int nconv = K * o->N;
    double *prev = NULL;
    int rc = 0;
    if (conv_tol >= 0) {
        prev = (double *)xalloc(sizeof(double) * nconv);
        memcpy(prev, o->c, sizeof(double) * nconv);
    }
    maybe_limit(o, o->c);
    for (;;) {
        if (max_steps >= 0 && step >= max_steps) break;
        if (t >= t_end) break;
        if (conv_tol >= 0 && !(conv > conv_tol)) break;
        if (fixed_dt > 0) {
            dt = fixed_dt;
        } else {
            orc_lambda(o, lam);
            double max_l = lam[0];
            for (i = 0; i < K; i++) max_l = (lam[i] > max_l) ? lam[i] : max_l;
            dt = o->cfl * o->min_r / max_l / (2. * o->p + 1.);
        }
        step++;
        if (isnan(dt)) {
            rc = 1;
            break;
        }
        if (t + dt > t_end) {
            dt = t_end - t;
            t = t_end;
        } else {
            t += dt;
        }
        switch (o->integ) {
        case ORC_RK4: {
            int n = orc_dofs(o), q;
            eval_stage(o, o->c, t, dt, o->k1);
            maybe_limit(o, o->k1);
            tempstorage(o, o->c, o->k2, o->k1, 0.5);
            eval_stage(o, o->k2, t + 0.5 * dt, dt, o->k2);
            maybe_limit(o, o->k2);
            tempstorage(o, o->c, o->k3, o->k2, 0.5);
            eval_stage(o, o->k3, t + 0.5 * dt, dt, o->k3);
            maybe_limit(o, o->k3);
            tempstorage(o, o->c, o->k4, o->k3, 1.0);
            eval_stage(o, o->k4, t + dt, dt, o->k4);
            maybe_limit(o, o->k4);
            /* rk4 time_integrator.cu:124-130 */
#pragma omp parallel for if (g_par) schedule(static)
            for (q = 0; q < n; q++)
                o->c[q] += o->k1[q] / 6. + o->k2[q] / 3. + o->k3[q] / 3. + o->k4[q] / 6.;
            maybe_limit(o, o->c);
            break;
        }
        case ORC_RK2MID: {
            int n = orc_dofs(o), q;
            eval_stage(o, o->c, t, dt, o->k1);
            maybe_limit(o, o->k1);
            tempstorage(o, o->c, o->k1, o->k1, 0.5);
            eval_stage(o, o->k1, t + 0.5 * dt, dt, o->k1);
            maybe_limit(o, o->k1);
            /* rk2 time_integrator.cu:503-509 */
#pragma omp parallel for if (g_par) schedule(static)
            for (q = 0; q < n; q++) o->c[q] += o->k1[q];
            maybe_limit(o, o->c);
            break;
        }
        case ORC_SSPRK2:
            eval_stage(o, o->c, t, dt, o->k1);
            maybe_limit(o, o->k1);
            ssp_combine(o, o->k2, NULL, 0., o->c, 1., o->k1, 1.);
            eval_stage(o, o->k2, t + dt, dt, o->k1);
            maybe_limit(o, o->k1);
            ssp_combine(o, o->c, o->c, 0.5, o->k2, 0.5, o->k1, 0.5);
            maybe_limit(o, o->c);
            break;
        case ORC_SSPRK3:
            eval_stage(o, o->c, t, dt, o->k1);
            maybe_limit(o, o->k1);
            ssp_combine(o, o->k2, NULL, 0., o->c, 1., o->k1, 1.);
            eval_stage(o, o->k2, t + dt, dt, o->k1);
            maybe_limit(o, o->k1);
            ssp_combine(o, o->k3, o->c, 0.75, o->k2, 0.25, o->k1, 0.25);
            eval_stage(o, o->k3, t + 0.5 * dt, dt, o->k1);
            maybe_limit(o, o->k1);
            ssp_combine(o, o->c, o->c, 1. / 3., o->k3, 2. / 3., o->k1, 2. / 3.);
            maybe_limit(o, o->c);
            break;
        default:
            rc = -1;
            break;
        }
        if (rc) break;
        if (prev) {
            double m = 0.;
            for (i = 0; i < nconv; i++) {
                double d = fabs(o->c[i] - prev[i]);
                m = d > m ? d : m;
            }
            conv = m;
            memcpy(prev, o->c, sizeof(double) * nconv);
        }
    }
    free(lam);
    free(prev);
    if (t_out) *t_out = t;
    if (steps_out) *steps_out = step;
    if (conv_out) *conv_out = conv;
    if (dt_last) *dt_last = dt;
    return rc;
}

int orc_run(orc *o, int nsteps, double fixed_dt, double *t_out, double *dt_last) {
    return orc_run_until(o, nsteps, fixed_dt, 0., HUGE_VAL, -1., t_out, NULL, NULL, dt_last);
}

int orc_geometry(const orc *o, double *J, double *xr, double *yr, double *xs, double *ys,
                 double *len, double *nx, double *ny, double *diam) {
    size_t K = (size_t)o->K, S = (size_t)o->S;
    if (J) memcpy(J, o->J, K * sizeof(double));
    if (xr) memcpy(xr, o->xr, K * sizeof(double));
    if (yr) memcpy(yr, o->yr, K * sizeof(double));
    if (xs) memcpy(xs, o->xs, K * sizeof(double));
    if (ys) memcpy(ys, o->ys, K * sizeof(double));
    if (len) memcpy(len, o->len, S * sizeof(double));
    if (nx) memcpy(nx, o->Nx, S * sizeof(double));
    if (ny) memcpy(ny, o->Ny, S * sizeof(double));
    if (diam) memcpy(diam, o->diam, K * sizeof(double));
    return 0;
}

int orc_tables(const orc *o, double *basis, double *gx, double *gy, double *side, double *vertex,
               double *w, double *r1, double *r2, double *w1, double *r1d) {
    size_t np = o->n_p, nq = o->nq, nq1 = o->nq1;
    if (basis) memcpy(basis, o->basis, np * nq * sizeof(double));
    if (gx) memcpy(gx, o->gx, np * nq * sizeof(double));
    if (gy) memcpy(gy, o->gy, np * nq * sizeof(double));
    if (side) memcpy(side, o->side, 3 * np * nq1 * sizeof(double));
    if (vertex) memcpy(vertex, o->vertex, 3 * np * sizeof(double));
    if (w) memcpy(w, o->w, nq * sizeof(double));
    if (r1) memcpy(r1, o->r1, nq * sizeof(double));
    if (r2) memcpy(r2, o->r2, nq * sizeof(double));
    if (w1) memcpy(w1, o->w1, nq1 * sizeof(double));
    if (r1d) memcpy(r1d, o->r1d, nq1 * sizeof(double));
    return 0;
}

/* eval_error_L2 conserv_kernels.cu:1432-1479 + host sum main.cu:859-874 */
int orc_l2_error(orc *o, double t, double *err) {
    int K = o->K, np = o->n_p, nq = o->nq, n, idx;
    double *e = (double *)xalloc(sizeof(double) * K);
    for (n = 0; n < o->N; n++) {
#pragma omp parallel for if (g_par) schedule(static)
        for (idx = 0; idx < K; idx++) {
            double V[6], X[2], u[NMAX], U, acc = 0.;
            int i, j;
            double detJ = o->J[idx];
            elem_V(o, idx, V);
            for (j = 0; j < nq; j++) {
                coords2d(o, X, V, j);
                U = 0.;
                for (i = 0; i < np; i++) U += o->c[K * np * n + i * K + idx] * o->basis[i * nq + j];
                exact_state(o, u, X[0], X[1], t);
                acc += o->w[j] * (u[n] - U) * (u[n] - U) * detJ;
            }
            e[idx] = acc;
        }
        double total = 0.;
        for (idx = 0; idx < K; idx++) total += e[idx];
        err[n] = sqrt(total);
    }
    free(e);
    return 0;
}

/* eval_error_L1 conserv_kernels.cu:1378-1430 + host sum main.cu:880-896 */
int orc_l1_error(orc *o, double t, double *err) {
    int K = o->K, np = o->n_p, nq = o->nq, n, idx;
    double *e = (double *)xalloc(sizeof(double) * K);
    for (n = 0; n < o->N; n++) {
#pragma omp parallel for if (g_par) schedule(static)
        for (idx = 0; idx < K; idx++) {
            double V[6], X[2], u[NMAX], U, acc = 0.;
            int i, j;
            double detJ = o->J[idx];
            elem_V(o, idx, V);
            for (j = 0; j < nq; j++) {
                coords2d(o, X, V, j);
                U = 0.;
                for (i = 0; i < np; i++) U += o->c[K * np * n + i * K + idx] * o->basis[i * nq + j];
                exact_state(o, u, X[0], X[1], t);
                acc += o->w[j] * fabs(u[n] - U) * detJ;
            }
            e[idx] = acc;
        }
        double total = 0.;
        for (idx = 0; idx < K; idx++) total += e[idx];
        err[n] = total;
    }
    free(e);
    return 0;
}

/* plot_exact conserv_kernels.cu:1315-1331: variable n of U_exact at the vertices */
int orc_eval_exact(orc *o, int n, double t, double *u1, double *u2, double *u3) {
    int K = o->K, idx;
    for (idx = 0; idx < K; idx++) {
        double V[6], u[NMAX];
        elem_V(o, idx, V);
        exact_state(o, u, V[0], V[1], t);
        u1[idx] = u[n];
        exact_state(o, u, V[2], V[3], t);
        u2[idx] = u[n];
        exact_state(o, u, V[4], V[5], t);
        u3[idx] = u[n];
    }
    return 0;
}

/* one numerical flux evaluation (the physics' riemann_solver) for tests */
int orc_flux_point(orc *o, const double *UL, const double *UR, double nx, double ny,
                   double epsL, double epsR, double *F) {
    double A[2] = {o->ax, o->ay};
    riemann(o, F, UL, UR, nx, ny, epsL, epsR, A);
    return 0;
}

/* physical flux of one state (eval_flux) for tests */
int orc_flux_phys(orc *o, const double *U, double *fx, double *fy) {
    double A[2] = {o->ax, o->ay};
    eval_flux(o, U, fx, fy, A);
    return 0;
}

/* eval_u conserv_kernels.cu:1284-1307 */
int orc_eval_u(orc *o, int n, double *u1, double *u2, double *u3) {
    int K = o->K, np = o->n_p, idx;
    for (idx = 0; idx < K; idx++) {
        double a = 0., b = 0., c = 0.;
        int i;
        for (i = 0; i < np; i++) {
            double v = o->c[K * np * n + i * K + idx];
            a += v * o->vertex[i * 3 + 0];
            b += v * o->vertex[i * 3 + 1];
            c += v * o->vertex[i * 3 + 2];
        }
        u1[idx] = a;
        u2[idx] = b;
        u3[idx] = c;
    }
    return 0;
}
