/* dg_oracle.h -- CPU restatement of DGCUDA's DG right-hand side (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  It is never part of the product path.
 *
 * Every function follows the reference's GPU tree (src/2d/conservation/) in
 * loop order, summation order and data layout (SoA: C[n*K*n_p + i*K + e]);
 * the file:line each one restates is cited next to its definition in
 * dg_oracle.c.  Extensions the reference lacks (SSP-RK2/3, periodic meshes,
 * the SWE Roe flux, Maxwell TE with per-element permittivity and its upwind
 * flux) are marked as such and pinned by the tests instead.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* physics */
#define ORC_ADVECTION 0
#define ORC_EULER 1
#define ORC_SWE 2
#define ORC_MAXWELL_TM 3
#define ORC_MAXWELL_TE 4

/* numerical flux */
#define ORC_FLUX_DEFAULT 0 /* the reference's choice for the physics */
#define ORC_FLUX_LLF 1
#define ORC_FLUX_UPWIND 2
#define ORC_FLUX_ROE 3

/* time integrators */
#define ORC_RK4 1
#define ORC_RK2MID 2
#define ORC_SSPRK2 3
#define ORC_SSPRK3 4

/* problems (initial, boundary and exact states) */
#define ORC_PROB_ADV_SINSIN 0     /* u0 = sin(2 pi x) sin(2 pi y), a = (ax, ay) */
#define ORC_PROB_ADV_LINEAR 1     /* advection/exact: u0 = x - y, a = (1, 1) */
#define ORC_PROB_EULER_VORTEX 2   /* Shu isentropic vortex on a periodic box */
#define ORC_PROB_EULER_SUPERSONIC 3 /* euler/supersonic steady vortex */
#define ORC_PROB_EULER_CONST 4    /* free stream */
#define ORC_PROB_SWE_BUMP 5       /* h = h0 + A exp(-|x-c|^2/s^2), zero velocity */
#define ORC_PROB_MAXWELL_PULSE 6  /* Gaussian pulse in the last variable */
#define ORC_PROB_CONST 7          /* any physics: constant state params[0..N-1] */
#define ORC_PROB_ADV_ROTHILL 8    /* advection/rotatinghill: a = 2 pi (-y, x), rotated-hill inflow */

typedef struct orc orc;

orc *orc_new(int phys, int p, int flux, int integrator, double cfl, int limiter,
             int posfix);
void orc_free(orc *o);
void orc_set_threads(int n);

/* params: gamma, G, eps, mu, ax, ay (unused entries ignored) */
int orc_set_params(orc *o, double gamma, double G, double eps, double mu,
                   double ax, double ay);
/* problem id and up to 8 parameters (see dg_oracle.c) */
int orc_set_problem(orc *o, int prob, const double *pp, int npp);
/* mesh in the reference's .pmsh arrays (main.cu:205-277) */
int orc_set_mesh(orc *o, int K, const double *V, const int *elem_s, int S,
                 const double *side_xy, const int *left, const int *right,
                 const int *lsn, const int *rsn);
/* per-element relative permittivity (Maxwell TE only) */
int orc_set_eps(orc *o, const double *eps_elem);

int orc_project_ic(orc *o);
int orc_set_coeffs(orc *o, const double *C);
int orc_get_coeffs(const orc *o, double *C);
int orc_dofs(const orc *o);
int orc_nvars(const orc *o);
int orc_np(const orc *o);

/* k = dt * L(C) as eval_surface + eval_volume + eval_rhs do it; C is not
   modified unless a positivity fix triggers (as in the reference). */
int orc_eval_rhs(orc *o, double *C, double t, double dt, double *k);

/* nsteps time steps; fixed_dt > 0 skips the CFL rule.  Returns 0, or 1 when
   the NaN-dt guard stopped the run early. */
int orc_run(orc *o, int nsteps, double fixed_dt, double *t_out, double *dt_last);
/* the reference's full loop: max_steps (< 0: none), t_end (HUGE_VAL: none),
   conv_tol (< 0: off); starts at t0 */
int orc_run_until(orc *o, int max_steps, double fixed_dt, double t0, double t_end,
                  double conv_tol, double *t_out, int *steps_out, double *conv_out,
                  double *dt_last);

double orc_min_r(const orc *o);
/* geometry outputs, each sized K or S */
int orc_geometry(const orc *o, double *J, double *xr, double *yr, double *xs,
                 double *ys, double *len, double *nx, double *ny, double *diam);
/* reference-element tables: basis[n_p*n_q], gx[n_p*n_q], gy[n_p*n_q],
   side[3*n_p*n_q1], vertex[3*n_p], w[n_q], r1[n_q], r2[n_q], w1[n_q1], r1d[n_q1] */
int orc_tables(const orc *o, double *basis, double *gx, double *gy, double *side,
               double *vertex, double *w, double *r1, double *r2, double *w1,
               double *r1d);
int orc_nq(const orc *o);
int orc_nq1(const orc *o);

/* per-variable L2 error against the problem's exact state at time t */
int orc_l2_error(orc *o, double t, double *err);
int orc_l1_error(orc *o, double t, double *err);
int orc_eval_exact(orc *o, int n, double t, double *u1, double *u2, double *u3);
/* per-element CFL speed (eval_global_lambda) of the current coefficients */
int orc_lambda(orc *o, double *lambda);
/* p=1 limiter applied in place to C (limit_c) */
int orc_limit(orc *o, double *C);
/* one numerical flux F(UL, UR).n and one physical flux (for tests) */
int orc_flux_point(orc *o, const double *UL, const double *UR, double nx, double ny,
                   double epsL, double epsR, double *F);
int orc_flux_phys(orc *o, const double *U, double *fx, double *fy);
/* values at the three vertices of each element for variable n (eval_u) */
int orc_eval_u(orc *o, int n, double *u1, double *u2, double *u3);

#ifdef __cplusplus
}
#endif
#endif
