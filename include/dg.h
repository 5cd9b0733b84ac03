/* dg.h -- C ABI of the MI355X DG right-hand-side library (libdgamd.so).
 *
 * Drop-in boundary for DGCUDA's explicit DG path (reference paths relative
 * to src/2d/conservation/).  The reference binds this path in three layers,
 * all replaced here by explicit handles (no globals, no exit()):
 *   - program API  run_dgcuda(argc, argv) + read_mesh  (main.cu:205-277, 600-974)
 *   - plugin API   eval_flux / riemann_solver / U0 / U_inflow / U_outflow /
 *                  U_reflection / U_exact / check_physical
 *                  (conserv_headers.cuh:5-34), chosen here by ids
 *   - kernel API   eval_surface, eval_volume, eval_rhs, rk_tempstorage, rk4,
 *                  limit_c, eval_global_lambda (conserv_kernels.cu:640-1083,
 *                  time_integrator.cu:26-130, euler.cu:248-277)
 *
 * Coefficients cross this boundary in the reference's SoA layout
 * C[n*K*n_p + i*K + e] (conserv_kernels.cu:224,1013), as host or device
 * pointers (DG_HOST / DG_DEVICE).  Every call returns DG_OK or a negative
 * error code; dg_last_error() gives the message.  One handle = one HIP
 * stream; calls on one handle must be serialised by the caller.
 */
#ifndef DG_H
#define DG_H

#ifdef __cplusplus
extern "C" {
#endif

#define DG_OK 0
#define DG_ERR_ARG -1     /* bad argument */
#define DG_ERR_STATE -2   /* call out of order (no mesh, ...) */
#define DG_ERR_HIP -3     /* HIP runtime error */
#define DG_ERR_IO -4      /* file error (.pmsh) */
#define DG_ERR_NAN -5     /* the NaN-dt guard stopped the run (time_integrator.cu:219-223) */

/* physics (euler/euler.cu, advection/advection.cu, shallowwater/shallowwater.cu,
   maxwells/maxwells.cu; MAXWELL_TE is new: (Ex, Ey, Hz) with per-element eps) */
#define DG_ADVECTION 0
#define DG_EULER 1
#define DG_SWE 2
#define DG_MAXWELL_TM 3
#define DG_MAXWELL_TE 4

/* numerical flux: DEFAULT = the reference's riemann_solver for the physics
   (upwind for advection, LLF otherwise; upwind for MAXWELL_TE) */
#define DG_FLUX_DEFAULT 0
#define DG_FLUX_LLF 1
#define DG_FLUX_UPWIND 2
#define DG_FLUX_ROE 3 /* shallow water only */

/* time integrators: RK4 and RK2MID are the reference's (time_integrator.cu:139,511);
   SSPRK2/SSPRK3 are Shu-Osher extensions */
#define DG_RK4 1
#define DG_RK2MID 2
#define DG_SSPRK2 3
#define DG_SSPRK3 4

/* built-in problems: initial, inflow/outflow and exact states, wall rule */
#define DG_PROB_ADV_SINSIN 0       /* sin(2 pi x) sin(2 pi y), a = (ax, ay) */
#define DG_PROB_ADV_LINEAR 1       /* advection/exact/exact.cu: u = x - y */
#define DG_PROB_EULER_VORTEX 2     /* isentropic vortex: beta, x0, y0, u, v, period */
#define DG_PROB_EULER_SUPERSONIC 3 /* euler/supersonic/supersonic.cu */
#define DG_PROB_EULER_CONST 4      /* free stream: rho, u, v, p */
#define DG_PROB_SWE_BUMP 5         /* h0 + A exp(-|x-c|^2/(k s^2)): h0, A, cx, cy, s, k */
#define DG_PROB_MAXWELL_PULSE 6    /* A exp(-|x-c|^2/s^2) in the last variable: A, cx, cy, s */
#define DG_PROB_CONST 7            /* constant state params[0..N-1] */
#define DG_PROB_ADV_ROTHILL 8      /* advection/rotatinghill: hill 5 exp(-|x-(0.2,0)|^2/(2 0.15^2))
                                      in the rotation a = 2 pi (-y, x); rotated-hill inflow,
                                      zero outflow; exact state = initial (period 1) */

/* memory kinds for pointer arguments */
#define DG_HOST 0
#define DG_DEVICE 1

typedef struct dg_config {
    int physics;
    int p;          /* polynomial order 0..5 (main.cu:148-167) */
    int flux;
    int integrator;
    int limiter;    /* 1: limit_c after every stage and step (p = 1 only) */
    int positivity; /* 1: element-local form of check_physical (Euler, SWE) */
    double cfl;     /* 0: reference default, 0.5 for RK4 and 0.7 for RK2MID */
    double gamma, g, eps, mu, ax, ay;
    int problem;
    double problem_params[8];
    int device;     /* HIP device ordinal */
    int volume;     /* volume integral of the stage: 0 the library's choice, 1 vector FP64
                       FMAs, 2 FP64 matrix cores (v_mfma_f64_16x16x4f64; Euler p = 4 only) */
} dg_config;

typedef struct dg_info_t {
    int K, S, N, n_p, n_q, n_q1;
    long long dofs;       /* K * n_p * N */
    double min_r;         /* smallest inscribed diameter (main.cu:722-736) */
    int vgprs_stage;      /* registers of the stage kernel in use (diagnostic) */
    int adj_fix_fired;    /* 1: the positivity fix fired somewhere in the last
                             dg_adjoint_solve's forward map, which the gradient
                             takes as inactive (it is then not exact there) */
    int adj_resident;     /* segments of the last dg_adjoint_solve whose stage data
                             stayed in device memory from the forward pass (the
                             reverse sweep recomputed only the others) */
} dg_info_t;

typedef struct dg_handle dg_handle;

/* fills cfg with the reference defaults (Euler, p = 1, RK4, gamma 1.4) */
void dg_config_default(dg_config *cfg);

/* replaces init_gpu + set_quadrature + preval_basis (main.cu:279-408, 124-194; basis.cu:168) */
int dg_create(dg_handle **h, const dg_config *cfg);
void dg_destroy(dg_handle *h);
const char *dg_last_error(const dg_handle *h);

/* mesh as read_mesh produces it (main.cu:205-277): V is [6][K]
   (V1x, V1y, V2x, V2y, V3x, V3y), elem_sides [3][K], side_xy [4][S]
   (x1, y1, x2, y2); right < 0 marks a boundary (-1 wall, -2 outflow, -3 inflow).
   Runs the geometry precompute of main.cu:698-769 (preval_* with the
   reference's FP32 sqrtf) on the host, then copies the mesh to the device. */
int dg_set_mesh(dg_handle *h, int K, const double *V, const int *elem_sides, int S,
                const double *side_xy, const int *left, const int *right, const int *lsn,
                const int *rsn);
/* read_mesh on a .pmsh file (main.cu:205-277) followed by dg_set_mesh */
int dg_load_pmsh(dg_handle *h, const char *path);
/* per-element relative permittivity for DG_MAXWELL_TE ([K], host) */
int dg_set_eps(dg_handle *h, const double *eps_elem);

/* init_conditions (conserv_kernels.cu:201-228) with the problem's U0 */
int dg_project_ic(dg_handle *h);
int dg_set_coeffs(dg_handle *h, const double *C, int where);
int dg_get_coeffs(dg_handle *h, double *C, int where);

/* k = dt * L(u): eval_surface + eval_volume + eval_rhs (conserv_kernels.cu:899-1083,
   time_integrator.cu:26-101) at time t.  u and k are [N][n_p][K]. */
int dg_eval_rhs(dg_handle *h, const double *u, double t, double dt, double *k, int where);

/* ---- user boundary states: the reference's U_inflow / U_outflow /
   U_reflection plugin hooks (conserv_headers.cuh:5-34, compiled into each
   problem) as data.  The mesh's boundary sides (right < 0) are numbered
   0..nb-1 in the order of their elements (element e's local sides 0, 1, 2
   before element e + 1's); side b's n_q1 = p + 1 quadrature points are
   numbered in the direction the element walks the side.
   dg_boundary_points: *nb always; then, for each array that is not NULL,
   the side index [nb], the code (-1 wall, -2 outflow, -3 inflow) [nb] and
   the physical points x, y [nb][n_q1] (host). */
int dg_boundary_points(dg_handle *h, int *nb, int *side, int *code, double *x, double *y);
/* U [nb][n_q1][N] (host or device): at every outflow and inflow point the
   state outside the domain (eval_boundary's U_outflow / U_inflow, conserv_
   kernels.cu:820-843); at every wall point U_reflection's data: in entries 0
   and 1 the unit wall direction (the side's normal by default), and for
   Euler in entry 2 the rule: 0 mirrors the momentum about the wall, 1 keeps
   only its tangential part (euler/supersonic, supersonic.cu:65-87).
   Time-independent: used at every stage until the mesh changes or
   U == NULL restores the problem's boundary data. */
int dg_set_boundary_states(dg_handle *h, const double *U, int where);
/* the boundary data in use, in dg_set_boundary_states' layout ([nb][n_q1][N],
   host): the caller's, or the problem's at time t */
int dg_get_boundary_states(dg_handle *h, double t, double *U);

/* dg_eval_rhs without allocation.  No counterpart in the reference: this
   entry point is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   u and k are DEVICE pointers ([N][n_p][K], not overlapping); the input is
   read in place, and the stage state and the positivity flags live in a
   fixed scratch array that each GPU holds for the life of the process,
   sized for K <= 4,194,304 elements (larger meshes: DG_ERR_ARG).  No
   hipMalloc, hipFree, memset or copy is issued; the call only enqueues
   kernels on the handle's stream and returns (dg_sync or a stream wait
   orders what follows).  Calls that overlap on one GPU share the scratch
   and race: serialise them (one stream, or events). */
int dg_eval_rhs_noalloc(dg_handle *h, const double *u, double t, double dt, double *k);

/* nsteps steps of time_integrate_rk4/rk2 (time_integrator.cu:139-717): dt from the
   CFL rule every step (fixed_dt <= 0) or fixed.  *t_inout: time in/out (reference
   convention: t advances by dt before the stages).  dt_out: the last step's dt. */
int dg_step(dg_handle *h, int nsteps, double fixed_dt, double *t_inout, double *dt_out);

/* the reference's whole time loop (time_integrate_rk4/rk2 with -t / -T / -c,
   time_integrator.cu:193-433): steps until max_steps are done (< 0: no limit),
   until t reaches t_end (HUGE_VAL: none; the step that would pass t_end is
   shortened to land on it, time_integrator.cu:225-228), or until the
   step-to-step change max |c_new - c_old| over the first K*N entries of the
   state drops to conv_tol or below (conv_tol < 0: off; time_integrator.cu:403-417).
   At least one stop condition is required.  *steps_done and *conv_out (the
   last change) may be NULL.  The stop tests run on the device. */
int dg_run(dg_handle *h, int max_steps, double t_end, double conv_tol, double fixed_dt,
           double *t_inout, int *steps_done, double *conv_out);

/* per-variable L2 error against the problem's exact state (eval_error_L2,
   conserv_kernels.cu:1432-1479; main.cu:859-874).  err is [N]. */
int dg_l2_error(dg_handle *h, double t, double *err);
/* per-variable L1 error (eval_error_L1, conserv_kernels.cu:1378-1430; host sum
   main.cu:880-896).  err is [N]. */
int dg_l1_error(dg_handle *h, double t, double *err);
/* variable n of the problem's exact state at the three vertices of each
   element (plot_exact, conserv_kernels.cu:1315-1331); with dg_eval_u this
   gives plot_error (conserv_kernels.cu:1341-1372, which reads its basis at a
   wrong stride; the difference taken here uses the vertex values).  Each
   output is [K] on the host. */
int dg_eval_exact(dg_handle *h, int n, double t, double *u1, double *u2, double *u3);
/* values at the three vertices of each element, variable n (eval_u,
   conserv_kernels.cu:1284-1307); each output is [K] on the host */
int dg_eval_u(dg_handle *h, int n, double *u1, double *u2, double *u3);
/* per-element CFL speed of the current coefficients (eval_global_lambda) */
int dg_lambda(dg_handle *h, double *lambda);
/* p = 1 limiter in place on the current coefficients (limit_c) */
int dg_limit(dg_handle *h);

int dg_info(dg_handle *h, dg_info_t *info);
/* geometry outputs (each host [K] or [S], any may be NULL) */
int dg_geometry(dg_handle *h, double *J, double *len, double *nx, double *ny, double *diam);
int dg_sync(dg_handle *h);
/* the HIP stream the handle launches on (hipStream_t) */
void *dg_stream(dg_handle *h);
/* average duration in ms of the last dg_step's stage launches, measured with
   HIP events on the handle's stream; also the number of stage launches timed */
int dg_last_timing(dg_handle *h, double *ms_per_stage, int *stages);

/* ---- discrete adjoint (no counterpart in the reference: it differentiates
   the map dg_step computes).  The forward solve uses a fixed dt so the discrete
   map is differentiable; the positivity fix is taken as inactive.  Checkpoints
   of the step state are kept every `every` steps (every <= 0: sqrt(nsteps));
   the reverse sweep re-runs each segment from its checkpoint.  Not available
   with the limiter. */
#define DG_OBJ_LINEAR 0   /* <g, u_T> with weights g [N][n_p][K] */
#define DG_OBJ_ENTROPY 1  /* integral of -rho log(p / rho^gamma) over the domain (Euler) */
#define DG_OBJ_PATCH_SQ 2 /* integral of u_var^2 over elements with centroid in
                             [x0, x1] x [y0, y1]; params x0, x1, y0, y1, var */
int dg_adjoint_set_objective(dg_handle *h, int kind, const double *params, int nparams,
                             const double *g_weights, int where);
/* nsteps forward steps of size dt from the current coefficients, then the
   reverse sweep.  Outputs J, dJ/du0 [N][n_p][K] and, for DG_MAXWELL_TE,
   dJ/deps [K] (either gradient pointer may be NULL).  The coefficients are left
   at the final state. */
int dg_adjoint_solve(dg_handle *h, int nsteps, double dt, int every, double *J,
                     double *grad_u0, double *grad_eps, int where);
/* one vector-Jacobian product of the right-hand side: gu = (dk/du)^T gk for
   k = dt * L(u) at time t (u, gk, gu [N][n_p][K]) */
int dg_adjoint_rhs(dg_handle *h, const double *u, double t, double dt, const double *gk,
                   double *gu, int where);
/* wall-clock ms of the last dg_adjoint_solve: forward solve (with checkpoints)
   and reverse sweep (recompute + adjoint stages) */
int dg_adjoint_timing(dg_handle *h, double *fwd_ms, double *rev_ms);
/* the objective set by dg_adjoint_set_objective at the current coefficients
   (what dg_adjoint_solve reports as J after its forward solve) */
int dg_adjoint_objective(dg_handle *h, double *J);

/* ---- ensembles (no counterpart in the reference, which runs one mesh per
   process): many independent meshes of one physics, integrator and flux,
   each with its own order p (0..5), problem parameters and CFL
   dt, stepped together.  A stage of the whole ensemble is ONE kernel launch
   (the instances are merged by order, every block inside one instance; with
   instances of order 4 or 5, a launch per order group) and one
   step-closing launch sets each instance's next dt; a step is replayed as a
   captured graph.  Each instance evolves exactly as it would alone on a
   handle (same arithmetic per element).  Boundary data that depend on time
   are evaluated per instance at its own stage times, with its own problem
   parameters.  The limiter is not available. */
typedef struct dg_ensemble dg_ensemble;
typedef struct dg_ens_info_t {
    int ninst;
    long long dofs, elements;
    int block;            /* elements per block of the stage launch */
    double tune_ms[3];    /* ms per step at blocks of 64, 128, 256 (-1: not timed) */
    int vgprs;            /* registers of the stage kernel in use */
    int launches_per_step;
    double ms_per_step;   /* the last dg_ens_step, HIP events */
    double tune_split_ms[3]; /* the same with a stage launch per order group (-1: not timed) */
    int split;            /* 1: a stage launch per order group (each at its order's registers) */
    int block_per_order[6]; /* split: each order group's elements per block, tuned one group
                               at a time from the best uniform candidate (0: no instance) */
    double tune_best_ms;  /* the chosen layout's ms per step in the tuning (-1: not tuned) */
} dg_ens_info_t;

/* dg_ens_build's block | DG_ENS_SPLIT: a stage launch per order group */
#define DG_ENS_SPLIT 0x10000

/* cfg: shared physics, flux, integrator, cfl, problem (cfg.p is ignored) */
int dg_ens_create(dg_ensemble **e, const dg_config *cfg);
void dg_ens_destroy(dg_ensemble *e);
const char *dg_ens_last_error(const dg_ensemble *e);
/* one instance: order p, its problem_params (NULL: cfg's) and its mesh in
   the dg_set_mesh form; the initial condition is projected.  *index: its
   number (instances are numbered in the order they are added) */
int dg_ens_add(dg_ensemble *e, int p, const double *problem_params, int K, const double *V,
               const int *elem_sides, int S, const double *side_xy, const int *left,
               const int *right, const int *lsn, const int *rsn, int *index);
/* merges the instances; block 64 / 128 / 256 elements per block (one stage
   launch for every order group), the same | DG_ENS_SPLIT (a stage launch per
   order group, each at its order's register budget), or <= 0 to time all six
   candidates on two steps and keep the fastest (the state is restored); when
   the split form wins, each order group's block size is then tuned in turn
   (dg_ens_info's block_per_order).  Instances of order 4 or 5 run only as
   64 | DG_ENS_SPLIT (their groups stay at 64; other blocks get DG_ERR_ARG). */
int dg_ens_build(dg_ensemble *e, int block);
/* nsteps steps of every instance: dt from each instance's CFL rule
   (fixed_dt <= 0) or fixed_dt for all */
int dg_ens_step(dg_ensemble *e, int nsteps, double fixed_dt);
int dg_ens_get_coeffs(dg_ensemble *e, int inst, double *C, int where);
int dg_ens_set_coeffs(dg_ensemble *e, int inst, const double *C, int where);
/* time reached, last dt and steps taken by one instance */
int dg_ens_state(dg_ensemble *e, int inst, double *t, double *dt, long long *steps);
int dg_ens_info(dg_ensemble *e, dg_ens_info_t *info);
/* dg_l2_error of one instance at time t */
int dg_ens_l2_error(dg_ensemble *e, int inst, double t, double *err);

#ifdef __cplusplus
}
#endif
#endif
