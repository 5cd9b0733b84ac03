// dg_internal.h -- structures shared by the host API (dg_api.cpp) and the
// HIP kernels (dg_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dgk {

constexpr int PMAX = 5;
constexpr int NMAX = 4;
constexpr int BLK = 256;  // threads per block of the element kernels (4 waves)

__host__ __device__ constexpr int np_of(int p) { return (p + 1) * (p + 2) / 2; }
__host__ __device__ constexpr int nq_of(int p) {
  return p == 0 ? 1 : p == 1 ? 3 : p == 2 ? 6 : p == 3 ? 12 : p == 4 ? 16 : 25;
}
__host__ __device__ constexpr int nvars_of(int phys) { return phys == 0 ? 1 : phys == 1 ? 4 : 3; }

// reference-element operators of order P (the reference's __constant__ tables,
// conserv_kernels.cu:49-76, built like preval_basis basis.cu:168-244)
template <int P>
struct Tab {
  static constexpr int NP = np_of(P), NQ = nq_of(P), NQ1 = P + 1;
  double B[NP * NQ];         // phi_i(r_j)              [i][j]
  double GX[NP * NQ];        // w_j d phi_i / dr (r_j)  [i][j]
  double GY[NP * NQ];        // w_j d phi_i / ds (r_j)  [i][j]
  double SIDE[3 * NP * NQ1];  // phi_i on side k, point j [k][i][j]
  double VERT[3 * NP];       // phi_i at the vertices    [i][v]
  double W[NQ], R1[NQ], R2[NQ];
  double W1[NQ1], X1[NQ1];   // 1D Gauss weights and points on [-1, 1]
};

struct PhysParams {
  double gamma, G, eps, mu, ax, ay;
  int prob, flux;
  double pp[8];
};

// device-resident step bookkeeping (one per handle)
struct DevState {
  double t;       // stage base time of the running step (advanced by dt, as time_integrator.cu:226-231)
  double dt;      // dt of the running step
  double tdone;   // time reached by the last completed step
  double dt_last; // dt of the last completed step
  int nan_flag;                 // set when the NaN-dt guard fires
  double lam_max;               // last reduced max speed
  double t_end;                 // stop time (+inf: none); the step that would pass it lands on it
  double conv_tol;              // stop when the step-to-step change is <= conv_tol (< 0: off)
  double conv;                  // last reduced step-to-step change
  int halt;                     // 1: t_end reached, 2: converged; every stage launch then exits
  int steps;                    // steps completed since the state was written
};

enum StageMode { M_KONLY = 0, M_US_CK = 1, M_A_UN_B_US_CK = 2, M_UN_CK = 3 };
enum AccMode { ACC_NONE = 0, ACC_SET6 = 1, ACC_ADD3 = 2, ACC_FINAL6 = 3 };

struct MeshDev {
  int K;
  const double *xr, *yr, *xs, *ys, *J;  // [K]
  const double *V;                       // [6][K]
  const int2 *slot;                      // [3][K]: neighbour (or -BC code), packed info
  const double *snx, *sny, *slen;        // [S]: normal (outward from the side's left element), length
  const double *eps;                     // [K] (Maxwell TE) or nullptr
  // boundary slots (right < 0)
  int nb_bnd;                            // number of boundary slots
  const int *bnd_elem;                   // [nb_bnd] element of the slot
  const int *bnd_code;                   // [nb_bnd] local side | (BC code << 2), code 1..3
  int S;
  const int4 *side;                      // [S]: left, right (or -BC code), lsn | rsn << 2,
                                         // boundary slot index (-1: interior)
  double *bnd_val;                       // [nb_bnd][NQ1][NMAX]: Dirichlet state or wall direction
};
// slot info bits: 0-1 neighbour's local side, 2 role (1: this element is the side's right),
// 3.. the side's index

struct StageArgs {
  MeshDev m;
  const double *us;  // stage input  [N][NP][K]
  const double *un;  // step base    [N][NP][K]
  double *out;       // stage output
  double *acc;       // RK4 running combination
  DevState *st;
  PhysParams ph;
  int mode, accmode;
  double A, B, C;
  double cstage;     // stage time = st->t + cstage * st->dt
  int last_stage;    // 1: finish the step (bookkeeping, CFL reduction)
  int cfl_mode;      // 1: next dt from the CFL rule; 0: keep dt
  // element-local positivity fix (check_physical, race-free): on the stage
  // input and the gathered neighbours (fix_in), on the step base read by the
  // combination (fix_un), on the written intermediate stage state (fix_out)
  int fix_in, fix_un, fix_out;
  // per-element flags of the stored step state (1: it fails the point check),
  // always a valid [K] array; flags_in: the stage's input and step base carry
  // valid flags, read instead of testing; flags_out: the last stage writes
  // them for the state it stores
  int *bad;
  int flags_in, flags_out;
  double cfl_num;    // cfl * min_r (dt = cfl_num / lam / (2p + 1))
  int conv;          // 1 (last stage): reduce the step-to-step change (rows n * n_p + i < N)
  unsigned long long *part;  // [2 * nblk] per-block maxima of the last stage (step reduction)
  int nblk;          // blocks in the grid (XCD remap)
  double *kout;      // Maxwell TE: the E rows of k = dt L(us) are also stored here (or nullptr)
  // ensemble: per logical block {instance, its first element, its end, 0}
  // (nullptr: one mesh, instance 0, elements [0, K))
  const int4 *ens;
  int vol_mfma;      // 1: the volume integral on the matrix cores (Euler p = 4)
  // stored traces (four-lane kernel, Euler p >= 4): every element's traces on
  // its three sides at the n_q1 side points (layout in dg_stage_q.h); tr_in:
  // those of us (nullptr: the stage gathers the neighbours' modes instead),
  // tr_out: written for the stage's output
  const double *tr_in;
  double *tr_out;
  int *fired;  // set to 1 wherever the positivity fix fires or a stored state is flagged
};

// doubles of stored traces per element: 3 sides x n_q1 points x 4 variables
__host__ __device__ constexpr int tr_elem(int p) { return 12 * (p + 1); }

// parameters of the step-closing reduction (k_finish)
struct FinishArgs {
  int cfl_mode;     // 1: dt from the CFL rule
  double cfl_num;   // cfl * min_r
  int p;
  double fixed_dt;  // dt of the first step when cfl_mode == 0
  int initial;      // 1: before the first step (no step completed)
  double t0;        // time of the initial state
  int conv;         // 1: the partials hold the step-to-step change too
};

// ensemble (dg_ensemble.hip): group p of the merged mesh runs blocks
// [gb[p], gb[p+1]) of one stage launch with stage arguments a[p]
// order groups of an ensemble: p = 0..5 (Euler instances of p >= 4 run the
// four-lane kernel's ensemble form, in split launches only)
constexpr int kEnsGroups = 6;
struct EnsStageArgs {
  StageArgs a[kEnsGroups];
  int gb[kEnsGroups + 1];
};
// one instance: its group (= its order p), element range [e0, e1) in the
// group's arrays, and the CFL constant cfl * min_r of its mesh
struct EnsInst {
  int g, e0, e1, p;
  double cfl_num;
};
struct EnsParts {
  unsigned long long *part[kEnsGroups];  // per group: the step-reduction partials
  int bs[kEnsGroups];                    // per group: elements per partial (its block size)
};
void launch_stage_ens(int phys, int B, const EnsStageArgs &ea, int nblocks_total, hipStream_t s);
void launch_stage_ens_split(int phys, const int *bs, const EnsStageArgs &ea, hipStream_t s);
void launch_lambda_ens(int phys, int p, const MeshDev &m, const PhysParams &ph, const double *c,
                       const int4 *blk256, unsigned long long *part, int *bad, hipStream_t s);
void launch_bnd_ens(int phys, int p, const MeshDev &m, const PhysParams *ph, const int *inst,
                    const DevState *st, double cstage, hipStream_t s);
void launch_finish_ens(DevState *st, const EnsInst *inst, int ninst, const EnsParts &parts,
                       const FinishArgs &f, hipStream_t s);
int ens_vgprs(int phys, int B);

// fixed scratch of dg_eval_rhs_noalloc (dg_kernels.hip): the DevState at
// byte 0, the positivity flags ([K] ints, K <= kScratchMaxK) at byte 256
constexpr int kScratchMaxK = 1 << 22;
int scratch_address(void **p);
void launch_set_state(DevState *st, double t, double dt, hipStream_t s);

// launchers (dg_kernels.hip)
int upload_tables(int p, const void *tab, size_t bytes);
void launch_stage(int phys, int p, const StageArgs &a, hipStream_t s);
void launch_init(int phys, int p, const MeshDev &m, const PhysParams &ph, double *c, hipStream_t s);
// per-element error integrals (norm 1 or 2), [N][K]
void launch_err(int phys, int p, int norm, const MeshDev &m, const PhysParams &ph,
                const double *c, double t, double *err_elem, hipStream_t s);
void launch_exact_vertex(int phys, const MeshDev &m, const PhysParams &ph, double t, int n,
                         double *u1, double *u2, double *u3, hipStream_t s);
// CFL speed of every element into lam_out (if not null); with st, closes a
// step (initial: opens the first) through part and k_finish
void launch_lambda_init(int phys, int p, const MeshDev &m, const PhysParams &ph, const double *c,
                        DevState *st, int cfl_mode, double cfl_num, double fixed_dt, double t0,
                        int initial, double *lam_out, hipStream_t s, const double *cprev,
                        int conv_rows, unsigned long long *part, int *bad_out = nullptr);
void launch_finish(DevState *st, const unsigned long long *part, int nblk, const FinishArgs &f,
                   hipStream_t s);
void launch_copy(long long n, double *dst, const double *src, hipStream_t s);
void launch_posfix(int phys, int p, int K, const PhysParams &ph, double *c, hipStream_t s);
void launch_flags(int phys, int p, int K, const PhysParams &ph, const double *c, int *bad,
                  hipStream_t s, int *fired = nullptr);
void launch_limit(int phys, int p, const MeshDev &m, double *c, hipStream_t s);
void launch_axpy(long long n, double *out, const double *c, double alpha, const double *k,
                 hipStream_t s);
void launch_ssp(long long n, double *out, const double *un, double a, const double *us, double b,
                const double *k, double cc, hipStream_t s);
void launch_rk4_final(long long n, double *c, const double *k1, const double *k2,
                      const double *k3, const double *k4, hipStream_t s);
void launch_eval_u(int phys, int p, int K, const double *c, int n, double *u1, double *u2,
                   double *u3, hipStream_t s);
// boundary data of every boundary slot at the stage time (or a fixed time)
void launch_bnd(int phys, int p, const MeshDev &m, const PhysParams &ph, const DevState *st,
                double cstage, int use_state, double t, hipStream_t s);
int stage_vgprs(int phys, int p, bool vol_mfma);
void launch_traces(int phys, int p, int K, const double *u, double *tr, hipStream_t s);

// adjoint launchers (dg_adjoint.hip)
struct AdjArgs {
  MeshDev m;
  const double *us;   // forward stage input
  const double *lam;  // adjoint of k is lc * lam, [N][NP][K]
  double lc;
  double *gout;       // output [N][NP][K]: scale * (dk/du)^T lam + sum_t oc[t] ot[t], t < nout
                      // (nullptr: not written)
  const double *ot[4];
  double oc[4];
  int nout;
  // optional second output (own-element terms only):
  // gout2 = c2g * (dk/du)^T lam + sum_t o2c[t] o2t[t], t < n2
  double *gout2;
  double c2g;
  const double *o2t[4];
  double o2c[4];
  int n2;
  // Maxwell TE eps gradient: geps [K] accumulates it (nullptr: not formed),
  // from the forward stage's k = dt L(us) (kst) and the per-side impedance
  // terms gzs [2][S] (left, right element) that k_eps_side writes
  double *geps;
  const double *kst;
  double *gzs;
  PhysParams ph;
  double t, dt;
  double scale;
};
void launch_adjoint_rhs(int phys, int p, const AdjArgs &a, hipStream_t s);
void launch_objective(int phys, int p, int kind, const MeshDev &m, const PhysParams &ph,
                      const double *c, const double *prm, double *obj_elem, double *grad,
                      hipStream_t s);

}  // namespace dgk
