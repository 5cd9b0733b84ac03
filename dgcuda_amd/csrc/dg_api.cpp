// dg_api.cpp -- host side of libdgamd.so: the C ABI declared in include/dg.h.
//
// Replaces the reference's driver layer (paths relative to src/2d/conservation/):
//   read_mesh         main.cu:205-277      -> dg_load_pmsh / dg_set_mesh
//   init_gpu, free_gpu main.cu:279-479     -> handle-owned device buffers
//   set_quadrature    main.cu:124-194      -> tables_for()
//   preval_basis      basis.cu:168-244     -> tables_for()
//   preval_* kernels  conserv_kernels.cu:412-585 -> host geometry (FP32 sqrtf kept)
//   time_integrate_rk4 / rk2 time_integrator.cu:139-717 -> dg_step
// Every error is returned as a code with a message (no exit()).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dg.h"
#include "dg_internal.h"
#include "dg_ref_tables.h"

using namespace dgk;

// a run of steps replays graphs of 1, 2, 4, ..., 2^(kGraphLevels - 1) steps
// (binary decomposition of the count: at most kGraphLevels graph launches
// per 2^kGraphLevels - 1 steps, one gap between graph launches each)
static constexpr int kGraphLevels = 6;

struct dg_handle {
  dg_config cfg{};
  PhysParams ph{};
  int N = 0, np = 0, nq = 0, nq1 = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // mesh
  int K = 0, S = 0;
  double min_r = 0.;
  std::vector<double> hJ, hlen, hnx, hny, hdiam;
  double *d_xr = nullptr, *d_yr = nullptr, *d_xs = nullptr, *d_ys = nullptr, *d_J = nullptr;
  double *d_V = nullptr, *d_snx = nullptr, *d_sny = nullptr, *d_slen = nullptr, *d_eps = nullptr;
  int2 *d_slot = nullptr;
  int nb_bnd = 0;
  bool bnd_dirichlet = false;  // any outflow / inflow slot
  int *d_bnd_elem = nullptr, *d_bnd_code = nullptr;
  int4 *d_side = nullptr;
  double *d_bnd_val = nullptr;
  // host copies for the boundary-point queries: vertices [6][K], and per
  // boundary slot its element, local side | code << 2 and side index
  std::vector<double> hV;
  std::vector<int> h_bnd_elem, h_bnd_code, h_bnd_side;
  bool user_bnd = false;  // the caller's boundary states replace the problem's
  bool use_traces = true;  // runs of steps may store traces (traces_wanted)
  // state: d_c, d_b1, d_b2 (and d_acc) are views into one allocation,
  // d_state (block_views)
  double *d_state = nullptr;
  double *d_c = nullptr, *d_b1 = nullptr, *d_b2 = nullptr, *d_acc = nullptr;
  // stored traces of d_c, d_b1, d_b2 (four-lane stage kernels; allocated by
  // the first run of steps that uses them): views into d_trblk
  double *d_trblk = nullptr;
  double *d_tr[3] = {nullptr, nullptr, nullptr};
  double *d_k[4] = {nullptr, nullptr, nullptr, nullptr};
  double *d_err = nullptr, *d_lam = nullptr;
  double *d_prev = nullptr;  // step base copy (convergence norm on the limiter path)
  DevState *d_st = nullptr;
  unsigned long long *d_part = nullptr;  // per-block maxima of the step reduction
  int *d_bad = nullptr;                  // positivity flags of the stored state d_c
  // graphs of 2^q steps, q < kGraphLevels (same configuration key)
  hipGraphExec_t graph[kGraphLevels] = {};
  int graph_key = -1;
  double last_ms_per_stage = 0.;
  int last_stages = 0;
  double last_dt = 0.;
  // adjoint
  int obj_kind = -1;
  std::vector<double> obj_params;
  double *d_gw = nullptr;
  double adj_fwd_ms = 0., adj_rev_ms = 0.;
  // set by the stage and flag kernels wherever the positivity fix fires or a
  // stored state is flagged (allocated by the first adjoint solve); the last
  // adjoint solve's reading
  int *d_fired = nullptr;
  int adj_fix_fired = 0;
  int adj_resident = 0;  // the last adjoint solve's segments kept from its forward pass
};

// ------------------------------------------------------------ helpers

static int fail(dg_handle *h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

#define HIPCHK(h, x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return fail(h, DG_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));         \
  } while (0)

static int alloc(dg_handle *h, void **p, size_t bytes) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess)
    return fail(h, DG_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return DG_OK;
}

// n arrays of `bytes` each in one allocation *base (freed first if set),
// at 256-byte-aligned offsets; their addresses in v[0..n-1].  The state
// buffers allocated this way ran the stage kernels 1.7 % faster than with a
// hipMalloc each (Euler p = 2 on 524,288 and p = 4 on 2,097,152 triangles),
// and the adjoint's buffers (adj_alloc) its reverse sweep 5-8 % faster.
// Offsets padded by 4 KiB or 68 KiB past the 256-byte multiple measured the
// same as none.
static int block_views(dg_handle *h, double **base, size_t bytes, int n, double **v) {
  const size_t stride = (bytes + 255) / 256 * 256;
  const int rc = alloc(h, (void **)base, stride * n);
  if (rc != DG_OK) return rc;
  for (int q = 0; q < n; q++) v[q] = (double *)((char *)*base + q * stride);
  return DG_OK;
}

template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// ------------------------------------------------------------ tables

// phi, phi_x, phi_y of basis.cu:68-143: graded monomials with host pow
static double phi(double r, double s, int n) {
  double R[64];
  int k = 0, power = 0;
  while (k <= n) {
    for (int j = 0; j <= power; j++) R[k++] = std::pow(r, power - j) * std::pow(s, j);
    power++;
  }
  double sum = 0.;
  for (int i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
  return sum;
}
static double phi_r(double r, double s, int n) {
  double R[64];
  int k = 1, power = 1;
  R[0] = 0;
  while (k <= n) {
    for (int j = 0; j <= power; j++) {
      R[k] = (power - j) * std::pow(r, power - j - 1) * std::pow(s, j);
      if (j == power) R[k] = 0;
      k++;
    }
    power++;
  }
  double sum = 0.;
  for (int i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
  return sum;
}
static double phi_s(double r, double s, int n) {
  double R[64];
  int k = 1, power = 1;
  R[0] = 0;
  while (k <= n) {
    for (int j = 0; j <= power; j++) {
      R[k] = std::pow(r, power - j) * j * std::pow(s, j - 1);
      if (j == 0) R[k] = 0;
      k++;
    }
    power++;
  }
  double sum = 0.;
  for (int i = 0; i <= n; i++) sum += dgt_basis[n][i] * R[i];
  return sum;
}

template <int P>
static void fill_tab(Tab<P> &T) {
  constexpr int NP = Tab<P>::NP, NQ = Tab<P>::NQ, NQ1 = Tab<P>::NQ1;
  const double *q2 = dgt_q2[P], *q1 = dgt_q1[P];
  for (int j = 0; j < NQ; j++) {
    T.R1[j] = q2[3 * j];
    T.R2[j] = q2[3 * j + 1];
    T.W[j] = q2[3 * j + 2] / 2.;  // main.cu:181
  }
  for (int j = 0; j < NQ1; j++) {
    T.X1[j] = q1[2 * j];
    T.W1[j] = q1[2 * j + 1];
  }
  for (int i = 0; i < NP; i++) {
    T.VERT[i * 3 + 0] = phi(0., 0., i);
    T.VERT[i * 3 + 1] = phi(1., 0., i);
    T.VERT[i * 3 + 2] = phi(0., 1., i);
    for (int j = 0; j < NQ; j++) {
      T.B[i * NQ + j] = phi(T.R1[j], T.R2[j], i);
      T.GX[i * NQ + j] = T.W[j] * phi_r(T.R1[j], T.R2[j], i);
      T.GY[i * NQ + j] = T.W[j] * phi_s(T.R1[j], T.R2[j], i);
    }
    for (int j = 0; j < NQ1; j++) {
      const double *x = T.X1;
      T.SIDE[(0 * NP + i) * NQ1 + j] = phi(0.5 + 0.5 * x[j], 0., i);
      T.SIDE[(1 * NP + i) * NQ1 + j] = phi((1. - x[j]) / 2., (1. + x[j]) / 2., i);
      T.SIDE[(2 * NP + i) * NQ1 + j] = phi(0., 0.5 + 0.5 * x[NQ1 - 1 - j], i);
    }
  }
}

static std::mutex g_tab_mu;
static bool g_tab_done[64][PMAX + 1];

template <int P>
static int upload_p(int dev) {
  static Tab<P> T;
  static bool filled = false;
  if (!filled) {
    fill_tab<P>(T);
    filled = true;
  }
  if (g_tab_done[dev][P]) return 0;
  if (upload_tables(P, &T, sizeof(T)) != 0) return -1;
  g_tab_done[dev][P] = true;
  return 0;
}

// device address of each GPU's fixed scratch (dg_eval_rhs_noalloc), found
// once per device with that device current
static std::mutex g_scratch_mu;
static void *g_scratch[64];

static int ensure_scratch(int dev) {
  if (dev < 0 || dev >= 64) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return 0;
  void *p = nullptr;
  if (scratch_address(&p) != 0) return -1;
  g_scratch[dev] = p;
  return 0;
}

static int ensure_tables(int dev, int p) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (dev < 0 || dev >= 64) return -1;
  switch (p) {
    case 0: return upload_p<0>(dev);
    case 1: return upload_p<1>(dev);
    case 2: return upload_p<2>(dev);
    case 3: return upload_p<3>(dev);
    case 4: return upload_p<4>(dev);
    case 5: return upload_p<5>(dev);
  }
  return -1;
}

// ------------------------------------------------------------ API

extern "C" void dg_config_default(dg_config *c) {
  std::memset(c, 0, sizeof *c);
  c->physics = DG_EULER;
  c->p = 1;
  c->flux = DG_FLUX_DEFAULT;
  c->integrator = DG_RK4;
  c->limiter = 0;
  c->positivity = 1;
  c->cfl = 0.;
  c->gamma = 1.4;
  c->g = 9.8;  // shallowwater.cu:7
  c->eps = 1.;
  c->mu = 1.;
  c->ax = 1.;
  c->ay = 1.;
  c->problem = DG_PROB_EULER_CONST;
  c->problem_params[0] = 1.;
  c->problem_params[1] = 1.;
  c->problem_params[2] = 1.;
  c->problem_params[3] = 1.;
  c->device = 0;
}

extern "C" const char *dg_last_error(const dg_handle *h) { return h ? h->err.c_str() : "null handle"; }

extern "C" int dg_create(dg_handle **out, const dg_config *cfg) {
  if (!out || !cfg) return DG_ERR_ARG;
  *out = nullptr;
  if (cfg->physics < 0 || cfg->physics > 4) return DG_ERR_ARG;
  if (cfg->p < 0 || cfg->p > PMAX) return DG_ERR_ARG;
  if (cfg->integrator < DG_RK4 || cfg->integrator > DG_SSPRK3) return DG_ERR_ARG;
  if (cfg->limiter && cfg->p != 1) return DG_ERR_ARG;  // main.cu:647-650
  if (cfg->flux == DG_FLUX_ROE && cfg->physics != DG_SWE) return DG_ERR_ARG;
  if (cfg->problem < DG_PROB_ADV_SINSIN || cfg->problem > DG_PROB_ADV_ROTHILL) return DG_ERR_ARG;
  if (cfg->problem == DG_PROB_ADV_ROTHILL && cfg->physics != DG_ADVECTION) return DG_ERR_ARG;
  if (cfg->volume < 0 || cfg->volume > 2) return DG_ERR_ARG;
  if (cfg->volume == 2 && !(cfg->physics == DG_EULER && cfg->p == 4)) return DG_ERR_ARG;
  dg_handle *h = new dg_handle();
  h->cfg = *cfg;
  if (h->cfg.cfl <= 0) h->cfg.cfl = (cfg->integrator == DG_RK2MID) ? 0.7 : 0.5;
  h->N = nvars_of(cfg->physics);
  h->np = np_of(cfg->p);
  h->nq = nq_of(cfg->p);
  h->nq1 = cfg->p + 1;
  h->ph.gamma = cfg->gamma;
  h->ph.G = cfg->g;
  h->ph.eps = cfg->eps;
  h->ph.mu = cfg->mu;
  h->ph.ax = cfg->ax;
  h->ph.ay = cfg->ay;
  h->ph.prob = cfg->problem;
  h->use_traces = std::getenv("DG_STAGE_GATHER") == nullptr;
  h->ph.flux = cfg->flux;
  for (int i = 0; i < 8; i++) h->ph.pp[i] = cfg->problem_params[i];
  hipError_t e = hipSetDevice(cfg->device);
  if (e != hipSuccess) {
    delete h;
    return DG_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      ensure_tables(cfg->device, cfg->p) != 0 || ensure_scratch(cfg->device) != 0 ||
      hipMalloc((void **)&h->d_st, sizeof(DevState)) != hipSuccess) {
    dg_destroy(h);
    return DG_ERR_HIP;
  }
  *out = h;
  return DG_OK;
}

static void drop_graphs(dg_handle *h) {
  for (auto &g : h->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  h->graph_key = -1;
}

static void free_mesh(dg_handle *h) {
  dfree(h->d_xr); dfree(h->d_yr); dfree(h->d_xs); dfree(h->d_ys); dfree(h->d_J);
  dfree(h->d_V); dfree(h->d_snx); dfree(h->d_sny); dfree(h->d_slen); dfree(h->d_eps);
  dfree(h->d_slot);
  dfree(h->d_bnd_elem); dfree(h->d_bnd_code); dfree(h->d_bnd_val); dfree(h->d_side);
  h->nb_bnd = 0;
  h->bnd_dirichlet = false;
  h->user_bnd = false;
  dfree(h->d_state);
  h->d_c = h->d_b1 = h->d_b2 = h->d_acc = nullptr;
  dfree(h->d_trblk);
  for (auto &t : h->d_tr) t = nullptr;
  for (auto &k : h->d_k) dfree(k);
  dfree(h->d_err); dfree(h->d_lam); dfree(h->d_gw); dfree(h->d_prev); dfree(h->d_part); dfree(h->d_bad);
  drop_graphs(h);
  // no mesh from here until a new one is fully uploaded: every call that
  // needs one checks K, and the objective's weights went with the buffers
  h->K = 0;
  h->S = 0;
  h->obj_kind = -1;
}

extern "C" void dg_destroy(dg_handle *h) {
  if (!h) return;
  (void)hipSetDevice(h->cfg.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_mesh(h);
  dfree(h->d_st);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static MeshDev mesh_dev(const dg_handle *h) {
  MeshDev m;
  m.K = h->K;
  m.xr = h->d_xr; m.yr = h->d_yr; m.xs = h->d_xs; m.ys = h->d_ys; m.J = h->d_J;
  m.V = h->d_V;
  m.slot = h->d_slot;
  m.snx = h->d_snx; m.sny = h->d_sny; m.slen = h->d_slen;
  m.eps = h->d_eps;
  m.nb_bnd = h->nb_bnd;
  m.bnd_elem = h->d_bnd_elem;
  m.bnd_code = h->d_bnd_code;
  m.S = h->S;
  m.side = h->d_side;
  m.bnd_val = h->d_bnd_val;
  return m;
}

static size_t dofs(const dg_handle *h) { return (size_t)h->K * h->np * h->N; }

extern "C" int dg_set_mesh(dg_handle *h, int K, const double *V, const int *es, int S,
                           const double *sxy, const int *left, const int *right, const int *lsn,
                           const int *rsn) {
  if (!h) return DG_ERR_ARG;
  if (K <= 0 || S <= 0 || !V || !es || !sxy || !left || !right || !lsn || !rsn)
    return fail(h, DG_ERR_ARG, "dg_set_mesh: empty mesh or null array");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // the new mesh is validated and precomputed on the host first; the old
  // mesh is freed only once the new one is known to be good
  const double *V1x = V, *V1y = V + (size_t)K, *V2x = V + 2 * (size_t)K, *V2y = V + 3 * (size_t)K;
  const double *V3x = V + 4 * (size_t)K, *V3y = V + 5 * (size_t)K;
  const double *sx1 = sxy, *sy1 = sxy + S, *sx2 = sxy + 2 * (size_t)S, *sy2 = sxy + 3 * (size_t)S;
  // side checks
  for (int s = 0; s < S; s++) {
    if (left[s] < 0 || left[s] >= K || right[s] >= K || right[s] < -3 || lsn[s] < 0 ||
        lsn[s] > 2 || (right[s] >= 0 && (rsn[s] < 0 || rsn[s] > 2)))
      return fail(h, DG_ERR_ARG, "dg_set_mesh: side %d has bad connectivity", s);
  }
  // geometry, the precompute order of main.cu:698-769
  std::vector<double> xr(K), yr(K), xs(K), ys(K), J(K), diam(K), len(S), Nx(S), Ny(S);
  for (int e = 0; e < K; e++) {  // preval_inscribed_circles conserv_kernels.cu:428-445
    double dx, dy;
    dx = V1x[e] - V2x[e]; dy = V1y[e] - V2y[e];
    double a = sqrtf((float)(dx * dx + dy * dy));
    dx = V2x[e] - V3x[e]; dy = V2y[e] - V3y[e];
    double b = sqrtf((float)(dx * dx + dy * dy));
    dx = V1x[e] - V3x[e]; dy = V1y[e] - V3y[e];
    double c = sqrtf((float)(dx * dx + dy * dy));
    double k = 0.5 * (a + b + c);
    diam[e] = 2 * sqrtf((float)(k * (k - a) * (k - b) * (k - c))) / k;
  }
  double min_r = diam[0];  // main.cu:722-736
  for (int e = 1; e < K; e++) min_r = diam[e] < min_r ? diam[e] : min_r;
  for (int e = 0; e < K; e++) {  // preval_jacobian :452-473, preval_partials :571-585
    J[e] = (V2x[e] - V1x[e]) * (V3y[e] - V1y[e]) - (V3x[e] - V1x[e]) * (V2y[e] - V1y[e]);
    xr[e] = V2x[e] - V1x[e];
    yr[e] = V2y[e] - V1y[e];
    xs[e] = V3x[e] - V1x[e];
    ys[e] = V3y[e] - V1y[e];
    if (!(J[e] > 0))
      return fail(h, DG_ERR_ARG, "dg_set_mesh: element %d is not counter-clockwise (J=%g)", e,
                  J[e]);
  }
  for (int s = 0; s < S; s++) {  // preval_side_length :412-421, preval_normals :481-511
    double dx = sx1[s] - sx2[s], dy = sy1[s] - sy2[s];
    len[s] = sqrtf((float)(dx * dx + dy * dy));
    double x = sx2[s] - sx1[s], y = sy2[s] - sy1[s];
    double l = sqrtf((float)(x * x + y * y));
    Nx[s] = -y / l;
    Ny[s] = x / l;
    // preval_normals_direction :513-569: point away from the left element
    int L = left[s];
    double tx, ty, ix, iy;
    switch (lsn[s]) {
      case 0: tx = V3x[L]; ty = V3y[L]; ix = (V1x[L] + V2x[L]) / 2.; iy = (V1y[L] + V2y[L]) / 2.; break;
      case 1: tx = V1x[L]; ty = V1y[L]; ix = (V2x[L] + V3x[L]) / 2.; iy = (V2y[L] + V3y[L]) / 2.; break;
      default: tx = V2x[L]; ty = V2y[L]; ix = (V1x[L] + V3x[L]) / 2.; iy = (V1y[L] + V3y[L]) / 2.; break;
    }
    if (Nx[s] * (tx - ix) + Ny[s] * (ty - iy) > 0) {
      Nx[s] *= -1;
      Ny[s] *= -1;
    }
  }
  // per-element side slots: slot k holds the element's local side k
  std::vector<int2> slot(3 * (size_t)K);
  std::vector<int> bnd_elem, bnd_code;
  std::vector<int4> side(S);
  for (int s = 0; s < S; s++)
    side[s] = make_int4(left[s], right[s], lsn[s] | ((right[s] >= 0 ? rsn[s] : 0) << 2), -1);
  if ((long long)S >= (1ll << 28))
    return fail(h, DG_ERR_ARG, "dg_set_mesh: %d sides exceed the 2^28 the slot encoding holds", S);
  // the element kernels address the state with 32-bit element offsets
  // (below 2^28); the four-lane Euler kernel (p >= 4) adds a variable's row
  // offset to them, so N * n_p * K must stay below 2^28 there
  if ((long long)K * h->N * h->np >= (1ll << 28) && h->cfg.physics == DG_EULER && h->cfg.p >= 4)
    return fail(h, DG_ERR_ARG, "dg_set_mesh: %d elements x %d coefficients exceed the 2^28 "
                "entries the p >= 4 Euler kernel addresses", K, h->N * h->np);
  if ((long long)K >= (1ll << 28))
    return fail(h, DG_ERR_ARG, "dg_set_mesh: %d elements exceed 2^28", K);
  bool dirichlet = false;
  std::vector<unsigned char> seen(S, 0);  // bit 0: listed by its left element, bit 1: by its right
  for (int e = 0; e < K; e++) {
    int filled = 0;
    for (int q = 0; q < 3; q++) {
      int s = es[(size_t)q * K + e];
      if (s < 0 || s >= S) return fail(h, DG_ERR_ARG, "dg_set_mesh: element %d side id %d", e, s);
      int ls, nb, nbls, role;
      if (left[s] == e) {
        ls = lsn[s]; nb = right[s]; nbls = nb >= 0 ? rsn[s] : 0; role = 0;
      } else if (right[s] == e) {
        ls = rsn[s]; nb = left[s]; nbls = lsn[s]; role = 1;
      } else {
        return fail(h, DG_ERR_ARG, "dg_set_mesh: side %d does not touch element %d", s, e);
      }
      if (filled & (1 << ls))
        return fail(h, DG_ERR_ARG, "dg_set_mesh: element %d has local side %d twice", e, ls);
      filled |= 1 << ls;
      seen[s] |= (unsigned char)(1 << role);
      const int info = (nbls & 3) | (role << 2) | (s << 3);
      if (nb < 0) {  // boundary slot: index into the boundary data
        if (role) return fail(h, DG_ERR_ARG, "dg_set_mesh: boundary side %d must have its element on the left", s);
        if (nb < -3) return fail(h, DG_ERR_ARG, "dg_set_mesh: side %d has boundary code %d", s, nb);
        side[s].w = (int)bnd_elem.size();
        bnd_elem.push_back(e);
        bnd_code.push_back(ls | ((-nb) << 2));
        if (nb != -1) dirichlet = true;
      }
      slot[(size_t)ls * K + e] = make_int2(nb, info);
    }
  }
  // every side is among its elements' sides (the per-side kernels rely on
  // both elements, and a boundary side on its boundary data)
  for (int s = 0; s < S; s++)
    if (seen[s] != (right[s] >= 0 ? 3 : 1))
      return fail(h, DG_ERR_ARG, "dg_set_mesh: side %d is not listed by its %s element", s,
                  (seen[s] & 1) ? "right" : "left");
  free_mesh(h);
  h->min_r = min_r;
  h->hJ = J;
  h->hlen = len;
  h->hnx = Nx;
  h->hny = Ny;
  h->hdiam = diam;
  const size_t kd = sizeof(double) * (size_t)K;
  int rc;
#define UP(dst, src, bytes)                                                        \
  if ((rc = alloc(h, (void **)&dst, bytes)) != DG_OK) return rc;                  \
  HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  UP(h->d_xr, xr.data(), kd);
  UP(h->d_yr, yr.data(), kd);
  UP(h->d_xs, xs.data(), kd);
  UP(h->d_ys, ys.data(), kd);
  UP(h->d_J, J.data(), kd);
  UP(h->d_V, V, 6 * kd);
  UP(h->d_slot, slot.data(), 3 * sizeof(int2) * (size_t)K);
  UP(h->d_snx, Nx.data(), sizeof(double) * (size_t)S);
  UP(h->d_sny, Ny.data(), sizeof(double) * (size_t)S);
  UP(h->d_slen, len.data(), sizeof(double) * (size_t)S);
  UP(h->d_side, side.data(), sizeof(int4) * (size_t)S);
  h->nb_bnd = (int)bnd_elem.size();
  h->bnd_dirichlet = dirichlet;
  h->hV.assign(V, V + 6 * (size_t)K);
  h->h_bnd_elem = bnd_elem;
  h->h_bnd_code = bnd_code;
  h->h_bnd_side.resize(bnd_elem.size());
  for (int q = 0; q < (int)bnd_elem.size(); q++)
    h->h_bnd_side[q] = slot[(size_t)(bnd_code[q] & 3) * K + bnd_elem[q]].y >> 3;
  if (h->nb_bnd) {
    UP(h->d_bnd_elem, bnd_elem.data(), sizeof(int) * bnd_elem.size());
    UP(h->d_bnd_code, bnd_code.data(), sizeof(int) * bnd_code.size());
    if ((rc = alloc(h, (void **)&h->d_bnd_val, sizeof(double) * bnd_elem.size() * h->nq1 * NMAX)) != DG_OK)
      return rc;
  }
#undef UP
  std::vector<double> ones(K, h->cfg.eps);
  if ((rc = alloc(h, (void **)&h->d_eps, kd)) != DG_OK) return rc;
  HIPCHK(h, hipMemcpy(h->d_eps, ones.data(), kd, hipMemcpyHostToDevice));
  const size_t sb = sizeof(double) * (size_t)K * h->np * h->N;
  {
    double *v[4];
    const int nv = h->cfg.integrator == DG_RK4 && !h->cfg.limiter ? 4 : 3;
    if ((rc = block_views(h, &h->d_state, sb, nv, v)) != DG_OK) return rc;
    h->d_c = v[0];
    h->d_b1 = v[1];
    h->d_b2 = v[2];
    h->d_acc = nv == 4 ? v[3] : nullptr;
  }
  if (h->cfg.limiter) {
    int nk = h->cfg.integrator == DG_RK4 ? 4 : 1;
    for (int q = 0; q < nk; q++)
      if ((rc = alloc(h, (void **)&h->d_k[q], sb)) != DG_OK) return rc;
  }
  if ((rc = alloc(h, (void **)&h->d_err, sizeof(double) * (size_t)K * h->N)) != DG_OK) return rc;
  if ((rc = alloc(h, (void **)&h->d_lam, kd)) != DG_OK) return rc;
  if ((rc = alloc(h, (void **)&h->d_bad, sizeof(int) * (size_t)K)) != DG_OK) return rc;
  if ((rc = alloc(h, (void **)&h->d_part,
                  sizeof(unsigned long long) * 2 * (size_t)((K + 63) / 64))) != DG_OK)
    return rc;
  HIPCHK(h, hipMemset(h->d_c, 0, sb));
  // every buffer is in place: the handle has a mesh from here on
  h->K = K;
  h->S = S;
  return DG_OK;
}

// read_mesh main.cu:205-277
extern "C" int dg_load_pmsh(dg_handle *h, const char *path) {
  if (!h || !path) return DG_ERR_ARG;
  FILE *f = std::fopen(path, "r");
  if (!f) return fail(h, DG_ERR_IO, "mesh file not found: %s", path);
  char line[1000];
  int K = 0, S = 0;
  if (!std::fgets(line, sizeof line, f) || std::sscanf(line, "%d", &K) != 1 || K <= 0) {
    std::fclose(f);
    return fail(h, DG_ERR_IO, "%s: bad element count", path);
  }
  std::vector<double> V(6 * (size_t)K);
  std::vector<int> es(3 * (size_t)K);
  for (int i = 0; i < K; i++) {
    double a[6];
    int s1, s2, s3;
    if (!std::fgets(line, sizeof line, f) ||
        std::sscanf(line, "%lf %lf %lf %lf %lf %lf %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4],
                    &a[5], &s1, &s2, &s3) != 9) {
      std::fclose(f);
      return fail(h, DG_ERR_IO, "%s: not enough items while reading element %d", path, i);
    }
    for (int q = 0; q < 6; q++) V[(size_t)q * K + i] = a[q];
    es[i] = s1;
    es[(size_t)K + i] = s2;
    es[2 * (size_t)K + i] = s3;
  }
  if (!std::fgets(line, sizeof line, f) || std::sscanf(line, "%d", &S) != 1 || S <= 0) {
    std::fclose(f);
    return fail(h, DG_ERR_IO, "%s: bad side count", path);
  }
  std::vector<double> sxy(4 * (size_t)S);
  std::vector<int> L(S), R(S), ls(S), rs(S);
  for (int i = 0; i < S; i++) {
    double a[4];
    if (!std::fgets(line, sizeof line, f) ||
        std::sscanf(line, "%lf %lf %lf %lf %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &L[i], &R[i],
                    &ls[i], &rs[i]) != 8) {
      std::fclose(f);
      return fail(h, DG_ERR_IO, "%s: not enough items while reading side %d", path, i);
    }
    for (int q = 0; q < 4; q++) sxy[(size_t)q * S + i] = a[q];
  }
  std::fclose(f);
  return dg_set_mesh(h, K, V.data(), es.data(), S, sxy.data(), L.data(), R.data(), ls.data(),
                     rs.data());
}

extern "C" int dg_set_eps(dg_handle *h, const double *eps) {
  if (!h || !eps) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_set_eps: no mesh");
  for (int e = 0; e < h->K; e++)
    if (!(eps[e] > 0)) return fail(h, DG_ERR_ARG, "dg_set_eps: eps[%d] must be positive", e);
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpy(h->d_eps, eps, sizeof(double) * h->K, hipMemcpyHostToDevice));
  return DG_OK;
}

static bool posfix_on(const dg_handle *h) {
  return h->cfg.positivity && (h->cfg.physics == DG_EULER || h->cfg.physics == DG_SWE);
}

extern "C" int dg_project_ic(dg_handle *h) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_project_ic: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  launch_init(h->cfg.physics, h->cfg.p, mesh_dev(h), h->ph, h->d_c, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

static hipMemcpyKind kind_to(int where) {
  return where == DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
static hipMemcpyKind kind_from(int where) {
  return where == DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

extern "C" int dg_set_coeffs(dg_handle *h, const double *C, int where) {
  if (!h || !C) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_set_coeffs: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpyAsync(h->d_c, C, sizeof(double) * dofs(h), kind_to(where), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

extern "C" int dg_get_coeffs(dg_handle *h, double *C, int where) {
  if (!h || !C) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_get_coeffs: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpyAsync(C, h->d_c, sizeof(double) * dofs(h), kind_from(where), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

// whether the stage runs its volume integral on the matrix cores
// (dg_config.volume; the library's choice is the vector FMAs)
static bool volume_mfma(const dg_config &c) { return c.volume == 2; }

static StageArgs base_args(dg_handle *h) {
  StageArgs a;
  std::memset(&a, 0, sizeof a);
  a.m = mesh_dev(h);
  a.st = h->d_st;
  a.ph = h->ph;
  a.cfl_num = h->cfg.cfl * h->min_r;
  a.part = h->d_part;
  a.bad = h->d_bad;  // valid always; used only where flags_in / flags_out say so
  a.vol_mfma = volume_mfma(h->cfg) ? 1 : 0;
  a.fired = h->d_fired;
  return a;
}

static int write_state(dg_handle *h, double t, double dt, double t_end = HUGE_VAL,
                       double conv_tol = -1.) {
  DevState s;
  std::memset(&s, 0, sizeof s);
  s.t = t;
  s.dt = dt;
  s.tdone = t;
  s.t_end = t_end;
  s.conv_tol = conv_tol;
  HIPCHK(h, hipMemcpyAsync(h->d_st, &s, sizeof s, hipMemcpyHostToDevice, h->stream));
  return DG_OK;
}

// boundary data depends on time only for these problems (the caller's
// boundary states do not)
static bool bnd_time_dependent(const dg_handle *h) {
  return h->bnd_dirichlet && !h->user_bnd &&
         (h->cfg.problem == DG_PROB_ADV_SINSIN || h->cfg.problem == DG_PROB_EULER_VORTEX ||
          h->cfg.problem == DG_PROB_ADV_ROTHILL);
}

// boundary data at a fixed time (once per call when it does not change)
static void refresh_bnd_fixed(dg_handle *h, double t) {
  if (h->nb_bnd && !h->user_bnd)
    launch_bnd(h->cfg.physics, h->cfg.p, mesh_dev(h), h->ph, h->d_st, 0., 0, t, h->stream);
}

// one stage launch, preceded by the boundary data at the stage time when needed
static void launch_stage_bnd(dg_handle *h, const StageArgs &a) {
  if (bnd_time_dependent(h))
    launch_bnd(h->cfg.physics, h->cfg.p, a.m, h->ph, h->d_st, a.cstage, 1, 0., h->stream);
  launch_stage(h->cfg.physics, h->cfg.p, a, h->stream);
}

extern "C" int dg_eval_rhs(dg_handle *h, const double *u, double t, double dt, double *k,
                           int where) {
  if (!h || !u || !k) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_eval_rhs: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  const size_t sb = sizeof(double) * dofs(h);
  HIPCHK(h, hipMemcpyAsync(h->d_b1, u, sb, kind_to(where), h->stream));
  int rc = write_state(h, t, dt);
  if (rc) return rc;
  StageArgs a = base_args(h);
  a.us = h->d_b1;
  a.un = h->d_b1;
  a.out = h->d_b2;
  a.mode = M_KONLY;
  if (posfix_on(h)) {
    // the positivity prepass on the input, through its flags (the stage
    // kernel takes a neighbour's fix from its flag)
    launch_flags(h->cfg.physics, h->cfg.p, h->K, h->ph, h->d_b1, h->d_bad, h->stream);
    a.fix_in = 1;
    a.flags_in = 1;
  }
  refresh_bnd_fixed(h, t);
  launch_stage(h->cfg.physics, h->cfg.p, a, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(k, h->d_b2, sb, kind_from(where), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

// ------------------------------------------------------------ user boundary states
// (the reference's U_inflow / U_outflow / U_reflection plugin hooks,
// conserv_headers.cuh:5-34, as data)

extern "C" int dg_boundary_points(dg_handle *h, int *nb, int *side, int *code, double *x,
                                  double *y) {
  if (!h || !nb) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_boundary_points: no mesh");
  *nb = h->nb_bnd;
  const int K = h->K, NQ1 = h->nq1;
  const double *q1 = dgt_q1[h->cfg.p];
  for (int b = 0; b < h->nb_bnd; b++) {
    const int e = h->h_bnd_elem[b], k = h->h_bnd_code[b] & 3;
    if (side) side[b] = h->h_bnd_side[b];
    if (code) code[b] = -(h->h_bnd_code[b] >> 2);
    if (!x || !y) continue;
    const double *V = h->hV.data();
    const double V0 = V[e], V1 = V[(size_t)K + e], V2 = V[2 * (size_t)K + e];
    const double V3 = V[3 * (size_t)K + e], V4 = V[4 * (size_t)K + e], V5 = V[5 * (size_t)K + e];
    for (int j = 0; j < NQ1; j++) {  // the points k_bnd evaluates (get_coordinates_1d)
      double ra, rb;
      if (k == 0) {
        ra = 0.5 + 0.5 * q1[2 * j];
        rb = 0.;
      } else if (k == 1) {
        ra = (1. - q1[2 * j]) / 2.;
        rb = (1. + q1[2 * j]) / 2.;
      } else {
        ra = 0.;
        rb = 0.5 + 0.5 * q1[2 * (NQ1 - 1 - j)];
      }
      x[(size_t)b * NQ1 + j] = V2 * ra + V4 * rb + V0 * (1 - ra - rb);
      y[(size_t)b * NQ1 + j] = V3 * ra + V5 * rb + V1 * (1 - ra - rb);
    }
  }
  return DG_OK;
}

extern "C" int dg_set_boundary_states(dg_handle *h, const double *U, int where) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_set_boundary_states: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  h->user_bnd = false;
  if (!U || !h->nb_bnd) return DG_OK;  // the problem's boundary data again
  const size_t NQ1 = h->nq1, N = h->N, nb = h->nb_bnd;
  std::vector<double> u(nb * NQ1 * N), bv(nb * NQ1 * NMAX, 0.);
  HIPCHK(h, hipMemcpy(u.data(), U, sizeof(double) * u.size(),
                      where == DG_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
  for (size_t b = 0; b < nb; b++)
    for (size_t j = 0; j < NQ1; j++)
      for (size_t n = 0; n < N; n++) bv[(b * NQ1 + j) * NMAX + n] = u[(b * NQ1 + j) * N + n];
  HIPCHK(h, hipMemcpy(h->d_bnd_val, bv.data(), sizeof(double) * bv.size(), hipMemcpyHostToDevice));
  h->user_bnd = true;
  return DG_OK;
}

extern "C" int dg_get_boundary_states(dg_handle *h, double t, double *U) {
  if (!h || !U) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_get_boundary_states: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  if (!h->nb_bnd) return DG_OK;
  refresh_bnd_fixed(h, t);  // the problem's at t (the caller's stay as set)
  const size_t NQ1 = h->nq1, N = h->N, nb = h->nb_bnd;
  std::vector<double> bv(nb * NQ1 * NMAX);
  HIPCHK(h, hipMemcpyAsync(bv.data(), h->d_bnd_val, sizeof(double) * bv.size(),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t b = 0; b < nb; b++)
    for (size_t j = 0; j < NQ1; j++)
      for (size_t n = 0; n < N; n++) U[(b * NQ1 + j) * N + n] = bv[(b * NQ1 + j) * NMAX + n];
  return DG_OK;
}

// dg_eval_rhs with device pointers and no allocation: the input is read in
// place and the step state and positivity flags live in the GPU's fixed
// scratch (its address found at dg_create).  Asynchronous on the handle's
// stream; every launch is a kernel (no memset or copy), so the sequence can
// be captured.
extern "C" int dg_eval_rhs_noalloc(dg_handle *h, const double *u, double t, double dt, double *k) {
  if (!h || !u || !k) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_eval_rhs_noalloc: no mesh");
  if (h->K > kScratchMaxK)
    return fail(h, DG_ERR_ARG,
                "dg_eval_rhs_noalloc: K = %d elements exceed the fixed scratch's %d", h->K,
                kScratchMaxK);
  const size_t sb = sizeof(double) * dofs(h);
  const char *ub = (const char *)u, *kb = (const char *)k;
  if (ub < kb + sb && kb < ub + sb)
    return fail(h, DG_ERR_ARG, "dg_eval_rhs_noalloc: u and k overlap (k is written while u is read)");
  const int dev = h->cfg.device;
  char *scr = (char *)g_scratch[dev];
  if (!scr) return fail(h, DG_ERR_STATE, "dg_eval_rhs_noalloc: no scratch address for device %d", dev);
  HIPCHK(h, hipSetDevice(dev));
  DevState *st = (DevState *)scr;
  int *bad = (int *)(scr + 256);
  launch_set_state(st, t, dt, h->stream);
  StageArgs a = base_args(h);
  a.st = st;
  a.bad = bad;
  a.us = u;
  a.un = u;
  a.out = k;
  a.mode = M_KONLY;
  if (posfix_on(h)) {
    launch_flags(h->cfg.physics, h->cfg.p, h->K, h->ph, u, bad, h->stream);
    a.fix_in = 1;
    a.flags_in = 1;
  }
  refresh_bnd_fixed(h, t);
  launch_stage(h->cfg.physics, h->cfg.p, a, h->stream);
  HIPCHK(h, hipGetLastError());
  return DG_OK;
}

// stage schedule of one fused step; io[s] is the input of stage s and
// io[S] the step's output (io[0] is the step's base state); returns S
static int make_stages(const StageArgs &a, int integrator, int pf, StageArgs *st,
                       double *const *io, double *acc, bool conv) {
  int n = 0;
  auto push = [&](int mode, double A, double B, double C, int accm, double cst) {
    StageArgs s = a;
    s.us = io[n];
    s.un = io[0];
    s.out = io[n + 1];
    s.acc = acc;
    s.mode = mode;
    s.A = A;
    s.B = B;
    s.C = C;
    s.accmode = accm;
    s.cstage = cst;
    // the first stage reads the step's state as computed; later stages read
    // fixed intermediate states and the step base (fixed again on read)
    s.fix_in = n == 0 ? pf : 0;
    s.fix_un = (mode == M_A_UN_B_US_CK || mode == M_UN_CK || accm == ACC_FINAL6) ? pf : 0;
    s.fix_out = pf;
    st[n++] = s;
  };
  switch (integrator) {
    case DG_RK4:  // time_integrator.cu:244-393
      push(M_US_CK, 0, 1, 0.5, ACC_SET6, 0.0);
      push(M_UN_CK, 0, 0, 0.5, ACC_ADD3, 0.5);
      push(M_UN_CK, 0, 0, 1.0, ACC_ADD3, 0.5);
      push(M_UN_CK, 0, 0, 1.0, ACC_FINAL6, 1.0);
      break;
    case DG_RK2MID:  // time_integrator.cu:599-674
      push(M_US_CK, 0, 1, 0.5, ACC_NONE, 0.0);
      push(M_UN_CK, 0, 0, 1.0, ACC_NONE, 0.5);
      break;
    case DG_SSPRK2:
      push(M_US_CK, 0, 1, 1.0, ACC_NONE, 0.0);
      push(M_A_UN_B_US_CK, 0.5, 0.5, 0.5, ACC_NONE, 1.0);
      break;
    case DG_SSPRK3:
      push(M_US_CK, 0, 1, 1.0, ACC_NONE, 0.0);
      push(M_A_UN_B_US_CK, 0.75, 0.25, 0.25, ACC_NONE, 1.0);
      push(M_A_UN_B_US_CK, 1. / 3., 2. / 3., 2. / 3., ACC_NONE, 0.5);
      break;
  }
  st[n - 1].last_stage = 1;
  st[n - 1].fix_out = 0;
  // the reference's convergence norm reads the first K * N entries of the
  // [N][n_p][K] state, i.e. rows n * n_p + i < N (time_integrator.cu:404-414)
  st[n - 1].conv = conv ? 1 : 0;
  return n;
}

static int build_step_io(dg_handle *h, StageArgs *st, int cfl_mode, double *const *io,
                         double *acc, bool conv = false) {
  StageArgs a = base_args(h);
  a.cfl_mode = cfl_mode;
  return make_stages(a, h->cfg.integrator, posfix_on(h) ? 1 : 0, st, io, acc, conv);
}

static int nstages_of(int integ) {
  return integ == DG_RK4 ? 4 : (integ == DG_SSPRK3 ? 3 : 2);
}

// Whether runs of steps use stored traces: the four-lane kernels (Euler
// p >= 4), with trace offsets within the 2^28 the element kernels address,
// unless DG_STAGE_GATHER is set in the environment when the handle is
// created (then every stage gathers the neighbours' modes).
static bool traces_wanted(const dg_handle *h) {
  return h->use_traces && h->cfg.physics == DG_EULER && h->cfg.p >= 4 &&
         (long long)tr_elem(h->cfg.p) * h->K < (1ll << 28);
}

// the trace buffer that goes with state buffer u (nullptr: none)
static double *traces_of(dg_handle *h, const double *u) {
  if (!h->d_tr[0]) return nullptr;
  if (u == h->d_c) return h->d_tr[0];
  if (u == h->d_b1) return h->d_tr[1];
  if (u == h->d_b2) return h->d_tr[2];
  return nullptr;
}

// the stored-trace buffers of the three state buffers (once per mesh)
static int alloc_traces(dg_handle *h) {
  if (h->d_trblk) return DG_OK;
  const int rc = block_views(h, &h->d_trblk, sizeof(double) * tr_elem(h->cfg.p) * (size_t)h->K,
                             3, h->d_tr);
  if (rc != DG_OK) return rc;
  drop_graphs(h);  // graphs captured before carry no trace pointers
  return DG_OK;
}

static int build_step(dg_handle *h, StageArgs *st, int cfl_mode, bool conv = false) {
  double *io[5];
  const int S = nstages_of(h->cfg.integrator);
  // ping-pong between the two stage buffers; the last stage writes the state
  io[0] = h->d_c;
  for (int s = 1; s < S; s++) io[s] = (s & 1) ? h->d_b1 : h->d_b2;
  io[S] = h->d_c;
  const int ns = build_step_io(h, st, cfl_mode, io, h->d_acc, conv);
  // each stage reads its input's stored traces and stores its output's
  for (int q = 0; q < ns; q++) {
    st[q].tr_in = traces_of(h, io[q]);
    st[q].tr_out = st[q].tr_in ? traces_of(h, io[q + 1]) : nullptr;
  }
  // the stored state d_c carries positivity flags (written with it)
  if (posfix_on(h))
    for (int q = 0; q < ns; q++) {
      st[q].flags_in = 1;
      st[q].flags_out = 1;
    }
  return ns;
}

// one step with separate launches and the limiter (the reference's order:
// eval -> limit(k) -> tempstorage, ..., combine -> limit(c))
static void limited_step(dg_handle *h, int cfl_mode, bool conv) {
  const int phys = h->cfg.physics, p = h->cfg.p;
  const long long n = (long long)dofs(h);
  MeshDev m = mesh_dev(h);
  StageArgs a = base_args(h);
  a.mode = M_KONLY;
  auto eval = [&](const double *us, double *k, double cst) {
    StageArgs s = a;
    s.us = us;
    s.un = us;
    s.out = k;
    s.cstage = cst;
    if (posfix_on(h)) launch_posfix(phys, p, h->K, h->ph, const_cast<double *>(us), h->stream);
    launch_stage_bnd(h, s);
    launch_limit(phys, p, m, k, h->stream);
  };
  double *c = h->d_c, **k = h->d_k;
  if (conv) launch_copy(n, h->d_prev, c, h->stream);
  switch (h->cfg.integrator) {
    case DG_RK4:
      eval(c, k[0], 0.0);
      launch_axpy(n, h->d_b1, c, 0.5, k[0], h->stream);
      eval(h->d_b1, k[1], 0.5);
      launch_axpy(n, h->d_b1, c, 0.5, k[1], h->stream);
      eval(h->d_b1, k[2], 0.5);
      launch_axpy(n, h->d_b1, c, 1.0, k[2], h->stream);
      eval(h->d_b1, k[3], 1.0);
      launch_rk4_final(n, c, k[0], k[1], k[2], k[3], h->stream);
      break;
    case DG_RK2MID:
      eval(c, k[0], 0.0);
      launch_axpy(n, h->d_b1, c, 0.5, k[0], h->stream);
      eval(h->d_b1, k[0], 0.5);
      launch_axpy(n, c, c, 1.0, k[0], h->stream);
      break;
    case DG_SSPRK2:
      eval(c, k[0], 0.0);
      launch_ssp(n, h->d_b1, nullptr, 0., c, 1., k[0], 1., h->stream);
      eval(h->d_b1, k[0], 1.0);
      launch_ssp(n, c, c, 0.5, h->d_b1, 0.5, k[0], 0.5, h->stream);
      break;
    case DG_SSPRK3:
      eval(c, k[0], 0.0);
      launch_ssp(n, h->d_b1, nullptr, 0., c, 1., k[0], 1., h->stream);
      eval(h->d_b1, k[0], 1.0);
      launch_ssp(n, h->d_b2, c, 0.75, h->d_b1, 0.25, k[0], 0.25, h->stream);
      eval(h->d_b2, k[0], 0.5);
      launch_ssp(n, c, c, 1. / 3., h->d_b2, 2. / 3., k[0], 2. / 3., h->stream);
      break;
  }
  launch_limit(phys, p, m, c, h->stream);
  // close the step: next dt from the limited state
  launch_lambda_init(phys, p, m, h->ph, c, h->d_st, cfl_mode, h->cfg.cfl * h->min_r, 0., 0., 0,
                     nullptr, h->stream, conv ? h->d_prev : nullptr, h->N, h->d_part);
}

// The time loop of time_integrate_rk4 / time_integrate_rk2
// (time_integrator.cu:193-433, 550-717): at most max_steps steps (< 0: no
// limit), stopping early at t_end or on convergence; the stop tests run on
// the device, so the host only reads the step state between chunks of steps.
static int run_steps(dg_handle *h, int max_steps, double fixed_dt, double t_end, double conv_tol,
                     double *t_inout, int *steps_done, double *conv_out) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  const int phys = h->cfg.physics, p = h->cfg.p;
  const int cfl_mode = fixed_dt > 0 ? 0 : 1;
  const bool conv = conv_tol >= 0.;
  const bool stops = conv || t_end < HUGE_VAL;
  const double t0 = t_inout ? *t_inout : 0.;
  MeshDev m = mesh_dev(h);
  if (conv && !h->d_prev && h->cfg.limiter) {
    int rc = alloc(h, (void **)&h->d_prev, sizeof(double) * dofs(h));
    if (rc) return rc;
  }
  int rc = write_state(h, t0, fixed_dt, t_end, conv_tol);
  if (rc) return rc;
  if (h->cfg.limiter) launch_limit(phys, p, m, h->d_c, h->stream);  // time_integrator.cu:179-184
  if (!h->cfg.limiter && traces_wanted(h)) {
    if ((rc = alloc_traces(h)) != DG_OK) return rc;
    launch_traces(phys, p, h->K, h->d_c, h->d_tr[0], h->stream);
  }
  launch_lambda_init(phys, p, m, h->ph, h->d_c, h->d_st, cfl_mode, h->cfg.cfl * h->min_r,
                     fixed_dt, t0, 1, nullptr, h->stream, nullptr, 0, h->d_part,
                     posfix_on(h) && !h->cfg.limiter ? h->d_bad : nullptr);
  refresh_bnd_fixed(h, t0);
  HIPCHK(h, hipGetLastError());
  DevState hs;
  std::memset(&hs, 0, sizeof hs);
  auto read_state = [&]() -> int {
    HIPCHK(h, hipMemcpyAsync(&hs, h->d_st, sizeof hs, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DG_OK;
  };
  if (stops && (rc = read_state())) return rc;  // t_end may already be reached
  long long launched = 0;
  const long long limit = max_steps < 0 ? (long long)1 << 62 : max_steps;
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  if (h->cfg.limiter) {
    // limit_c needs every neighbour's finished state: separate launches, one
    // step at a time when a stop test is active (these launches do not exit on halt)
    if (!h->d_k[0]) return fail(h, DG_ERR_STATE, "limiter buffers missing");
    while (launched < limit && !hs.halt && !hs.nan_flag) {
      limited_step(h, cfl_mode, conv);
      launched++;
      if (stops && (rc = read_state())) return rc;
    }
    launched *= nstages_of(h->cfg.integrator);
  } else {
    StageArgs st[4];
    const int ns = build_step(h, st, cfl_mode, conv);
    const bool use_graph = std::getenv("DG_NO_GRAPH") == nullptr;
    const int key = cfl_mode | (conv ? 2 : 0);
    if (use_graph && (limit > 1 || stops) && h->graph_key != key) {
      drop_graphs(h);
      // every step launches the same kernels with the same arguments (time,
      // dt and the stop state live on the device), so a run of steps is the
      // one-step sequence repeated
      for (int q = 0; q < kGraphLevels; q++) {
        hipGraph_t g;
        HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < (1 << q); r++)
          for (int i = 0; i < ns; i++) launch_stage_bnd(h, st[i]);
        HIPCHK(h, hipStreamEndCapture(h->stream, &g));
        hipError_t e = hipGraphInstantiate(&h->graph[q], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(h, DG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
        HIPCHK(h, hipGraphUpload(h->graph[q], h->stream));
      }
      h->graph_key = key;
    }
    const bool graph = use_graph && h->graph_key == key;
    long long steps = 0;
    while (steps < limit && !hs.halt && !hs.nan_flag) {
      // launches after a stop exit at once; chunks bound the wasted launches
      const long long chunk = stops ? std::min<long long>(limit - steps, 32) : limit - steps;
      for (long long s = 0; s < chunk;) {
        if (graph) {
          // the largest graph that fits the steps left
          int q = kGraphLevels - 1;
          while ((1ll << q) > chunk - s) q--;
          HIPCHK(h, hipGraphLaunch(h->graph[q], h->stream));
          s += 1ll << q;
        } else {
          for (int i = 0; i < ns; i++) launch_stage_bnd(h, st[i]);
          s++;
        }
      }
      steps += chunk;
      if (stops && (rc = read_state())) return rc;
    }
    launched = steps * ns;
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  if ((rc = read_state())) return rc;
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const long long nst = stops ? (long long)hs.steps * nstages_of(h->cfg.integrator) : launched;
  h->last_ms_per_stage = nst ? ms / nst : 0.;
  h->last_stages = (int)nst;
  if (t_inout) *t_inout = hs.tdone;
  if (steps_done) *steps_done = hs.steps;
  if (conv_out) *conv_out = hs.conv;
  h->last_dt = hs.dt_last;
  if (hs.nan_flag) return fail(h, DG_ERR_NAN, "Error: dt is NaN (time_integrator.cu:219-223)");
  return DG_OK;
}

extern "C" int dg_step(dg_handle *h, int nsteps, double fixed_dt, double *t_inout,
                       double *dt_out) {
  if (!h || nsteps < 0) return DG_ERR_ARG;
  int rc = run_steps(h, nsteps, fixed_dt, HUGE_VAL, -1., t_inout, nullptr, nullptr);
  if (dt_out) *dt_out = h->last_dt;
  return rc;
}

extern "C" int dg_run(dg_handle *h, int max_steps, double t_end, double conv_tol, double fixed_dt,
                      double *t_inout, int *steps_done, double *conv_out) {
  if (!h) return DG_ERR_ARG;
  if (max_steps < 0 && !(t_end < HUGE_VAL) && conv_tol < 0.)
    return fail(h, DG_ERR_ARG, "dg_run: no stop condition (max_steps, t_end or conv_tol)");
  if (t_end != t_end) return fail(h, DG_ERR_ARG, "dg_run: t_end is NaN");
  return run_steps(h, max_steps, fixed_dt, t_end, conv_tol, t_inout, steps_done, conv_out);
}

// per-variable error integral against the problem's exact state; the host
// sums the per-element values as main.cu:859-900 does
static int error_norm(dg_handle *h, int norm, double t, double *err) {
  if (!h || !err) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "error norm: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  launch_err(h->cfg.physics, h->cfg.p, norm, mesh_dev(h), h->ph, h->d_c, t, h->d_err, h->stream);
  HIPCHK(h, hipGetLastError());
  std::vector<double> e((size_t)h->K * h->N);
  HIPCHK(h, hipMemcpyAsync(e.data(), h->d_err, sizeof(double) * e.size(), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int n = 0; n < h->N; n++) {
    double tot = 0.;
    for (int i = 0; i < h->K; i++) tot += e[(size_t)n * h->K + i];
    err[n] = norm == 2 ? std::sqrt(tot) : tot;
  }
  return DG_OK;
}

extern "C" int dg_l2_error(dg_handle *h, double t, double *err) { return error_norm(h, 2, t, err); }

extern "C" int dg_l1_error(dg_handle *h, double t, double *err) { return error_norm(h, 1, t, err); }

extern "C" int dg_eval_exact(dg_handle *h, int n, double t, double *u1, double *u2, double *u3) {
  if (!h || !u1 || !u2 || !u3 || n < 0) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_eval_exact: no mesh");
  if (n >= h->N) return fail(h, DG_ERR_ARG, "dg_eval_exact: variable %d of %d", n, h->N);
  HIPCHK(h, hipSetDevice(h->cfg.device));
  double *d = nullptr;
  HIPCHK(h, hipMalloc((void **)&d, sizeof(double) * 3 * h->K));
  launch_exact_vertex(h->cfg.physics, mesh_dev(h), h->ph, t, n, d, d + h->K, d + 2 * h->K,
                      h->stream);
  hipError_t e = hipGetLastError();
  const size_t kb = sizeof(double) * h->K;
  if (e == hipSuccess) e = hipMemcpyAsync(u1, d, kb, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(u2, d + h->K, kb, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(u3, d + 2 * h->K, kb, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(h, DG_ERR_HIP, "dg_eval_exact: %s", hipGetErrorString(e));
  return DG_OK;
}

extern "C" int dg_eval_u(dg_handle *h, int n, double *u1, double *u2, double *u3) {
  if (!h || !u1 || !u2 || !u3 || n < 0) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_eval_u: no mesh");
  if (n >= h->N) return fail(h, DG_ERR_ARG, "dg_eval_u: variable %d of %d", n, h->N);
  HIPCHK(h, hipSetDevice(h->cfg.device));
  double *d = h->d_err;  // N*K >= 3*K is not guaranteed; use lam + err
  std::vector<double> tmp;
  double *o1 = h->d_lam, *o2 = d, *o3 = d + h->K;
  double *extra = nullptr;
  if (h->N < 2) {
    HIPCHK(h, hipMalloc((void **)&extra, sizeof(double) * 2 * h->K));
    o2 = extra;
    o3 = extra + h->K;
  }
  launch_eval_u(h->cfg.physics, h->cfg.p, h->K, h->d_c, n, o1, o2, o3, h->stream);
  HIPCHK(h, hipGetLastError());
  const size_t kb = sizeof(double) * h->K;
  HIPCHK(h, hipMemcpyAsync(u1, o1, kb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(u2, o2, kb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(u3, o3, kb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (extra) (void)hipFree(extra);
  return DG_OK;
}

extern "C" int dg_lambda(dg_handle *h, double *lam) {
  if (!h || !lam) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_lambda: no mesh");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  launch_lambda_init(h->cfg.physics, h->cfg.p, mesh_dev(h), h->ph, h->d_c, nullptr, 0, 0., 0., 0.,
                     0, h->d_lam, h->stream, nullptr, 0, nullptr);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(lam, h->d_lam, sizeof(double) * h->K, hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

extern "C" int dg_limit(dg_handle *h) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_limit: no mesh");
  if (h->cfg.p != 1) return fail(h, DG_ERR_ARG, "Error: limiter only enabled for p = 1");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  launch_limit(h->cfg.physics, h->cfg.p, mesh_dev(h), h->d_c, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

extern "C" int dg_info(dg_handle *h, dg_info_t *info) {
  if (!h || !info) return DG_ERR_ARG;
  info->K = h->K;
  info->S = h->S;
  info->N = h->N;
  info->n_p = h->np;
  info->n_q = h->nq;
  info->n_q1 = h->nq1;
  info->dofs = (long long)dofs(h);
  info->min_r = h->min_r;
  (void)hipSetDevice(h->cfg.device);
  info->vgprs_stage = stage_vgprs(h->cfg.physics, h->cfg.p, volume_mfma(h->cfg));
  info->adj_fix_fired = h->adj_fix_fired;
  info->adj_resident = h->adj_resident;
  return DG_OK;
}

extern "C" int dg_geometry(dg_handle *h, double *J, double *len, double *nx, double *ny,
                           double *diam) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_geometry: no mesh");
  if (J) std::memcpy(J, h->hJ.data(), sizeof(double) * h->K);
  if (diam) std::memcpy(diam, h->hdiam.data(), sizeof(double) * h->K);
  if (len) std::memcpy(len, h->hlen.data(), sizeof(double) * h->S);
  if (nx) std::memcpy(nx, h->hnx.data(), sizeof(double) * h->S);
  if (ny) std::memcpy(ny, h->hny.data(), sizeof(double) * h->S);
  return DG_OK;
}

extern "C" int dg_sync(dg_handle *h) {
  if (!h) return DG_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

extern "C" void *dg_stream(dg_handle *h) { return h ? (void *)h->stream : nullptr; }

extern "C" int dg_last_timing(dg_handle *h, double *ms, int *stages) {
  if (!h) return DG_ERR_ARG;
  if (ms) *ms = h->last_ms_per_stage;
  if (stages) *stages = h->last_stages;
  return DG_OK;
}

// ------------------------------------------------------------ adjoint
//
// Discrete adjoint of dg_step with a fixed dt (no counterpart in the
// reference).  The step is written in Butcher form
//   u_s = u^n + sum_{r<s} a_sr k_r,  u^{n+1} = u^n + sum_s b_s k_s,  k_s = k(u_s)
// (RK4; explicit midpoint; SSP-RK2: a10 = 1, b = 1/2, 1/2; SSP-RK3: a10 = 1,
// a20 = a21 = 1/4, b = 1/6, 1/6, 2/3), so with g the adjoint of u^{n+1}
//   gk_s = b_s g + sum_{r>s} a_rs l_r,  l_s = (dk/du(u_s))^T gk_s,
//   adjoint of u^n = g + sum_s l_s.
// Checkpoints of u^n are kept every `every` steps; the reverse sweep re-runs
// each segment forward from its checkpoint (same fused stage kernels)
// keeping every stage input, then walks the segment backwards.

namespace {
struct Butcher {
  int S;
  double a[4][4];
  double b[4];
  double c[4];
};

Butcher butcher_of(int integ) {
  Butcher t;
  std::memset(&t, 0, sizeof t);
  switch (integ) {
    case DG_RK4:
      t.S = 4;
      t.a[1][0] = 0.5; t.a[2][1] = 0.5; t.a[3][2] = 1.0;
      t.b[0] = 1. / 6; t.b[1] = 1. / 3; t.b[2] = 1. / 3; t.b[3] = 1. / 6;
      t.c[1] = 0.5; t.c[2] = 0.5; t.c[3] = 1.0;
      break;
    case DG_RK2MID:
      t.S = 2;
      t.a[1][0] = 0.5; t.b[1] = 1.0; t.c[1] = 0.5;
      break;
    case DG_SSPRK2:
      t.S = 2;
      t.a[1][0] = 1.0; t.b[0] = 0.5; t.b[1] = 0.5; t.c[1] = 1.0;
      break;
    default:
      t.S = 3;
      t.a[1][0] = 1.0; t.a[2][0] = 0.25; t.a[2][1] = 0.25;
      t.b[0] = 1. / 6; t.b[1] = 1. / 6; t.b[2] = 2. / 3;
      t.c[1] = 1.0; t.c[2] = 0.5;
      break;
  }
  return t;
}
}  // namespace


// Device buffers of the adjoint: each adj_alloc call makes one allocation
// and hands out its buffers at 256-byte-aligned offsets, each 8 KiB past a
// 256-byte multiple of the one before (one hipMalloc per group rather than
// per buffer: a sweep may hold thousands of buffers).  With the 8 KiB pad the
// config-5 sweep ran 1.8 % faster than without (2, 4 and 16 KiB: 1.3-1.5 %);
// the forward's state buffers (block_views) measured no gain from a pad.
struct AdjBuffers {
  std::vector<double *> bufs;
  std::vector<void *> blocks;
  AdjBuffers() = default;
  AdjBuffers(const AdjBuffers &) = delete;
  AdjBuffers &operator=(const AdjBuffers &) = delete;
  ~AdjBuffers() {
    for (void *p : blocks) (void)hipFree(p);
  }
};

static int adj_alloc(dg_handle *h, AdjBuffers &ab, size_t count, size_t bytes) {
  if (!count) return DG_OK;
  const size_t stride = (bytes + 255) / 256 * 256 + 8192;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, stride * count);
  if (e != hipSuccess)
    return fail(h, DG_ERR_HIP, "adjoint: hipMalloc of %zu buffers of %zu bytes failed (%s)", count,
                bytes, hipGetErrorString(e));
  ab.blocks.push_back(p);
  for (size_t q = 0; q < count; q++) ab.bufs.push_back((double *)((char *)p + q * stride));
  return DG_OK;
}

// How many of a checkpointed solve's segments keep the stage data of the
// forward pass itself (the last ones), so that the reverse sweep walks them
// without recomputing.  Off unless DG_ADJ_STORE_GB gives a memory budget in
// GB: then as many as it and the free device memory hold, less a margin (5 %
// of the device or 2 GiB, whichever is more), after the checkpoints and the
// recompute buffers that the other segments need.  (Off by default: fresh
// allocations of that size cost more than the recompute they save, e.g.
// 271 GB for 54 of config 4's 125 segments took 2.2 s to allocate and free
// against 0.23 s of recompute saved.)  seg_bytes(ns) is one segment's stage
// data for ns steps.
template <class F>
static int resident_segments(int nsteps, int every, int nseg, size_t sb, const F &seg_bytes) {
  const char *env = std::getenv("DG_ADJ_STORE_GB");
  if (!env) return 0;
  char *end = nullptr;
  const double gb = std::strtod(env, &end);
  if (end == env || !(gb > 0.)) return 0;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const size_t margin = std::max(tot / 20, (size_t)2 << 30);
  size_t budget = fr > margin ? fr - margin : 0;
  const double cap = gb * 1e9;  // compared as a double: no overflow for a large setting
  if (cap < (double)budget) budget = (size_t)cap;
  auto len = [&](int q) { return std::min(every, nsteps - q * every); };
  size_t all = 0;
  for (int q = 0; q < nseg; q++) all += seg_bytes(len(q));
  if (all <= budget) return nseg;  // no checkpoints, no recompute
  // segment 0 at least is recomputed: the recompute buffers and a checkpoint
  // per recomputed segment; each resident segment saves its checkpoint
  size_t used = seg_bytes(every) + (size_t)nseg * sb;
  int r = 0;
  for (int q = nseg - 1; q > 0; q--) {
    const size_t add = seg_bytes(len(q)) - sb;
    if (used + add > budget) break;
    used += add;
    r++;
  }
  return r;
}

extern "C" int dg_adjoint_set_objective(dg_handle *h, int kind, const double *params, int nparams,
                                        const double *g_weights, int where) {
  if (!h) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_adjoint_set_objective: no mesh");
  if (kind < DG_OBJ_LINEAR || kind > DG_OBJ_PATCH_SQ)
    return fail(h, DG_ERR_ARG, "dg_adjoint_set_objective: unknown objective %d", kind);
  if (kind == DG_OBJ_ENTROPY && h->cfg.physics != DG_EULER)
    return fail(h, DG_ERR_ARG, "entropy objective needs Euler");
  if (kind == DG_OBJ_PATCH_SQ && nparams < 5)
    return fail(h, DG_ERR_ARG, "patch objective needs x0, x1, y0, y1, variable");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  h->obj_kind = kind;
  h->obj_params.assign(8, 0.);
  for (int i = 0; i < nparams && i < 8; i++) h->obj_params[i] = params[i];
  if (kind == DG_OBJ_LINEAR) {
    if (!g_weights) return fail(h, DG_ERR_ARG, "linear objective needs weights");
    int rc = alloc(h, (void **)&h->d_gw, sizeof(double) * dofs(h));
    if (rc) return rc;
    HIPCHK(h, hipMemcpy(h->d_gw, g_weights, sizeof(double) * dofs(h), kind_to(where)));
  }
  return DG_OK;
}

static bool adjoint_supported(const dg_handle *h) { return !h->cfg.limiter; }

// objective value of d_c and its gradient into gout (device)
static int eval_objective(dg_handle *h, double *J, double *gout) {
  const size_t n = dofs(h);
  if (h->obj_kind == DG_OBJ_LINEAR) {
    std::vector<double> c(n), g(n);
    HIPCHK(h, hipMemcpyAsync(c.data(), h->d_c, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(g.data(), h->d_gw, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(gout, h->d_gw, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double s = 0.;
    for (size_t i = 0; i < n; i++) s += g[i] * c[i];
    *J = s;
    return DG_OK;
  }
  double *prm = nullptr;
  HIPCHK(h, hipMalloc((void **)&prm, 8 * sizeof(double)));
  HIPCHK(h, hipMemcpyAsync(prm, h->obj_params.data(), 8 * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  launch_objective(h->cfg.physics, h->cfg.p, h->obj_kind, mesh_dev(h), h->ph, h->d_c, prm,
                   h->d_lam, gout, h->stream);
  std::vector<double> v(h->K);
  HIPCHK(h, hipMemcpyAsync(v.data(), h->d_lam, sizeof(double) * h->K, hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  (void)hipFree(prm);
  double s = 0.;
  for (int e = 0; e < h->K; e++) s += v[e];
  *J = s;
  return DG_OK;
}

static AdjArgs adj_args(dg_handle *h, const double *us, const double *lam, double *gout,
                        double *geps, double dt) {
  AdjArgs a;
  std::memset(&a, 0, sizeof a);
  a.m = mesh_dev(h);
  a.us = us;
  a.lam = lam;
  a.lc = 1.0;
  a.gout = gout;
  a.nout = 0;
  a.geps = geps;
  a.ph = h->ph;
  a.dt = dt;
  a.scale = 1.0;
  return a;
}

extern "C" int dg_adjoint_rhs(dg_handle *h, const double *u, double t, double dt, const double *gk,
                              double *gu, int where) {
  if (!h || !u || !gk || !gu) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_adjoint_rhs: no mesh");
  if (!adjoint_supported(h))
    return fail(h, DG_ERR_ARG,
                "adjoint: the limiter is not differentiated");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  const size_t sb = sizeof(double) * dofs(h);
  AdjBuffers ab;
  int rc = adj_alloc(h, ab, 1, sb);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(h->d_b1, u, sb, kind_to(where), h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_b2, gk, sb, kind_to(where), h->stream));
  refresh_bnd_fixed(h, t);
  launch_adjoint_rhs(h->cfg.physics, h->cfg.p, adj_args(h, h->d_b1, h->d_b2, ab.bufs[0], nullptr, dt),
                     h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(gu, ab.bufs[0], sb, kind_from(where), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DG_OK;
}

extern "C" int dg_adjoint_solve(dg_handle *h, int nsteps, double dt, int every, double *J,
                                double *grad_u0, double *grad_eps, int where) {
  if (!h || nsteps <= 0 || !(dt > 0) || !J) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_adjoint_solve: no mesh");
  if (h->obj_kind < 0) return fail(h, DG_ERR_STATE, "dg_adjoint_solve: no objective set");
  if (!adjoint_supported(h))
    return fail(h, DG_ERR_ARG,
                "adjoint: the limiter is not differentiated");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  if (every <= 0) every = (int)std::ceil(std::sqrt((double)nsteps));
  if (every > nsteps) every = nsteps;
  const Butcher T = butcher_of(h->cfg.integrator);
  const int S = T.S;
  const size_t n = dofs(h), sb = sizeof(double) * n;
  const int nseg = (nsteps + every - 1) / every;
  const int phys = h->cfg.physics, p = h->cfg.p;
  const bool te = phys == DG_MAXWELL_TE;
  // buffers: adjoint work; checkpoints of the recomputed segments and one
  // segment's stage inputs (and, for the TE eps gradient, the E rows of the
  // stages' k) for their recompute; the resident segments' own stage data
  const bool want_eps = te && grad_eps;
  const size_t kst_b = sizeof(double) * 2 * h->np * h->K;
  AdjBuffers ck, store, kst, wk;
  int rc = adj_alloc(h, wk, S + 3, sb);
  if (rc) return rc;
  if (traces_wanted(h) && (rc = alloc_traces(h)) != DG_OK) return rc;
  const int nres = resident_segments(nsteps, every, nseg, sb, [&](int ns) {
    return ((size_t)ns * S + 1) * sb + (want_eps ? (size_t)ns * S * kst_b : 0);
  });
  const int nrec = nseg - nres;  // segments 0 .. nrec - 1 are recomputed
  h->adj_resident = nres;
  if (nrec > 0) {
    rc = adj_alloc(h, ck, nrec, sb);
    if (!rc) rc = adj_alloc(h, store, (size_t)every * S + 1, sb);
    if (!rc && want_eps) rc = adj_alloc(h, kst, (size_t)every * S, kst_b);
    if (rc) return rc;
  }
  std::vector<AdjBuffers> rstore(nres), rkst(nres);
  for (int j = 0; j < nres; j++) {
    const int ns = std::min(every, nsteps - (nrec + j) * every);
    rc = adj_alloc(h, rstore[j], (size_t)ns * S + 1, sb);
    if (!rc && want_eps) rc = adj_alloc(h, rkst[j], (size_t)ns * S, kst_b);
    if (rc) return rc;
  }
  // (each buffer's owner is set before the next call that can return early)
  struct Free {
    double *p;
    ~Free() {
      if (p) (void)hipFree(p);
    }
  } fgeps{nullptr}, fgzs{nullptr};
  double *d_geps = nullptr, *d_gzs = nullptr;
  if (want_eps) {
    HIPCHK(h, hipMalloc((void **)&d_geps, sizeof(double) * h->K));
    fgeps.p = d_geps;
    HIPCHK(h, hipMalloc((void **)&d_gzs, 2 * sizeof(double) * h->S));
    fgzs.p = d_gzs;
    HIPCHK(h, hipMemsetAsync(d_geps, 0, sizeof(double) * h->K, h->stream));
  }
  double *G = wk.bufs[S];
  double **L = wk.bufs.data();  // l_0 .. l_{S-1}
  // the positivity fix is taken as inactive: watch whether it fires anywhere
  // in the forward map or its recompute (stage launches captured before the
  // watch existed carry no pointer to it, so their graphs go)
  if (posfix_on(h)) {
    if (!h->d_fired) {
      if ((rc = alloc(h, (void **)&h->d_fired, sizeof(int))) != DG_OK) return rc;
      drop_graphs(h);
    }
    HIPCHK(h, hipMemsetAsync(h->d_fired, 0, sizeof(int), h->stream));
  }
  h->adj_fix_fired = 0;

  const bool tdep = bnd_time_dependent(h);
  // ---- the launches of the reverse sweep (run after the forward pass
  // below).  Per step, stage r (last to first) runs one VJP on
  // the single array gk_r; the launch for stage r >= 1 also writes the next
  // weight, gk_{r-1} = a_{r,r-1} l_r + b_{r-1} g + sum_{r' > r} a_{r',r-1} l_r'
  // (own-element terms), and the launch for stage 0 writes the new g,
  // g + sum_r l_r, into the other g buffer.  The last stage reads b_{S-1} g.
  double *Gn = wk.bufs[S + 1];
  double *GKb = L[0];  // gk_{r-1} (L[0] itself is never stored: stage 0 writes g)
  double *GKa = wk.bufs[S + 2];
  // One segment's launches: re-run it from its first state (in stb[0]: the
  // checkpoint, or for a resident segment the forward pass's state) keeping
  // every stage input (record), then walk it backwards (walk).  Without
  // time-dependent boundary data the launches of a segment depend only on
  // its length and on which g buffer it starts from, so they are captured
  // once as a graph and replayed.
  // a tableau with only the subdiagonal (RK4, midpoint RK2, SSP-RK2) lets
  // the stage adjoints l_r be recovered from the weights w_{r-1} that are
  // written anyway, so no l_r array is stored
  bool subdiag = true;
  for (int r = 1; r < S; r++) {
    if (T.a[r][r - 1] == 0.) subdiag = false;
    for (int c = 0; c + 1 < r; c++)
      if (T.a[r][c] != 0.) subdiag = false;
  }
  auto record = [&](int ns, AdjBuffers &stb, AdjBuffers &ksb) -> int {
    // positivity flags of the checkpoint; each step's last stage writes
    // those of the state it stores, as in the forward solve
    if (posfix_on(h))
      launch_flags(phys, p, h->K, h->ph, stb.bufs[0], h->d_bad, h->stream, h->d_fired);
    // stored traces, as in the forward solve: the stage-buffer trace arrays
    // (idle during the sweep) alternate between a stage's input and output
    double *tr_cur = nullptr, *tr_nxt = nullptr;
    if (traces_wanted(h) && h->d_tr[1] && h->d_tr[2]) {
      tr_cur = h->d_tr[1];
      tr_nxt = h->d_tr[2];
      launch_traces(phys, p, h->K, stb.bufs[0], tr_cur, h->stream);
    }
    for (int s = 0; s < ns; s++) {
      double *io[5];
      for (int r = 0; r < S; r++) io[r] = stb.bufs[(size_t)s * S + r];
      io[S] = stb.bufs[(size_t)(s + 1) * S];
      StageArgs st[4];
      build_step_io(h, st, 0, io, h->d_acc ? h->d_acc : h->d_b2);
      for (int r = 0; r < S; r++) {
        if (posfix_on(h)) {
          st[r].flags_in = 1;
          st[r].flags_out = 1;
        }
        if (want_eps) st[r].kout = ksb.bufs[(size_t)s * S + r];
        if (tr_cur) {
          st[r].tr_in = tr_cur;
          st[r].tr_out = tr_nxt;
          std::swap(tr_cur, tr_nxt);
        }
        launch_stage_bnd(h, st[r]);
      }
    }
    HIPCHK(h, hipGetLastError());
    return DG_OK;
  };
  auto walk = [&](int ns, double tseg, AdjBuffers &stb, AdjBuffers &ksb) -> int {
    // walk the segment backwards
    for (int s = ns - 1; s >= 0; s--) {
      const double tn = tseg + s * dt;
      const double *lam = G;
      double lcoef = T.b[S - 1];
      for (int r = S - 1; r >= 0; r--) {
        AdjArgs aa = adj_args(h, stb.bufs[(size_t)s * S + r], lam, r == 0 ? Gn : L[r], d_geps, dt);
        aa.lc = lcoef;
        if (want_eps) {
          aa.kst = ksb.bufs[(size_t)s * S + r];
          aa.gzs = d_gzs;
        }
        double *next = (r & 1) ? GKa : GKb;  // never the array this launch reads
        if (subdiag) {
          // w_{r-1} = b_{r-1} g + a_{r,r-1} l_r goes to L[r]; no l_r array
          next = L[r];
          aa.gout = r == 0 ? Gn : nullptr;
          if (r >= 1) {
            aa.gout2 = next;
            aa.c2g = T.a[r][r - 1];
            aa.n2 = 0;
            if (T.b[r - 1] != 0.) {
              aa.o2t[0] = G;
              aa.o2c[0] = T.b[r - 1];
              aa.n2 = 1;
            }
          } else {
            // g <- g + sum_r l_r with l_{r+1} = (w_r - b_r g) / a_{r+1,r}
            double cg = 1.;
            for (int r2 = 0; r2 + 1 < S; r2++) cg -= T.b[r2] / T.a[r2 + 1][r2];
            aa.ot[0] = G;
            aa.oc[0] = cg;
            for (int r2 = 0; r2 + 1 < S; r2++) {
              aa.ot[r2 + 1] = L[r2 + 1];
              aa.oc[r2 + 1] = 1. / T.a[r2 + 1][r2];
            }
            aa.nout = S;
          }
        } else if (r >= 1) {
          aa.gout2 = next;
          aa.c2g = T.a[r][r - 1];
          int n2 = 0;
          if (T.b[r - 1] != 0.) {
            aa.o2t[n2] = G;
            aa.o2c[n2++] = T.b[r - 1];
          }
          for (int r2 = r + 1; r2 < S; r2++)
            if (T.a[r2][r - 1] != 0.) {
              aa.o2t[n2] = L[r2];
              aa.o2c[n2++] = T.a[r2][r - 1];
            }
          aa.n2 = n2;
        } else {  // g <- g + sum_r l_r
          aa.ot[0] = G;
          aa.oc[0] = 1.;
          for (int r2 = 1; r2 < S; r2++) {
            aa.ot[r2] = L[r2];
            aa.oc[r2] = 1.;
          }
          aa.nout = S;
        }
        if (tdep)
          launch_bnd(phys, p, mesh_dev(h), h->ph, h->d_st, 0., 0, tn + dt + T.c[r] * dt, h->stream);
        launch_adjoint_rhs(phys, p, aa, h->stream);
        lam = next;
        lcoef = 1.0;
      }
      std::swap(G, Gn);
    }
    return DG_OK;
  };
  auto segment = [&](int ns, double tseg) -> int {
    const int r = record(ns, store, kst);
    return r ? r : walk(ns, tseg, store, kst);
  };

  // ---- forward with checkpoints (wall clock; dg_step synchronises per segment)
  HIPCHK(h, hipStreamSynchronize(h->stream));
  auto w0 = std::chrono::steady_clock::now();
  double t = 0.;
  for (int q = 0; q < nseg; q++) {
    const int ns = std::min(every, nsteps - q * every);
    if (q < nrec) {
      HIPCHK(h, hipMemcpyAsync(ck.bufs[q], h->d_c, sb, hipMemcpyDeviceToDevice, h->stream));
      rc = dg_step(h, ns, dt, &t, nullptr);
      if (rc) return rc;
      continue;
    }
    // a resident segment: the recompute's launches run here, from the
    // current state, and their stage data stay for the sweep
    const double tseg = q * every * dt;
    AdjBuffers &rs = rstore[q - nrec];
    HIPCHK(h, hipMemcpyAsync(rs.bufs[0], h->d_c, sb, hipMemcpyDeviceToDevice, h->stream));
    if ((rc = write_state(h, tseg + dt, dt)) != DG_OK) return rc;
    refresh_bnd_fixed(h, tseg);
    if ((rc = record(ns, rs, rkst[q - nrec])) != DG_OK) return rc;
    HIPCHK(h, hipMemcpyAsync(h->d_c, rs.bufs[(size_t)ns * S], sb, hipMemcpyDeviceToDevice,
                             h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  auto w1 = std::chrono::steady_clock::now();
  h->adj_fwd_ms = std::chrono::duration<double, std::milli>(w1 - w0).count();

  // ---- objective and its gradient at the final state
  rc = eval_objective(h, J, G);
  if (rc) return rc;
  struct Graphs {
    std::vector<std::pair<std::pair<int, double *>, hipGraphExec_t>> v;
    ~Graphs() {
      for (auto &g : v) (void)hipGraphExecDestroy(g.second);
    }
  } graphs;
  if (!tdep) refresh_bnd_fixed(h, 0.);
  for (int q = nseg - 1; q >= 0; q--) {
    const int ns = std::min(every, nsteps - q * every);
    const double tseg = q * every * dt;
    if (q >= nrec) {  // resident: the stage data are there; walk it (not captured)
      if ((rc = walk(ns, tseg, rstore[q - nrec], rkst[q - nrec])) != DG_OK) return rc;
      HIPCHK(h, hipGetLastError());
      continue;
    }
    HIPCHK(h, hipMemcpyAsync(store.bufs[0], ck.bufs[q], sb, hipMemcpyDeviceToDevice, h->stream));
    rc = write_state(h, tseg + dt, dt);
    if (rc) return rc;
    if (tdep) {
      refresh_bnd_fixed(h, tseg);
      if ((rc = segment(ns, tseg))) return rc;
      HIPCHK(h, hipGetLastError());
      continue;
    }
    const std::pair<int, double *> key(ns, G);
    hipGraphExec_t ex = nullptr;
    for (auto &g : graphs.v)
      if (g.first == key) ex = g.second;
    if (ex) {
      HIPCHK(h, hipGraphLaunch(ex, h->stream));
      if (ns & 1) std::swap(G, Gn);
    } else {
      hipGraph_t gr;
      HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      rc = segment(ns, tseg);
      hipError_t e = hipStreamEndCapture(h->stream, &gr);
      if (rc) return rc;
      if (e != hipSuccess) return fail(h, DG_ERR_HIP, "adjoint capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) return fail(h, DG_ERR_HIP, "adjoint graph: %s", hipGetErrorString(e));
      graphs.v.push_back({key, ex});
      HIPCHK(h, hipGraphLaunch(ex, h->stream));
    }
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  auto w2 = std::chrono::steady_clock::now();
  h->adj_rev_ms = std::chrono::duration<double, std::milli>(w2 - w1).count();
  if (grad_u0)
    HIPCHK(h, hipMemcpyAsync(grad_u0, G, sb, kind_from(where), h->stream));
  if (want_eps)
    HIPCHK(h, hipMemcpyAsync(grad_eps, d_geps, sizeof(double) * h->K, kind_from(where), h->stream));
  int fired = 0;
  if (posfix_on(h) && h->d_fired)
    HIPCHK(h, hipMemcpyAsync(&fired, h->d_fired, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->adj_fix_fired = fired ? 1 : 0;
  return DG_OK;
}

extern "C" int dg_adjoint_objective(dg_handle *h, double *J) {
  if (!h || !J) return DG_ERR_ARG;
  if (!h->K) return fail(h, DG_ERR_STATE, "dg_adjoint_objective: no mesh");
  if (h->obj_kind < 0) return fail(h, DG_ERR_STATE, "dg_adjoint_objective: no objective set");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  AdjBuffers ab;
  int rc = adj_alloc(h, ab, 1, sizeof(double) * dofs(h));
  if (rc) return rc;
  return eval_objective(h, J, ab.bufs[0]);
}

extern "C" int dg_adjoint_timing(dg_handle *h, double *fwd_ms, double *rev_ms) {
  if (!h) return DG_ERR_ARG;
  if (fwd_ms) *fwd_ms = h->adj_fwd_ms;
  if (rev_ms) *rev_ms = h->adj_rev_ms;
  return DG_OK;
}

// ------------------------------------------------------------ ensembles
//
// No counterpart in the reference (one mesh per process).  Each instance is
// first set up as a handle of its own (mesh precompute, boundary data,
// initial projection); dg_ens_build then merges the instances of each
// polynomial order into one group (ranges padded to 256 elements, indices
// offset) and a stage of the whole ensemble becomes one launch
// (k_stage_ens) followed by one step-closing launch for all instances.

static constexpr int kEnsPad = 256;
static const int kEnsB[3] = {64, 128, 256};

// a stage launch layout: merged (one launch, every group in blocks of
// kEnsB[bi[0]]) or split (a launch per order group, group g in blocks of
// kEnsB[bi[g]])
struct EnsCfg {
  bool split = false;
  int bi[kEnsGroups] = {0, 0, 0, 0, 0, 0};
  int key(int cfl_mode) const {
    int k = cfl_mode * 2 + (split ? 1 : 0);
    for (int g = 0; g < kEnsGroups; g++) k = k * 4 + bi[g];
    return k;
  }
};
static EnsCfg ens_uniform(int bi, bool split) {
  EnsCfg c;
  c.split = split;
  for (int &b : c.bi) b = bi;
  return c;
}

struct dg_ensemble {
  dg_config cfg{};
  PhysParams ph{};
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<dg_handle *> inst;
  std::vector<int> inst_off;  // element offset of each instance in its group
  bool built = false;
  int N = 0, nstage = 0;
  struct Group {
    int p = -1, K = 0, S = 0, nb = 0;  // padded elements, sides, boundary slots
    std::vector<int> members;
    double *xr = nullptr, *yr = nullptr, *xs = nullptr, *ys = nullptr, *J = nullptr,
           *V = nullptr, *snx = nullptr, *sny = nullptr, *slen = nullptr, *eps = nullptr,
           *bnd_val = nullptr, *c = nullptr, *b1 = nullptr, *b2 = nullptr, *acc = nullptr;
    double *state = nullptr;  // one allocation holding c, b1, b2 (and acc)
    int2 *slot = nullptr;
    int4 *side = nullptr;
    int *bnd_elem = nullptr, *bnd_code = nullptr, *bad = nullptr;
    int *bnd_inst = nullptr;  // instance of each boundary slot
    unsigned long long *part = nullptr;
    int4 *blk[3] = {nullptr, nullptr, nullptr};  // per logical block, for each block size
  } grp[kEnsGroups];
  DevState *d_st = nullptr;
  EnsInst *d_inst = nullptr;
  PhysParams *d_ph = nullptr;  // each instance's parameters (time-dependent boundary data)
  bool tdep = false;           // some instance's boundary data depend on time
  int B = 0;
  bool split = false;  // a stage launch per order group instead of one for all
  int Bg[kEnsGroups] = {0, 0, 0, 0, 0, 0};  // split: each group's block size index
  double tune_best_ms = -1.;                 // the tuned layout's ms per step
  double tune_ms[3] = {0., 0., 0.};
  double tune_split_ms[3] = {0., 0., 0.};
  hipGraphExec_t graph = nullptr;
  int graph_key = -1;
  double last_ms_per_step = 0.;
};

static int efail(dg_ensemble *E, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (E) E->err = buf;
  return code;
}

#define EHIPCHK(E, x)                                                                 \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return efail(E, DG_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));        \
  } while (0)

template <typename T>
static int ealloc(dg_ensemble *E, T *&p, size_t count) {
  dfree(p);
  hipError_t e = hipMalloc((void **)&p, sizeof(T) * (count ? count : 1));
  if (e != hipSuccess)
    return efail(E, DG_ERR_HIP, "hipMalloc(%zu) failed: %s", sizeof(T) * count,
                 hipGetErrorString(e));
  return DG_OK;
}

static int ens_lds_bytes(int phys, int B) {
  const int N = nvars_of(phys);
  int m = 0;
  for (int p = 0; p <= 3; p++) m = std::max(m, 3 * np_of(p) * (p + 1) + N * np_of(p) * B);
  return 8 * m;
}

extern "C" const char *dg_ens_last_error(const dg_ensemble *E) {
  return E ? E->err.c_str() : "null ensemble";
}

extern "C" void dg_ens_destroy(dg_ensemble *E) {
  if (!E) return;
  (void)hipSetDevice(E->cfg.device);
  if (E->stream) (void)hipStreamSynchronize(E->stream);
  for (dg_handle *h : E->inst) dg_destroy(h);
  for (auto &g : E->grp) {
    dfree(g.xr); dfree(g.yr); dfree(g.xs); dfree(g.ys); dfree(g.J); dfree(g.V);
    dfree(g.snx); dfree(g.sny); dfree(g.slen); dfree(g.eps); dfree(g.bnd_val);
    dfree(g.state);
    g.c = g.b1 = g.b2 = g.acc = nullptr;
    dfree(g.slot); dfree(g.side);
    dfree(g.bnd_elem); dfree(g.bnd_code); dfree(g.bad); dfree(g.part); dfree(g.bnd_inst);
    for (auto &b : g.blk) dfree(b);
  }
  dfree(E->d_st);
  dfree(E->d_inst);
  dfree(E->d_ph);
  if (E->graph) (void)hipGraphExecDestroy(E->graph);
  if (E->ev0) (void)hipEventDestroy(E->ev0);
  if (E->ev1) (void)hipEventDestroy(E->ev1);
  if (E->stream) (void)hipStreamDestroy(E->stream);
  delete E;
}

extern "C" int dg_ens_create(dg_ensemble **out, const dg_config *cfg) {
  if (!out || !cfg) return DG_ERR_ARG;
  *out = nullptr;
  if (cfg->physics < 0 || cfg->physics > 4) return DG_ERR_ARG;
  if (cfg->integrator < DG_RK4 || cfg->integrator > DG_SSPRK3) return DG_ERR_ARG;
  if (cfg->limiter) return DG_ERR_ARG;  // the limiter needs its own launches
  if (cfg->flux == DG_FLUX_ROE && cfg->physics != DG_SWE) return DG_ERR_ARG;
  dg_ensemble *E = new dg_ensemble();
  E->cfg = *cfg;
  if (E->cfg.cfl <= 0) E->cfg.cfl = (cfg->integrator == DG_RK2MID) ? 0.7 : 0.5;
  E->N = nvars_of(cfg->physics);
  E->nstage = cfg->integrator == DG_RK4 ? 4 : (cfg->integrator == DG_SSPRK3 ? 3 : 2);
  if (hipSetDevice(cfg->device) != hipSuccess ||
      hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess) {
    dg_ens_destroy(E);
    return DG_ERR_HIP;
  }
  *out = E;
  return DG_OK;
}

extern "C" int dg_ens_add(dg_ensemble *E, int p, const double *problem_params, int K,
                          const double *V, const int *es, int S, const double *sxy,
                          const int *left, const int *right, const int *lsn, const int *rsn,
                          int *index) {
  if (!E) return DG_ERR_ARG;
  if (E->built) return efail(E, DG_ERR_STATE, "dg_ens_add: the ensemble is already built");
  if (p < 0 || p > 5) return efail(E, DG_ERR_ARG, "dg_ens_add: order %d (ensembles take 0..5)", p);
  dg_config c = E->cfg;
  c.p = p;
  if (problem_params)
    for (int i = 0; i < 8; i++) c.problem_params[i] = problem_params[i];
  dg_handle *h = nullptr;
  int rc = dg_create(&h, &c);
  if (rc) return efail(E, rc, "dg_ens_add: instance configuration rejected");
  rc = dg_set_mesh(h, K, V, es, S, sxy, left, right, lsn, rsn);
  if (!rc) rc = dg_project_ic(h);
  if (!rc) {
    refresh_bnd_fixed(h, 0.);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(h, DG_ERR_HIP, "boundary data: %s", hipGetErrorString(e));
  }
  if (rc) {
    efail(E, rc, "dg_ens_add: %s", h->err.c_str());
    dg_destroy(h);
    return rc;
  }
  if (index) *index = (int)E->inst.size();
  E->inst.push_back(h);
  return DG_OK;
}

// the per-stage launch arguments of the ensemble at block size B
static void ens_stage_args(dg_ensemble *E, const EnsCfg &c, EnsStageArgs *ea, int *total,
                           int cfl_mode) {
  const int pf = (E->cfg.positivity && (E->cfg.physics == DG_EULER || E->cfg.physics == DG_SWE))
                     ? 1 : 0;
  std::memset(ea, 0, sizeof(EnsStageArgs) * 4);
  int gb = 0;
  for (int g = 0; g < kEnsGroups; g++) {
    auto &G = E->grp[g];
    for (int s = 0; s < E->nstage; s++) ea[s].gb[g] = gb;
    if (G.members.empty()) continue;
    StageArgs a;
    std::memset(&a, 0, sizeof a);
    MeshDev &m = a.m;
    m.K = G.K;
    m.xr = G.xr; m.yr = G.yr; m.xs = G.xs; m.ys = G.ys; m.J = G.J; m.V = G.V;
    m.slot = G.slot; m.snx = G.snx; m.sny = G.sny; m.slen = G.slen; m.eps = G.eps;
    m.nb_bnd = G.nb; m.bnd_elem = G.bnd_elem; m.bnd_code = G.bnd_code;
    m.S = G.S; m.side = G.side; m.bnd_val = G.bnd_val;
    a.st = E->d_st;
    a.ph = E->ph;
    a.cfl_mode = cfl_mode;
    a.part = G.part;
    a.bad = G.bad;
    a.ens = G.blk[c.bi[g]];
    double *io[5];
    io[0] = G.c;
    for (int s = 1; s < E->nstage; s++) io[s] = (s & 1) ? G.b1 : G.b2;
    io[E->nstage] = G.c;
    StageArgs st[4];
    make_stages(a, E->cfg.integrator, pf, st, io, G.acc, false);
    for (int s = 0; s < E->nstage; s++) {
      if (pf) {
        st[s].flags_in = 1;
        st[s].flags_out = 1;
      }
      ea[s].a[g] = st[s];
    }
    gb += G.K / kEnsB[c.bi[g]];
  }
  for (int s = 0; s < E->nstage; s++) ea[s].gb[kEnsGroups] = gb;
  *total = gb;
}

// the step-closing arguments shared by all instances
static FinishArgs ens_finish_args(dg_ensemble *E, int cfl_mode, double fixed_dt, int initial) {
  FinishArgs f;
  std::memset(&f, 0, sizeof f);
  f.cfl_mode = cfl_mode;
  f.fixed_dt = fixed_dt;
  f.initial = initial;
  return f;
}

// the partials of every group, in blocks of bs[g] (c's block sizes; 0: the
// 256-element blocks of the CFL pass at a step's start)
static EnsParts ens_parts(dg_ensemble *E, const EnsCfg *c) {
  EnsParts pp;
  for (int g = 0; g < kEnsGroups; g++) {
    pp.part[g] = E->grp[g].part;
    pp.bs[g] = c ? kEnsB[c->bi[g]] : kEnsPad;
  }
  return pp;
}

// the next dt of every instance from its current state (and positivity flags)
static int ens_begin(dg_ensemble *E, int cfl_mode, double fixed_dt) {
  for (int g = 0; g < kEnsGroups; g++) {
    auto &G = E->grp[g];
    if (G.members.empty()) continue;
    MeshDev m;
    std::memset(&m, 0, sizeof m);
    m.K = G.K;
    m.eps = G.eps;
    launch_lambda_ens(E->cfg.physics, g, m, E->ph, G.c, G.blk[2], G.part, G.bad, E->stream);
  }
  launch_finish_ens(E->d_st, E->d_inst, (int)E->inst.size(), ens_parts(E, nullptr),
                    ens_finish_args(E, cfl_mode, fixed_dt, 1), E->stream);
  EHIPCHK(E, hipGetLastError());
  return DG_OK;
}

static int ens_launch_step(dg_ensemble *E, const EnsCfg &c, const EnsStageArgs *ea, int total,
                           int cfl_mode) {
  int bs[kEnsGroups];
  for (int g = 0; g < kEnsGroups; g++) bs[g] = kEnsB[c.bi[g]];
  for (int s = 0; s < E->nstage; s++) {
    // time-dependent boundary data at each instance's own stage time
    if (E->tdep)
      for (int g = 0; g < kEnsGroups; g++)
        if (!E->grp[g].members.empty())
          launch_bnd_ens(E->cfg.physics, g, ea[s].a[g].m, E->d_ph, E->grp[g].bnd_inst, E->d_st,
                         ea[s].a[g].cstage, E->stream);
    if (c.split)
      launch_stage_ens_split(E->cfg.physics, bs, ea[s], E->stream);
    else
      launch_stage_ens(E->cfg.physics, bs[0], ea[s], total, E->stream);
  }
  launch_finish_ens(E->d_st, E->d_inst, (int)E->inst.size(), ens_parts(E, &c),
                    ens_finish_args(E, cfl_mode, 0., 0), E->stream);
  return DG_OK;
}

// nsteps steps of every instance in launch layout c (graph-replayed)
static int ens_run(dg_ensemble *E, const EnsCfg &c, int nsteps, double fixed_dt,
                   double *ms_per_step) {
  const int cfl_mode = fixed_dt > 0 ? 0 : 1;
  EnsStageArgs ea[4];
  int total = 0;
  ens_stage_args(E, c, ea, &total, cfl_mode);
  int rc = ens_begin(E, cfl_mode, fixed_dt);
  if (rc) return rc;
  const int key = c.key(cfl_mode);
  if (E->graph_key != key) {
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    E->graph = nullptr;
    E->graph_key = -1;
    hipGraph_t g;
    EHIPCHK(E, hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    ens_launch_step(E, c, ea, total, cfl_mode);
    EHIPCHK(E, hipStreamEndCapture(E->stream, &g));
    hipError_t e = hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return efail(E, DG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    E->graph_key = key;
  }
  EHIPCHK(E, hipEventRecord(E->ev0, E->stream));
  for (int s = 0; s < nsteps; s++) EHIPCHK(E, hipGraphLaunch(E->graph, E->stream));
  EHIPCHK(E, hipEventRecord(E->ev1, E->stream));
  EHIPCHK(E, hipStreamSynchronize(E->stream));
  float ms = 0.f;
  EHIPCHK(E, hipEventElapsedTime(&ms, E->ev0, E->ev1));
  if (ms_per_step) *ms_per_step = nsteps > 0 ? ms / nsteps : 0.;
  return DG_OK;
}

extern "C" int dg_ens_build(dg_ensemble *E, int block) {
  if (!E) return DG_ERR_ARG;
  if (E->built) return efail(E, DG_ERR_STATE, "dg_ens_build: already built");
  if (E->inst.empty()) return efail(E, DG_ERR_STATE, "dg_ens_build: no instances");
  const bool force_split = block > 0 && (block & DG_ENS_SPLIT);
  if (block > 0) block &= ~DG_ENS_SPLIT;
  if (block > 0 && block != 64 && block != 128 && block != 256)
    return efail(E, DG_ERR_ARG, "dg_ens_build: block %d (64, 128, 256, optionally | DG_ENS_SPLIT, "
                 "or <= 0 to tune)", block);
  // orders 4 and 5 run only split, in 64-element blocks
  bool high = false;
  for (dg_handle *h : E->inst) high = high || h->cfg.p > 3;
  if (high && block > 0 && !(block == 64 && force_split))
    return efail(E, DG_ERR_ARG, "dg_ens_build: instances of order 4 or 5 take block 64 | "
                 "DG_ENS_SPLIT (or <= 0 to tune)");
  EHIPCHK(E, hipSetDevice(E->cfg.device));
  E->ph = E->inst[0]->ph;
  const int ninst = (int)E->inst.size(), N = E->N;
  E->inst_off.assign(ninst, 0);
  std::vector<EnsInst> hinst(ninst);
  // layout of each group
  for (int i = 0; i < ninst; i++) {
    dg_handle *h = E->inst[i];
    auto &G = E->grp[h->cfg.p];
    G.p = h->cfg.p;
    G.members.push_back(i);
    E->inst_off[i] = G.K;
    const long long padded = (long long)(h->K + kEnsPad - 1) / kEnsPad * kEnsPad;
    if ((long long)G.K + padded >= (1ll << 28) || (long long)G.S + h->S >= (1ll << 28))
      return efail(E, DG_ERR_ARG, "dg_ens_build: a group exceeds 2^28 elements or sides");
    // the four-lane Euler kernel (p >= 4) adds a variable's row offset to the
    // 32-bit element offsets, as dg_set_mesh checks for one mesh
    if (h->cfg.physics == DG_EULER && h->cfg.p >= 4 &&
        ((long long)G.K + padded) * h->N * h->np >= (1ll << 28))
      return efail(E, DG_ERR_ARG, "dg_ens_build: the Euler order-%d group exceeds 2^28 "
                   "elements x coefficients", h->cfg.p);
    hinst[i].g = h->cfg.p;
    hinst[i].e0 = G.K;
    hinst[i].e1 = G.K + h->K;
    hinst[i].p = h->cfg.p;
    hinst[i].cfl_num = E->cfg.cfl * h->min_r;
    G.K += (int)padded;
    G.S += h->S;
    G.nb += h->nb_bnd;
  }
  int rc;
  for (int g = 0; g < kEnsGroups; g++) {
    auto &G = E->grp[g];
    if (G.members.empty()) continue;
    const int np = np_of(g), nq1 = g + 1;
    const size_t K = G.K, rows = (size_t)N * np;
    if ((rc = ealloc(E, G.xr, K)) || (rc = ealloc(E, G.yr, K)) || (rc = ealloc(E, G.xs, K)) ||
        (rc = ealloc(E, G.ys, K)) || (rc = ealloc(E, G.J, K)) || (rc = ealloc(E, G.V, 6 * K)) ||
        (rc = ealloc(E, G.eps, K)) || (rc = ealloc(E, G.slot, 3 * K)) ||
        (rc = ealloc(E, G.snx, G.S)) || (rc = ealloc(E, G.sny, G.S)) ||
        (rc = ealloc(E, G.slen, G.S)) || (rc = ealloc(E, G.side, G.S)) ||
        (rc = ealloc(E, G.bnd_elem, G.nb)) || (rc = ealloc(E, G.bnd_code, G.nb)) ||
        (rc = ealloc(E, G.bnd_inst, G.nb)) ||
        (rc = ealloc(E, G.bnd_val, (size_t)G.nb * nq1 * NMAX)) ||
        (rc = ealloc(E, G.bad, K)) || (rc = ealloc(E, G.part, 2 * (K / 64))))
      return rc;
    {
      // the state buffers in one allocation, as a handle's (block_views)
      const size_t stride = (rows * K + 31) / 32 * 32;  // doubles: 256-byte offsets
      const int nv = E->cfg.integrator == DG_RK4 ? 4 : 3;
      if ((rc = ealloc(E, G.state, stride * nv))) return rc;
      G.c = G.state;
      G.b1 = G.state + stride;
      G.b2 = G.state + 2 * stride;
      G.acc = nv == 4 ? G.state + 3 * stride : nullptr;
    }
    EHIPCHK(E, hipMemsetAsync(G.c, 0, sizeof(double) * rows * K, E->stream));
    EHIPCHK(E, hipMemsetAsync(G.bad, 0, sizeof(int) * K, E->stream));
    int soff = 0, boff = 0;
    std::vector<int4> blk[3];
    for (int b = 0; b < 3; b++) blk[b].resize(K / kEnsB[b]);
    for (int i : G.members) {
      dg_handle *h = E->inst[i];
      const int off = E->inst_off[i], Ki = h->K, Si = h->S, nbi = h->nb_bnd;
      const size_t kd = sizeof(double) * Ki;
      auto d2d = [&](double *dst, const double *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, E->stream);
      };
      EHIPCHK(E, d2d(G.xr + off, h->d_xr, kd));
      EHIPCHK(E, d2d(G.yr + off, h->d_yr, kd));
      EHIPCHK(E, d2d(G.xs + off, h->d_xs, kd));
      EHIPCHK(E, d2d(G.ys + off, h->d_ys, kd));
      EHIPCHK(E, d2d(G.J + off, h->d_J, kd));
      EHIPCHK(E, d2d(G.eps + off, h->d_eps, kd));
      for (int r = 0; r < 6; r++) EHIPCHK(E, d2d(G.V + r * K + off, h->d_V + (size_t)r * Ki, kd));
      EHIPCHK(E, d2d(G.snx + soff, h->d_snx, sizeof(double) * Si));
      EHIPCHK(E, d2d(G.sny + soff, h->d_sny, sizeof(double) * Si));
      EHIPCHK(E, d2d(G.slen + soff, h->d_slen, sizeof(double) * Si));
      if (nbi)
        EHIPCHK(E, d2d(G.bnd_val + (size_t)boff * nq1 * NMAX, h->d_bnd_val,
                       sizeof(double) * nbi * nq1 * NMAX));
      EHIPCHK(E, hipMemcpy2DAsync(G.c + off, sizeof(double) * K, h->d_c, kd, kd, rows,
                                  hipMemcpyDeviceToDevice, E->stream));
      // indices: neighbours into the group's elements, sides into its sides,
      // boundary slots into its boundary data
      std::vector<int2> slot(3 * (size_t)Ki);
      std::vector<int4> side(Si);
      std::vector<int> belem(nbi), bcode(nbi);
      EHIPCHK(E, hipMemcpy(slot.data(), h->d_slot, sizeof(int2) * slot.size(), hipMemcpyDeviceToHost));
      EHIPCHK(E, hipMemcpy(side.data(), h->d_side, sizeof(int4) * Si, hipMemcpyDeviceToHost));
      if (nbi) {
        EHIPCHK(E, hipMemcpy(belem.data(), h->d_bnd_elem, sizeof(int) * nbi, hipMemcpyDeviceToHost));
        EHIPCHK(E, hipMemcpy(bcode.data(), h->d_bnd_code, sizeof(int) * nbi, hipMemcpyDeviceToHost));
      }
      for (auto &q : slot) {
        if (q.x >= 0) q.x += off;
        q.y = (q.y & 7) | (((q.y >> 3) + soff) << 3);
      }
      for (auto &q : side) {
        q.x += off;
        if (q.y >= 0) q.y += off;
        if (q.w >= 0) q.w += boff;
      }
      for (auto &q : belem) q += off;
      for (int k = 0; k < 3; k++)
        EHIPCHK(E, hipMemcpy(G.slot + k * K + off, slot.data() + (size_t)k * Ki, sizeof(int2) * Ki,
                             hipMemcpyHostToDevice));
      EHIPCHK(E, hipMemcpy(G.side + soff, side.data(), sizeof(int4) * Si, hipMemcpyHostToDevice));
      if (nbi) {
        const std::vector<int> binst(nbi, i);
        EHIPCHK(E, hipMemcpy(G.bnd_elem + boff, belem.data(), sizeof(int) * nbi, hipMemcpyHostToDevice));
        EHIPCHK(E, hipMemcpy(G.bnd_code + boff, bcode.data(), sizeof(int) * nbi, hipMemcpyHostToDevice));
        EHIPCHK(E, hipMemcpy(G.bnd_inst + boff, binst.data(), sizeof(int) * nbi, hipMemcpyHostToDevice));
      }
      const int padded = (Ki + kEnsPad - 1) / kEnsPad * kEnsPad;
      for (int b = 0; b < 3; b++)
        for (int lb = off / kEnsB[b]; lb < (off + padded) / kEnsB[b]; lb++)
          blk[b][lb] = make_int4(i, off, off + Ki, 0);
      soff += Si;
      boff += nbi;
    }
    for (int b = 0; b < 3; b++) {
      if ((rc = ealloc(E, G.blk[b], blk[b].size()))) return rc;
      EHIPCHK(E, hipMemcpy(G.blk[b], blk[b].data(), sizeof(int4) * blk[b].size(),
                           hipMemcpyHostToDevice));
    }
  }
  if ((rc = ealloc(E, E->d_inst, ninst)) || (rc = ealloc(E, E->d_st, ninst)) ||
      (rc = ealloc(E, E->d_ph, ninst)))
    return rc;
  {
    std::vector<PhysParams> hph(ninst);
    E->tdep = false;
    for (int i = 0; i < ninst; i++) {
      hph[i] = E->inst[i]->ph;
      E->tdep = E->tdep || bnd_time_dependent(E->inst[i]);
    }
    EHIPCHK(E, hipMemcpy(E->d_ph, hph.data(), sizeof(PhysParams) * ninst, hipMemcpyHostToDevice));
  }
  EHIPCHK(E, hipMemcpy(E->d_inst, hinst.data(), sizeof(EnsInst) * ninst, hipMemcpyHostToDevice));
  std::vector<DevState> hst(ninst);
  for (auto &s : hst) {
    std::memset(&s, 0, sizeof s);
    s.t_end = HUGE_VAL;
    s.conv_tol = -1.;
  }
  EHIPCHK(E, hipMemcpy(E->d_st, hst.data(), sizeof(DevState) * ninst, hipMemcpyHostToDevice));
  EHIPCHK(E, hipStreamSynchronize(E->stream));
  E->built = true;
  // block size: given, or the fastest of the candidates on this ensemble
  // (two steps each from the built state, which is then restored)
  if (block > 0) {
    for (int b = 0; b < 3; b++)
      if (kEnsB[b] == block) E->B = b;
    E->split = force_split;
    for (int &b : E->Bg) b = E->B;
    if (!force_split && ens_lds_bytes(E->cfg.physics, block) > 65536)
      return efail(E, DG_ERR_ARG, "dg_ens_build: block %d needs more LDS than a block has", block);
    return DG_OK;
  }
  // DG_ENS_BLOCKS="b0,b1,...,b5" (elements per block of the order groups
  // 0..5) fixes a split layout instead of tuning (for diagnosis and tests)
  if (const char *env = std::getenv("DG_ENS_BLOCKS")) {
    E->split = true;
    E->B = 0;
    const char *q = env;
    for (int g = 0; g < kEnsGroups; g++) {
      char *end = nullptr;
      const long v = std::strtol(q, &end, 10);
      int bi = -1;
      for (int b = 0; b < 3; b++)
        if (end != q && v == kEnsB[b]) bi = b;
      if (bi < 0 || (g > 3 && bi != 0))
        return efail(E, DG_ERR_ARG, "DG_ENS_BLOCKS: order %d needs 64, 128 or 256 (orders 4, 5: 64)",
                     g);
      E->Bg[g] = bi;
      q = *end == ',' ? end + 1 : end;
    }
    return DG_OK;
  }
  std::vector<std::vector<double>> save(kEnsGroups);
  for (int g = 0; g < kEnsGroups; g++) {
    auto &G = E->grp[g];
    if (G.members.empty()) continue;
    save[g].resize((size_t)N * np_of(g) * G.K);
    EHIPCHK(E, hipMemcpy(save[g].data(), G.c, sizeof(double) * save[g].size(), hipMemcpyDeviceToHost));
  }
  // one candidate: two steps from the built state (the first builds the
  // graph), then the state is restored
  auto time_cfg = [&](const EnsCfg &c, double *ms) -> int {
    int r = ens_run(E, c, 1, 0., nullptr);
    if (!r) r = ens_run(E, c, 2, 0., ms);
    if (r) return r;
    for (int g = 0; g < kEnsGroups; g++) {
      auto &G = E->grp[g];
      if (G.members.empty()) continue;
      EHIPCHK(E, hipMemcpy(G.c, save[g].data(), sizeof(double) * save[g].size(), hipMemcpyHostToDevice));
    }
    EHIPCHK(E, hipMemcpy(E->d_st, hst.data(), sizeof(DevState) * ninst, hipMemcpyHostToDevice));
    return DG_OK;
  };
  // candidates: the block sizes, each as one launch for all groups and as
  // a launch per group (each group at its own order's register budget)
  double best = HUGE_VAL;
  for (int b = 0; b < 3; b++)
    for (int sp = 0; sp < 2; sp++) {
      double &slot = sp ? E->tune_split_ms[b] : E->tune_ms[b];
      slot = -1.;
      if (!sp && ens_lds_bytes(E->cfg.physics, kEnsB[b]) > 65536) continue;
      if (high && !(sp && kEnsB[b] == 64)) continue;
      double ms = 0.;
      if ((rc = time_cfg(ens_uniform(b, sp != 0), &ms))) return rc;
      slot = ms;
      if (ms < best) {
        best = ms;
        E->B = b;
        E->split = sp != 0;
      }
    }
  for (int &b : E->Bg) b = E->B;
  // split: each order group's own block size, one group at a time from the
  // best uniform layout (orders 4 and 5 stay at 64 elements)
  if (E->split) {
    EnsCfg c = ens_uniform(E->B, true);
    for (int g = 0; g < kEnsGroups && g <= 3; g++) {
      if (E->grp[g].members.empty()) continue;
      const int keep = c.bi[g];
      for (int b = 0; b < 3; b++) {
        if (b == keep) continue;
        EnsCfg t = c;
        t.bi[g] = b;
        double ms = 0.;
        if ((rc = time_cfg(t, &ms))) return rc;
        if (ms < best) {
          best = ms;
          c = t;
        }
      }
    }
    for (int g = 0; g < kEnsGroups; g++) E->Bg[g] = c.bi[g];
  }
  E->tune_best_ms = best;
  return DG_OK;
}

extern "C" int dg_ens_step(dg_ensemble *E, int nsteps, double fixed_dt) {
  if (!E || nsteps < 0) return DG_ERR_ARG;
  if (!E->built) return efail(E, DG_ERR_STATE, "dg_ens_step: build the ensemble first");
  EHIPCHK(E, hipSetDevice(E->cfg.device));
  if (nsteps == 0) return DG_OK;
  EnsCfg c = ens_uniform(E->B, E->split);
  if (E->split)
    for (int g = 0; g < kEnsGroups; g++) c.bi[g] = E->Bg[g];
  return ens_run(E, c, nsteps, fixed_dt, &E->last_ms_per_step);
}

static int ens_check_inst(dg_ensemble *E, int i) {
  if (!E) return DG_ERR_ARG;
  if (!E->built) return efail(E, DG_ERR_STATE, "build the ensemble first");
  if (i < 0 || i >= (int)E->inst.size()) return efail(E, DG_ERR_ARG, "no instance %d", i);
  return DG_OK;
}

extern "C" int dg_ens_get_coeffs(dg_ensemble *E, int i, double *C, int where) {
  int rc = ens_check_inst(E, i);
  if (rc) return rc;
  if (!C) return DG_ERR_ARG;
  dg_handle *h = E->inst[i];
  auto &G = E->grp[h->cfg.p];
  const size_t kd = sizeof(double) * h->K;
  EHIPCHK(E, hipSetDevice(E->cfg.device));
  // a row at a time: under rocprofv3's kernel trace, hipMemcpy2DAsync to
  // pageable host memory returned non-finite values for 76 of 256 instances
  // whose device state was intact (their t and dt matched the instances
  // stepped alone); the row copies read back correctly there
  const int rows = E->N * h->np;
  for (int r = 0; r < rows; r++)
    EHIPCHK(E, hipMemcpyAsync(C + (size_t)r * h->K, G.c + (size_t)r * G.K + E->inst_off[i], kd,
                              kind_from(where), E->stream));
  EHIPCHK(E, hipStreamSynchronize(E->stream));
  return DG_OK;
}

extern "C" int dg_ens_set_coeffs(dg_ensemble *E, int i, const double *C, int where) {
  int rc = ens_check_inst(E, i);
  if (rc) return rc;
  if (!C) return DG_ERR_ARG;
  dg_handle *h = E->inst[i];
  auto &G = E->grp[h->cfg.p];
  const size_t kd = sizeof(double) * h->K;
  EHIPCHK(E, hipSetDevice(E->cfg.device));
  const int rows = E->N * h->np;  // a row at a time, as dg_ens_get_coeffs
  for (int r = 0; r < rows; r++)
    EHIPCHK(E, hipMemcpyAsync(G.c + (size_t)r * G.K + E->inst_off[i], C + (size_t)r * h->K, kd,
                              kind_to(where), E->stream));
  EHIPCHK(E, hipStreamSynchronize(E->stream));
  return DG_OK;
}

extern "C" int dg_ens_state(dg_ensemble *E, int i, double *t, double *dt, long long *steps) {
  int rc = ens_check_inst(E, i);
  if (rc) return rc;
  DevState s;
  EHIPCHK(E, hipSetDevice(E->cfg.device));
  EHIPCHK(E, hipMemcpy(&s, E->d_st + i, sizeof s, hipMemcpyDeviceToHost));
  if (t) *t = s.tdone;
  if (dt) *dt = s.dt_last;
  if (steps) *steps = s.steps;
  if (s.nan_flag) return efail(E, DG_ERR_NAN, "instance %d: the NaN-dt guard stopped it", i);
  return DG_OK;
}

extern "C" int dg_ens_info(dg_ensemble *E, dg_ens_info_t *info) {
  if (!E || !info) return DG_ERR_ARG;
  std::memset(info, 0, sizeof *info);
  info->ninst = (int)E->inst.size();
  for (dg_handle *h : E->inst) {
    info->dofs += (long long)dofs(h);
    info->elements += h->K;
  }
  info->block = E->built ? kEnsB[E->B] : 0;
  for (int b = 0; b < 3; b++) {
    info->tune_ms[b] = E->tune_ms[b];
    info->tune_split_ms[b] = E->tune_split_ms[b];
  }
  info->split = E->split ? 1 : 0;
  info->vgprs = E->built && !E->split ? ens_vgprs(E->cfg.physics, kEnsB[E->B]) : -1;
  int groups = 0;
  for (int g = 0; g < kEnsGroups; g++) groups += E->grp[g].members.empty() ? 0 : 1;
  info->launches_per_step = E->nstage * (E->split ? groups : 1) + 1;
  info->ms_per_step = E->last_ms_per_step;
  for (int g = 0; g < kEnsGroups; g++)
    info->block_per_order[g] =
        !E->built || E->grp[g].members.empty() ? 0 : kEnsB[E->split ? E->Bg[g] : E->B];
  info->tune_best_ms = E->tune_best_ms;
  return DG_OK;
}

// the instance's own handle (mesh queries, errors against its problem)
extern "C" int dg_ens_l2_error(dg_ensemble *E, int i, double t, double *err) {
  int rc = ens_check_inst(E, i);
  if (rc) return rc;
  dg_handle *h = E->inst[i];
  std::vector<double> C(dofs(h));
  if ((rc = dg_ens_get_coeffs(E, i, C.data(), DG_HOST))) return rc;
  if ((rc = dg_set_coeffs(h, C.data(), DG_HOST))) return efail(E, rc, "%s", h->err.c_str());
  if ((rc = dg_l2_error(h, t, err))) return efail(E, rc, "%s", h->err.c_str());
  return DG_OK;
}
