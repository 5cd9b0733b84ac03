// dg_kernels.hip -- HIP kernels of the DG right-hand side for gfx950.
//
// The reference evaluates one RK stage with four launches and two large
// side-indexed intermediates (paths relative to src/2d/conservation/):
//   eval_surface  conserv_kernels.cu:899-985   one thread per side, writes
//                                              rhs_surface_{left,right}[N][n_p][S]
//   eval_volume   conserv_kernels.cu:992-1083  one thread per element
//   eval_rhs      time_integrator.cu:26-101    gathers the side arrays, k = dt/J (...)
//   rk_tempstorage / rk4  time_integrator.cu:107-130
//   eval_global_lambda + host max  euler.cu:248-277, time_integrator.cu:200-217
// Here one kernel (k_stage) does all of that per element: the element
// evaluates its volume integral, recomputes the numerical flux on each of
// its three sides from its own trace and the neighbour's (the neighbour's
// coefficients are gathered; no side arrays, no atomics), applies dt/J,
// combines the RK stage, and -- on the last stage -- reduces the CFL speed
// of the new state and computes the next dt on the device.  Both elements of
// a side evaluate its flux in the same code path from the same values, so
// the two contributions are exact negatives and the scheme stays
// conservative to the last bit.
//
// Layout: coefficients in the reference's SoA order C[(n*n_p + i)*K + e], so
// one thread per element reads and writes fully coalesced 512-byte lines
// per wavefront.  Element blocks are remapped so that each of the 8 XCDs
// gets a contiguous range of elements (neighbour gathers then hit the XCD's
// own L2).
#include <type_traits>
#include <utility>

#include "dg_internal.h"
#include "dg_physics.h"
#include "dg_tabvals.h"
#include "dg_device.h"
#include "dg_stage.h"
#include "dg_stage_q.h"

namespace dgk {

__constant__ Tab<0> c_tab0;
__constant__ Tab<1> c_tab1;
__constant__ Tab<2> c_tab2;
__constant__ Tab<3> c_tab3;
__constant__ Tab<4> c_tab4;
__constant__ Tab<5> c_tab5;

template <int P>
__device__ __forceinline__ const Tab<P> &tab();
template <> __device__ __forceinline__ const Tab<0> &tab<0>() { return c_tab0; }
template <> __device__ __forceinline__ const Tab<1> &tab<1>() { return c_tab1; }
template <> __device__ __forceinline__ const Tab<2> &tab<2>() { return c_tab2; }
template <> __device__ __forceinline__ const Tab<3> &tab<3>() { return c_tab3; }
template <> __device__ __forceinline__ const Tab<4> &tab<4>() { return c_tab4; }
template <> __device__ __forceinline__ const Tab<5> &tab<5>() { return c_tab5; }

int upload_tables(int p, const void *t, size_t bytes) {
  hipError_t e = hipErrorInvalidValue;
  switch (p) {
    case 0: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab0), t, bytes); break;
    case 1: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab1), t, bytes); break;
    case 2: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab2), t, bytes); break;
    case 3: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab3), t, bytes); break;
    case 4: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab4), t, bytes); break;
    case 5: e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab5), t, bytes); break;
  }
  return e == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ fixed scratch
//
// dg_eval_rhs_noalloc (for callers that cannot allocate: code under graph
// capture, serving loops) takes its scratch from this file-scope array
// instead of the handle's buffers: the step state (t, dt) and the positivity
// flags of its input.  Each GPU holds one copy for the life of the process;
// every call on a GPU shares it, so calls that overlap on one GPU race.
constexpr size_t kScratchState = 256;  // DevState, padded
static_assert(sizeof(DevState) <= kScratchState, "DevState must fit its scratch slot");
static __device__ __align__(256) unsigned char scratch[kScratchState + sizeof(int) * kScratchMaxK];

// the device address of this GPU's scratch (the array's name is a host-side
// stand-in and must never be passed to a kernel); call with the device current
int scratch_address(void **p) {
  *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(p, HIP_SYMBOL(scratch));
  return e == hipSuccess && *p ? 0 : -1;
}

// the stage state of a single right-hand-side evaluation at (t, dt), written
// by a kernel (no memset or copy: the launch sequence stays capturable)
__global__ void k_set_state(DevState *st, double t, double dt) {
  DevState s;
  __builtin_memset(&s, 0, sizeof s);
  s.t = t;
  s.dt = dt;
  s.tdone = t;
  s.t_end = __builtin_huge_val();
  s.conv_tol = -1.;
  *st = s;
}

void launch_set_state(DevState *st, double t, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, s, st, t, dt);
}

__global__ __launch_bounds__(BLK) void k_finish(DevState *st, const unsigned long long *part,
                                                int nblk, FinishArgs f) {
  finish_body(st, part, f.conv ? part + nblk : nullptr, nblk, f, f.t0);
}

void launch_finish(DevState *st, const unsigned long long *part, int nblk, const FinishArgs &f,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLK), 0, s, st, part, nblk, f);
}

template <int PHYS, int P>
__global__ __launch_bounds__((stage_blk<PHYS, P>())) void k_stage(const StageArgs a) {
  constexpr int NP = np_of(P), NQ1 = P + 1, N = nvars_of(PHYS);
  constexpr int B = stage_blk<PHYS, P>();
  __shared__ double s_side[3 * NP * NQ1];
  __shared__ double s_c[N * NP][B];
  stage_body<PHYS, P, B>(a, blockIdx.x, a.nblk, s_side, s_c);
}

void launch_stage(int phys, int p, const StageArgs &a0, hipStream_t s) {
  StageArgs a = a0;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    if constexpr (stage_quad<ph, pp>()) {
      a.nblk = nblocks(a.m.K, kQuadBlk / 4);
      const dim3 g(a.nblk), b(kQuadBlk);
      const bool tr = a.tr_in != nullptr;
      if constexpr (volume_mfma_ok<ph, pp>()) {
        if (a.vol_mfma) {
          if (tr)
            hipLaunchKernelGGL((k_stage_q<ph, pp, true, true>), g, b, 0, s, a);
          else
            hipLaunchKernelGGL((k_stage_q<ph, pp, true, false>), g, b, 0, s, a);
          return;
        }
      }
      if (tr)
        hipLaunchKernelGGL((k_stage_q<ph, pp, false, true>), g, b, 0, s, a);
      else
        hipLaunchKernelGGL((k_stage_q<ph, pp, false, false>), g, b, 0, s, a);
    } else {
      constexpr int B = stage_blk<ph, pp>();
      a.nblk = nblocks(a.m.K, B);
      hipLaunchKernelGGL((k_stage<ph, pp>), dim3(a.nblk), dim3(B), 0, s, a);
    }
  });
  if (a.last_stage) {
    FinishArgs f{};
    f.cfl_mode = a.cfl_mode;
    f.cfl_num = a.cfl_num;
    f.p = p;
    f.conv = a.conv;
    launch_finish(a.st, a.part, a.nblk, f, s);
  }
}

// the traces of state u (four-lane kernels only; a no-op elsewhere)
void launch_traces(int phys, int p, int K, const double *u, double *tr, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    if constexpr (stage_quad<ph, pp>())
      hipLaunchKernelGGL((k_traces_q<ph, pp>), dim3(nblocks(K, kQuadBlk / 4)), dim3(kQuadBlk), 0, s,
                         u, tr, K);
  });
}

int stage_vgprs(int phys, int p, bool vol_mfma) {
  int v = -1;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    hipFuncAttributes attr;
    const void *fn;
    if constexpr (volume_mfma_ok<ph, pp>())
      fn = vol_mfma ? reinterpret_cast<const void *>(&k_stage_q<ph, pp, true, true>)
                    : reinterpret_cast<const void *>(&k_stage_q<ph, pp, false, true>);
    else if constexpr (stage_quad<ph, pp>())
      fn = reinterpret_cast<const void *>(&k_stage_q<ph, pp, false, true>);
    else
      fn = reinterpret_cast<const void *>(&k_stage<ph, pp>);
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) v = attr.numRegs;
  });
  return v;
}

// ------------------------------------------------------------ setup kernels

// boundary data of every boundary slot (eval_boundary conserv_kernels.cu:820-843)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_bnd(const MeshDev m, const PhysParams ph,
                                             const DevState *st, double cstage, int use_state,
                                             double t0) {
  const int b = blockIdx.x * BLK + threadIdx.x;
  if (b >= m.nb_bnd) return;
  const double t = use_state ? st->t + cstage * st->dt : t0;
  bnd_slot<PHYS, P>(m, ph, t, b);
}

void launch_bnd(int phys, int p, const MeshDev &m, const PhysParams &ph, const DevState *st,
                double cstage, int use_state, double t, hipStream_t s) {
  if (m.nb_bnd <= 0) return;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_bnd<ph_, pp>), dim3(nblocks(m.nb_bnd, BLK)), dim3(BLK), 0, s, m, ph,
                       st, cstage, use_state, t);
  });
}


// init_conditions conserv_kernels.cu:201-228 (evalU0: euler.cu:31-55)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_init(const MeshDev m, const PhysParams ph, double *c) {
  constexpr int NP = np_of(P), NQ = nq_of(P), N = nvars_of(PHYS);
  const Tab<P> &T = tab<P>();
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  const double V0 = m.V[e], V1 = m.V[(size_t)K + e], V2 = m.V[2 * (size_t)K + e];
  const double V3 = m.V[3 * (size_t)K + e], V4 = m.V[4 * (size_t)K + e],
               V5 = m.V[5 * (size_t)K + e];
  double u0[NQ][NMAX];
  for (int j = 0; j < NQ; j++) {
    const double X = V2 * T.R1[j] + V4 * T.R2[j] + V0 * (1 - T.R1[j] - T.R2[j]);
    const double Y = V3 * T.R1[j] + V5 * T.R2[j] + V1 * (1 - T.R1[j] - T.R2[j]);
    exact_state<N>(ph, X, Y, 0., u0[j]);
  }
  for (int i = 0; i < NP; i++)
    for (int n = 0; n < N; n++) {
      double U = 0.;
      for (int j = 0; j < NQ; j++) U += T.W[j] * u0[j][n] * T.B[i * NQ + j];
      c[(size_t)(n * NP + i) * K + e] = U;
    }
}

void launch_init(int phys, int p, const MeshDev &m, const PhysParams &ph, double *c,
                 hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_init<ph_, pp>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, m, ph, c);
  });
}

// eval_error_L2 conserv_kernels.cu:1432-1479 (NORM 2) and eval_error_L1
// conserv_kernels.cu:1378-1430 (NORM 1), per element, all variables
template <int PHYS, int P, int NORM>
__global__ __launch_bounds__(BLK) void k_err(const MeshDev m, const PhysParams ph,
                                             const double *c, double t, double *err) {
  constexpr int NP = np_of(P), NQ = nq_of(P), N = nvars_of(PHYS);
  const Tab<P> &T = tab<P>();
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  const double V0 = m.V[e], V1 = m.V[(size_t)K + e], V2 = m.V[2 * (size_t)K + e];
  const double V3 = m.V[3 * (size_t)K + e], V4 = m.V[4 * (size_t)K + e],
               V5 = m.V[5 * (size_t)K + e];
  const double detJ = m.J[e];
  double acc[NMAX] = {0., 0., 0., 0.};
  for (int j = 0; j < NQ; j++) {
    const double X = V2 * T.R1[j] + V4 * T.R2[j] + V0 * (1 - T.R1[j] - T.R2[j]);
    const double Y = V3 * T.R1[j] + V5 * T.R2[j] + V1 * (1 - T.R1[j] - T.R2[j]);
    double u[NMAX];
    exact_state<N>(ph, X, Y, t, u);
    for (int n = 0; n < N; n++) {
      double U = 0.;
      for (int i = 0; i < NP; i++) U += c[(size_t)(n * NP + i) * K + e] * T.B[i * NQ + j];
      if constexpr (NORM == 2)
        acc[n] += T.W[j] * (u[n] - U) * (u[n] - U) * detJ;
      else
        acc[n] += T.W[j] * fabs(u[n] - U) * detJ;
    }
  }
  for (int n = 0; n < N; n++) err[(size_t)n * K + e] = acc[n];
}

void launch_err(int phys, int p, int norm, const MeshDev &m, const PhysParams &ph,
                const double *c, double t, double *err, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    if (norm == 1)
      hipLaunchKernelGGL((k_err<ph_, pp, 1>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, m, ph, c,
                         t, err);
    else
      hipLaunchKernelGGL((k_err<ph_, pp, 2>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, m, ph, c,
                         t, err);
  });
}

// plot_exact conserv_kernels.cu:1315-1331: the exact state's variable n at the
// three vertices
template <int PHYS>
__global__ __launch_bounds__(BLK) void k_exact_vertex(const MeshDev m, const PhysParams ph,
                                                      double t, int n, double *u1, double *u2,
                                                      double *u3) {
  constexpr int N = nvars_of(PHYS);
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  double U[NMAX];
  double *out[3] = {u1, u2, u3};
  for (int v = 0; v < 3; v++) {
    exact_state<N>(ph, m.V[(size_t)(2 * v) * K + e], m.V[(size_t)(2 * v + 1) * K + e], t, U);
    out[v][e] = U[n];
  }
}

void launch_exact_vertex(int phys, const MeshDev &m, const PhysParams &ph, double t, int n,
                         double *u1, double *u2, double *u3, hipStream_t s) {
  dispatch(phys, 1, [&](auto PH, auto) {
    constexpr int ph_ = decltype(PH)::value;
    hipLaunchKernelGGL((k_exact_vertex<ph_>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, m, ph, t,
                       n, u1, u2, u3);
  });
}

// CFL speed of every element (eval_global_lambda) + device max + first dt
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_lambda_init(const MeshDev m, const PhysParams ph,
                                                     const double *c,
                                                     double *lam_out, const double *cprev,
                                                     int conv_rows, unsigned long long *part,
                                                     int *bad_out) {
  constexpr int NP = np_of(P), N = nvars_of(PHYS);
  const Tab<P> &T = tab<P>();
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  double lam = 0., dc = 0.;
  if constexpr (has_positivity<PHYS>()) {
    if (bad_out && e < K) {
      double o[N][NP];
      for (int n = 0; n < N; n++)
        for (int i = 0; i < NP; i++) o[n][i] = c[(size_t)(n * NP + i) * K + e];
      bad_out[e] = needs_fix<PHYS, P>(ph, o) ? 1 : 0;
    }
  }
  if (e < K) {
    double U[NMAX];
    for (int n = 0; n < N; n++) U[n] = c[(size_t)(n * NP) * K + e] * T.B[0];
    PhysParams pv = ph;
    if (PHYS == PH_ADV && ph.prob == PB_ADV_ROTHILL) {  // advection.cu:176-181
      double V[6], x, y;
      for (int q = 0; q < 6; q++) V[q] = m.V[(size_t)q * K + e];
      elem_point(V, T.R1[nq_of(P) / 2], T.R2[nq_of(P) / 2], x, y);
      rothill_velocity(x, y, pv.ax, pv.ay);
    }
    lam = cfl_speed<PHYS>(pv, U, PHYS == PH_TE ? m.eps[e] : 1.);
    if (lam_out) lam_out[e] = lam;
    lam = ref_max_operand(lam, e);
    if (cprev)
      for (int r = 0; r < conv_rows && r < N * NP; r++)
        dc = fmax(dc, fabs(c[(size_t)r * K + e] - cprev[(size_t)r * K + e]));
  }
  if (part) block_partials(lam, dc, cprev != nullptr, part, blockIdx.x, gridDim.x);
}

void launch_lambda_init(int phys, int p, const MeshDev &m, const PhysParams &ph, const double *c,
                        DevState *st, int cfl_mode, double cfl_num, double fixed_dt, double t0,
                        int initial, double *lam_out, hipStream_t s, const double *cprev,
                        int conv_rows, unsigned long long *part, int *bad_out) {
  const int nblk = nblocks(m.K, BLK);
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_lambda_init<ph_, pp>), dim3(nblk), dim3(BLK), 0, s, m, ph, c, lam_out,
                       cprev, conv_rows, st ? part : nullptr, bad_out);
  });
  if (st) {
    FinishArgs f{};
    f.cfl_mode = cfl_mode;
    f.cfl_num = cfl_num;
    f.p = p;
    f.fixed_dt = fixed_dt;
    f.initial = initial;
    f.t0 = t0;
    f.conv = cprev != nullptr;
    launch_finish(st, part, nblk, f, s);
  }
}

// element-local positivity fix of a state (race-free check_physical)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_posfix(int K, const PhysParams ph, double *c) {
  constexpr int NP = np_of(P), N = nvars_of(PHYS);
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  if constexpr (has_positivity<PHYS>()) {
    double o[N][NP];
    for (int n = 0; n < N; n++)
      for (int i = 0; i < NP; i++) o[n][i] = c[(size_t)(n * NP + i) * K + e];
    if (!elem_physical<PHYS, P>(ph, o))
      for (int n = 0; n < N; n++)
        for (int i = 1; i < NP; i++) c[(size_t)(n * NP + i) * K + e] = 0.;
  }
}

void launch_posfix(int phys, int p, int K, const PhysParams &ph, double *c, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_posfix<ph_, pp>), dim3(nblocks(K, BLK)), dim3(BLK), 0, s, K, ph, c);
  });
}

// positivity flags of a state (1: it fails the point check, as the stage
// kernel writes them for a stored step state)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_flags(int K, const PhysParams ph, const double *c,
                                               int *bad, int *fired) {
  constexpr int NP = np_of(P), N = nvars_of(PHYS);
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  bool b = false;
  if constexpr (has_positivity<PHYS>()) {
    double o[N][NP];
    for (int n = 0; n < N; n++)
      for (int i = 0; i < NP; i++) o[n][i] = c[(size_t)(n * NP + i) * K + e];
    b = needs_fix<PHYS, P>(ph, o);
  }
  bad[e] = b ? 1 : 0;
  if (b && fired) atomicOr(fired, 1);
}

void launch_flags(int phys, int p, int K, const PhysParams &ph, const double *c, int *bad,
                  hipStream_t s, int *fired) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_flags<ph_, pp>), dim3(nblocks(K, BLK)), dim3(BLK), 0, s, K, ph, c, bad,
                       fired);
  });
}

// limit_c conserv_kernels.cu:640-720 (p = 1)
template <int PHYS>
__global__ __launch_bounds__(BLK) void k_limit(const MeshDev m, double *C) {
  constexpr int NP = 3, NQ1 = 2, N = nvars_of(PHYS);
  const Tab<1> &T = tab<1>();
  const int K = m.K;
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  int nb[3];
  for (int k = 0; k < 3; k++) {
    int v = m.slot[(size_t)k * K + e].x;
    nb[k] = v < 0 ? e : v;
  }
  for (int n = 0; n < N; n++) {
    const double Uc = C[(size_t)(n * NP) * K + e] * T.B[0];
    double Umin = Uc, Umax = Uc, alpha, mina = 1.;
    for (int k = 0; k < 3; k++) {
      double Ui = C[(size_t)(n * NP) * K + nb[k]] * T.B[0];
      Umin = Ui < Umin ? Ui : Umin;
      Umax = Ui > Umax ? Ui : Umax;
    }
    for (int j = 0; j < NQ1; j++)
      for (int sd = 0; sd < 3; sd++) {
        double Ui = 0.;
        for (int i = 0; i < NP; i++)
          Ui += C[(size_t)(n * NP + i) * K + e] * T.SIDE[(sd * NP + i) * NQ1 + j];
        if (Ui > Uc)
          alpha = fmin(1., (Uc - Umin) / (Ui - Uc));
        else if (Ui < Uc)
          alpha = fmin(1., (Uc - Umax) / (Ui - Uc));
        else
          alpha = 1.;
        if (alpha < mina) mina = alpha;
      }
    if (mina < 0) mina = 0.;
    C[(size_t)(n * NP + 1) * K + e] *= mina;
    C[(size_t)(n * NP + 2) * K + e] *= mina;
  }
}

void launch_limit(int phys, int p, const MeshDev &m, double *c, hipStream_t s) {
  if (p != 1) return;
  dispatch(phys, 1, [&](auto PH, auto) {
    constexpr int ph_ = decltype(PH)::value;
    hipLaunchKernelGGL((k_limit<ph_>), dim3(nblocks(m.K, BLK)), dim3(BLK), 0, s, m, c);
  });
}

// ------------------------------------------------------------ elementwise

__global__ void k_axpy(long long n, double *out, const double *c, double alpha, const double *k) {
  long long i = (long long)blockIdx.x * BLK + threadIdx.x;
  if (i < n) out[i] = c[i] + alpha * k[i];  // rk_tempstorage time_integrator.cu:107-113
}

void launch_axpy(long long n, double *out, const double *c, double alpha, const double *k,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(n, BLK)), dim3(BLK), 0, s, n, out, c, alpha, k);
}

__global__ void k_ssp(long long n, double *out, const double *un, double a, const double *us,
                      double b, const double *k, double cc) {
  long long i = (long long)blockIdx.x * BLK + threadIdx.x;
  if (i < n) {
    double v = (a != 0.) ? a * un[i] + b * us[i] : b * us[i];
    out[i] = v + cc * k[i];
  }
}

void launch_ssp(long long n, double *out, const double *un, double a, const double *us, double b,
                const double *k, double cc, hipStream_t s) {
  hipLaunchKernelGGL(k_ssp, dim3(nblocks(n, BLK)), dim3(BLK), 0, s, n, out, un, a, us, b, k, cc);
}

__global__ void k_rk4_final(long long n, double *c, const double *k1, const double *k2,
                            const double *k3, const double *k4) {
  long long i = (long long)blockIdx.x * BLK + threadIdx.x;
  // rk4 time_integrator.cu:124-130
  if (i < n) c[i] += k1[i] / 6. + k2[i] / 3. + k3[i] / 3. + k4[i] / 6.;
}

void launch_rk4_final(long long n, double *c, const double *k1, const double *k2,
                      const double *k3, const double *k4, hipStream_t s) {
  hipLaunchKernelGGL(k_rk4_final, dim3(nblocks(n, BLK)), dim3(BLK), 0, s, n, c, k1, k2, k3, k4);
}

__global__ void k_copy(long long n, double *dst, const double *src) {
  long long i = (long long)blockIdx.x * BLK + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

void launch_copy(long long n, double *dst, const double *src, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n, BLK)), dim3(BLK), 0, s, n, dst, src);
}

// eval_u conserv_kernels.cu:1284-1307
template <int P>
__global__ void k_eval_u(int K, const double *c, int n, double *u1, double *u2, double *u3) {
  constexpr int NP = np_of(P);
  const Tab<P> &T = tab<P>();
  const int e = blockIdx.x * BLK + threadIdx.x;
  if (e >= K) return;
  double a = 0., b = 0., d = 0.;
  for (int i = 0; i < NP; i++) {
    double v = c[(size_t)(n * NP + i) * K + e];
    a += v * T.VERT[i * 3 + 0];
    b += v * T.VERT[i * 3 + 1];
    d += v * T.VERT[i * 3 + 2];
  }
  u1[e] = a;
  u2[e] = b;
  u3[e] = d;
}

void launch_eval_u(int phys, int p, int K, const double *c, int n, double *u1, double *u2,
                   double *u3, hipStream_t s) {
  dispatch(phys, p, [&](auto, auto PP) {
    constexpr int pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_eval_u<pp>), dim3(nblocks(K, BLK)), dim3(BLK), 0, s, K, c, n, u1, u2,
                       u3);
  });
}

}  // namespace dgk
