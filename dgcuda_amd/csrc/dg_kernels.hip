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

// The reference takes max_l = lambda[0], then lambda[i] > max_l over all i
// (time_integrator.cu:208-211): a NaN speed counts only when it is element
// 0's.  As an operand of the key max (+NaN above every number) that is:
// element 0's NaN as +NaN, any other NaN as 0.
__device__ __forceinline__ double ref_max_operand(double lam, int e) {
  if (lam != lam) return e == 0 ? __longlong_as_double(0x7ff8000000000000ll) : 0.;
  return lam;
}

// Step reduction, part 1 (every block): the block's max of the CFL speed and
// of the step-to-step change, stored as keys in part[block] and
// part[nblk + block] with plain stores.  The kernel boundary before k_finish
// makes them visible; no block needs a device-wide fence.
template <int B = BLK>
__device__ void block_partials(double lam, double dc, bool conv, unsigned long long *part) {
  __shared__ unsigned long long s_l[B / 64], s_c[B / 64];
  unsigned long long kl = wave_max_key(ord_key(lam));
  unsigned long long kc = conv ? wave_max_key(ord_key(dc)) : 0ull;
  if ((threadIdx.x & 63) == 0) {
    s_l[threadIdx.x >> 6] = kl;
    s_c[threadIdx.x >> 6] = kc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < B / 64; w++) {
      kl = s_l[w] > kl ? s_l[w] : kl;
      kc = s_c[w] > kc ? s_c[w] : kc;
    }
    part[blockIdx.x] = kl;
    if (conv) part[gridDim.x + blockIdx.x] = kc;
  }
}

// Step reduction, part 2 (one block): the max over the partials closes the
// step: dt of the next step from the CFL rule (time_integrator.cu:217) or
// kept, time bookkeeping (time_integrator.cu:226-231), the NaN guard
// (:219-223), the stop at t_end with the last dt shortened to land on it
// (:225-228) and the convergence stop (:403-417).  `initial` marks the call
// before the first step.
__global__ __launch_bounds__(BLK) void k_finish(DevState *st, const unsigned long long *part,
                                                int nblk, FinishArgs f) {
  __shared__ unsigned long long s_l[BLK], s_c[BLK];
  unsigned long long kl = 0ull, kc = 0ull;
  for (int b = threadIdx.x; b < nblk; b += BLK) {
    const unsigned long long l = part[b];
    kl = l > kl ? l : kl;
    if (f.conv) {
      const unsigned long long c = part[nblk + b];
      kc = c > kc ? c : kc;
    }
  }
  s_l[threadIdx.x] = kl;
  s_c[threadIdx.x] = kc;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s_l[threadIdx.x] = s_l[threadIdx.x + w] > s_l[threadIdx.x] ? s_l[threadIdx.x + w]
                                                                 : s_l[threadIdx.x];
      s_c[threadIdx.x] = s_c[threadIdx.x + w] > s_c[threadIdx.x] ? s_c[threadIdx.x + w]
                                                                 : s_c[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (st->nan_flag | st->halt) return;
  const double lmax = ord_val(s_l[0]);
  double dtn;
  if (f.cfl_mode)
    dtn = f.cfl_num / lmax / (2. * f.p + 1.);
  else
    dtn = f.initial ? f.fixed_dt : st->dt;
  if (f.initial) {
    st->tdone = f.t0;
    st->dt_last = 0.;
    if (st->conv_tol >= 0.) {  // the reference starts from conv = 1 (time_integrator.cu:193)
      st->conv = 1.;
      if (!(1. > st->conv_tol)) st->halt = 2;
    }
  } else {
    st->tdone = st->t;
    st->dt_last = st->dt;
    st->steps += 1;
  }
  st->lam_max = lmax;
  if (dtn != dtn) st->nan_flag = 1;
  if (f.conv) {
    const double cv = ord_val(s_c[0]);
    st->conv = cv;
    if (!(cv > st->conv_tol)) st->halt = 2;
  }
  double tn = st->tdone + dtn;
  if (st->tdone >= st->t_end) {
    st->halt = 1;
  } else if (tn > st->t_end) {
    dtn = st->t_end - st->tdone;
    tn = st->t_end;
  }
  st->dt = dtn;
  st->t = tn;
}

void launch_finish(DevState *st, const unsigned long long *part, int nblk, const FinishArgs &f,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLK), 0, s, st, part, nblk, f);
}

// ------------------------------------------------------------ stage

// max over the check points (volume and side quadrature points) of |phi_i|
template <int P>
constexpr double phi_bound(int i) {
  using TC = TabC<P>;
  double m = 0.;
  for (int j = 0; j < TC::NQ; j++) {
    double v = TC::B[i * TC::NQ + j];
    v = v < 0 ? -v : v;
    m = v > m ? v : m;
  }
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < TC::NQ1; j++) {
      double v = TC::SIDE[(k * TC::NP + i) * TC::NQ1 + j];
      v = v < 0 ? -v : v;
      m = v > m ? v : m;
    }
  return m;
}

// Sufficient condition for elem_physical: every point value lies within
// mean +- sum_i |c_i| max|phi_i|; if those bounds keep the density (and for
// Euler the pressure) positive with a wide margin over rounding, every point
// passes is_physical and the point-by-point check can be skipped.
template <int PHYS, int P>
__device__ __forceinline__ bool surely_physical(const PhysParams &ph,
                                                const double (&o)[nvars_of(PHYS)][np_of(P)]) {
  constexpr int NP = np_of(P), N = nvars_of(PHYS);
  constexpr double margin = 1e-10;
  double mean[NMAX], dev[NMAX];
#pragma unroll
  for (int n = 0; n < N; n++) {
    mean[n] = o[n][0] * TabC<P>::B[0];
    double d = 0.;
#pragma unroll
    for (int i = 1; i < NP; i++) d = fma(fabs(o[n][i]), phi_bound<P>(i), d);
    dev[n] = d * (1. + margin);
  }
  const double rmin = mean[0] - dev[0];
  const double rsc = fabs(mean[0]) + dev[0];
  if (!(rmin > margin * rsc)) return false;
  if constexpr (PHYS == PH_EULER) {
    const double mx = fabs(mean[1]) + dev[1], my = fabs(mean[2]) + dev[2];
    const double emin = mean[3] - dev[3];
    const double kin = 0.5 * (mx * mx + my * my) * frcp(rmin);
    return emin - kin > margin * (fabs(mean[3]) + dev[3] + kin);
  }
  return true;
}

// the point check itself (check_physical's test at every volume and side
// quadrature point); loops stay rolled: this path is rare
template <int PHYS, int P>
__device__ __forceinline__ bool elem_physical(const PhysParams &ph,
                                              const double (&o)[nvars_of(PHYS)][np_of(P)]) {
  constexpr int NP = np_of(P), NQ = nq_of(P), NQ1 = P + 1, N = nvars_of(PHYS);
  const Tab<P> &T = tab<P>();
  bool ok = true;
#pragma unroll 1
  for (int j = 0; j < NQ + 3 * NQ1; j++) {
    double U[NMAX];
#pragma unroll
    for (int n = 0; n < N; n++) {
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const double phi = j < NQ ? T.B[i * NQ + j]
                                  : T.SIDE[((j - NQ) / NQ1 * NP + i) * NQ1 + (j - NQ) % NQ1];
        s = fma(o[n][i], phi, s);
      }
      U[n] = s;
    }
    ok = ok && is_physical<PHYS>(ph, U);
  }
  return ok;
}

// the element-local positivity fix: a state that fails the point check keeps
// only its means (check_physical euler.cu:92-116 zeroes the high modes)
template <int PHYS, int P>
__device__ __forceinline__ bool needs_fix(const PhysParams &ph,
                                          const double (&o)[nvars_of(PHYS)][np_of(P)]) {
  if constexpr (has_positivity<PHYS>())
    return !surely_physical<PHYS, P>(ph, o) && !elem_physical<PHYS, P>(ph, o);
  return false;
}

template <int PHYS, int P>
__device__ __forceinline__ void keep_means(bool fix, double (&o)[nvars_of(PHYS)][np_of(P)]) {
  if constexpr (has_positivity<PHYS>()) {
    if (fix) {
#pragma unroll
      for (int n = 0; n < nvars_of(PHYS); n++)
#pragma unroll
        for (int i = 1; i < np_of(P); i++) o[n][i] = 0.;
    }
  }
}

template <int PHYS, int P>
__device__ __forceinline__ void posfix(const PhysParams &ph,
                                       double (&o)[nvars_of(PHYS)][np_of(P)]) {
  keep_means<PHYS, P>(needs_fix<PHYS, P>(ph, o), o);
}

// the coefficients of element e ([N][NP] of the SoA state u)
template <int N, int NP>
__device__ __forceinline__ void load_state(const double *u, int K, int e, double (&x)[N][NP]) {
#pragma unroll
  for (int n = 0; n < N; n++)
#pragma unroll
    for (int i = 0; i < NP; i++) x[n][i] = gld(u + (size_t)(n * NP + i) * K, e);
}

__device__ __forceinline__ int gld_int(const int *p, int i) {
  return ((const __attribute__((address_space(1))) int *)p)[(unsigned)i & kIdxMask];
}

// the positivity fix of a stored state read by this stage: from its flag
// (loaded ahead) when the state carries flags, else by the test
template <int PHYS, int P>
__device__ __forceinline__ void fix_read(const PhysParams &ph, bool flags, int flag,
                                         double (&o)[nvars_of(PHYS)][np_of(P)]) {
  if constexpr (has_positivity<PHYS>()) {
    if (flags)
      keep_means<PHYS, P>(flag != 0, o);
    else
      posfix<PHYS, P>(ph, o);
  }
}

// One RK stage over all elements, one thread per element (see the header
// comment).  The element's memory accesses are software-pipelined: its own
// state, its three side slots and side 0's neighbour are requested at entry;
// while side k is evaluated the neighbour of side k+1 (and during the last
// side the step base and the metric) are in flight.  The registers for that
// come from parking the element's own coefficients in LDS (a per-thread
// column, no sharing) while the sides are evaluated.
// The own coefficients are parked in LDS: 256-thread blocks up to 24
// doubles per element, 128-thread blocks up to 48, so a block's copy stays
// within 48 KB and two waves per SIMD still fit the 160 KB of a CU.
// Larger elements (N * n_p > 48: Euler p >= 4, the 3-variable systems at
// p = 5) run 64-thread blocks and stream: the neighbour's coefficients are
// folded into its traces as they arrive instead of being held, and the own
// traces are formed from the parked copy one mode at a time, so no second
// coefficient set is ever live beside the accumulators.
template <int PHYS, int P>
__host__ __device__ constexpr bool stage_stream() {
  return nvars_of(PHYS) * np_of(P) > 48;
}
template <int PHYS, int P>
__host__ __device__ constexpr bool stage_park() {
  return true;
}
template <int PHYS, int P>
__host__ __device__ constexpr int stage_blk() {
  return stage_stream<PHYS, P>() ? 64 : (nvars_of(PHYS) * np_of(P) <= 24 ? 256 : 128);
}

template <int PHYS, int P>
__global__ __launch_bounds__((stage_blk<PHYS, P>())) void k_stage(const StageArgs a) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, NQ1 = TC::NQ1, N = nvars_of(PHYS);
  constexpr int B = stage_blk<PHYS, P>();
  constexpr bool kPark = stage_park<PHYS, P>();
  constexpr bool kStream = stage_stream<PHYS, P>();
  // requests one side ahead only while a second neighbour state fits the
  // register budget (N * n_p <= 24 doubles); larger elements load at use
  constexpr bool kPipe = N * NP <= 24;
  __shared__ double s_side[3 * NP * NQ1];
  __shared__ double s_c[kPark ? N * NP : 1][B];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3 * NP * NQ1; q++) s_side[q] = TC::SIDE[q];
  }
  __syncthreads();
  DevState *st = a.st;
  if (st->nan_flag | st->halt) return;
  const int K = a.m.K;
  const int e = remap_block(blockIdx.x, a.nblk) * B + threadIdx.x;
  const bool act = e < K;
  const int ee = act ? e : K - 1;
  const double dt = st->dt;
  const bool flags = a.flags_in != 0;
  const bool use_un = a.accmode == ACC_FINAL6 || a.mode == M_A_UN_B_US_CK || a.mode == M_UN_CK;

  // ---- requested at entry: the slots (first: loads complete in order, and
  // the neighbour index waits only for them), own state, side 0's neighbour
  int2 sl[3];
  int nbk[3];
#pragma unroll
  for (int k = 0; k < 3; k++) sl[k] = a.m.slot[(size_t)k * K + ee];
  double c[N][NP];
  load_state(a.us, K, ee, c);
#pragma unroll
  for (int k = 0; k < 3; k++) nbk[k] = sl[k].x >= 0 ? sl[k].x : ee;  // boundary: a stand-in
  // flags are read without branches (the array is always valid) and before
  // the neighbour, so the wait for the own state's flag leaves it in flight
  int fl_me = 0, fl_nb[3] = {0, 0, 0};
  if constexpr (has_positivity<PHYS>()) {
    fl_me = gld_int(a.bad, ee);
#pragma unroll
    for (int k = 0; k < 3; k++) fl_nb[k] = gld_int(a.bad, nbk[k]);
  }
  double cn[kStream ? 1 : N][kStream ? 1 : NP];
  if constexpr (!kStream) load_state(a.us, K, nbk[0], cn);
  double eps_me = 1.;
  if constexpr (PHYS == PH_TE) eps_me = gld(a.m.eps, ee);
  // advection whose velocity varies in space: the element's vertices give
  // the physical points where it is evaluated
  double Vv[6] = {0., 0., 0., 0., 0., 0.};
  const bool vary = PHYS == PH_ADV && a.ph.prob == PB_ADV_ROTHILL;
  if constexpr (PHYS == PH_ADV) {
    if (vary) {
#pragma unroll
      for (int q = 0; q < 6; q++) Vv[q] = gld(a.m.V + (size_t)q * K, ee);
    }
  }

  if (a.fix_in) fix_read<PHYS, P>(a.ph, flags, fl_me, c);
  if constexpr (kPark) {
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) s_c[n * NP + i][threadIdx.x] = c[n][i];
    asm volatile("" ::: "memory");  // the copy in registers is dead from here
  }

  double acc[N][NP];
#pragma unroll
  for (int n = 0; n < N; n++)
#pragma unroll
    for (int i = 0; i < NP; i++) acc[n][i] = 0.;
  double ub[N][NP];
  double xr = 0., yr = 0., xs = 0., ys = 0., Jm = 1.;

  // ---- surface integrals (eval_surface + eval_left_right conserv_kernels.cu:850-985);
  // slot k is the element's local side k
#pragma unroll
  for (int k = 0; k < 3; k++) {
    __builtin_amdgcn_sched_barrier(0);
    const int nb = sl[k].x, nbls = sl[k].y & 3, role = (sl[k].y >> 2) & 1;
    // the next requests go out before this side's arithmetic
    double cx[N][NP];
    if (k < 2) {
      if constexpr (kPipe) load_state(a.us, K, nbk[k + 1], cx);
    } else {
      if constexpr (kPipe) {
        if (use_un) load_state(a.un, K, ee, ub);
      }
      xr = gld(a.m.xr, ee);
      yr = gld(a.m.yr, ee);
      xs = gld(a.m.xs, ee);
      ys = gld(a.m.ys, ee);
      Jm = gld(a.m.J, ee);
    }
    const int sid = sl[k].y >> 3;
    const double nx = gld(a.m.snx, sid), ny = gld(a.m.sny, sid), len = gld(a.m.slen, sid);
    double eps_nb = eps_me;
    if constexpr (PHYS == PH_TE) {
      if (nb >= 0) eps_nb = gld(a.m.eps, nb);
    }
    double Un[NQ1][NMAX];
    if (nb >= 0) {
      // neighbour trace; the neighbour walks the side the other way, so its
      // point NQ1-1-jj is our point jj
      if constexpr (kStream) {
        // modes folded in as they arrive; a neighbour whose stored state
        // fails the positivity test contributes its means only (its flag:
        // the host gives every stage that fixes its input valid flags)
        const bool hi = !(a.fix_in && fl_nb[k]);
#pragma unroll
        for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
          for (int n = 0; n < N; n++) Un[jj][n] = 0.;
#pragma unroll
        for (int i = 0; i < NP; i++) {
          double ph[NQ1];
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++) ph[jj] = s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj];
#pragma unroll
          for (int n = 0; n < N; n++) {
            double v = gld(a.us + (size_t)(n * NP + i) * K, nb);
            if (i > 0) v = hi ? v : 0.;
#pragma unroll
            for (int jj = 0; jj < NQ1; jj++) Un[jj][n] = fma(v, ph[jj], Un[jj][n]);
          }
        }
      } else {
        if (a.fix_in) fix_read<PHYS, P>(a.ph, flags, fl_nb[k], cn);
#pragma unroll
        for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
          for (int n = 0; n < N; n++) {
            double s = cn[n][0] * s_side[(nbls * NP) * NQ1 + NQ1 - 1 - jj];
#pragma unroll
            for (int i = 1; i < NP; i++)
              s = fma(cn[n][i], s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj], s);
            Un[jj][n] = s;
          }
      }
    }
    // own trace
    double Umk[NQ1][NMAX];
    if constexpr (kStream) {
#pragma unroll
      for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
        for (int n = 0; n < N; n++) Umk[jj][n] = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++)
#pragma unroll
        for (int n = 0; n < N; n++) {
          const double v = s_c[n * NP + i][threadIdx.x];
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++)
            Umk[jj][n] = fma(v, TC::SIDE[(k * NP + i) * NQ1 + jj], Umk[jj][n]);
        }
    } else {
      double cc[N][NP];
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) cc[n][i] = kPark ? s_c[n * NP + i][threadIdx.x] : c[n][i];
#pragma unroll
      for (int jj = 0; jj < NQ1; jj++)
#pragma unroll
        for (int n = 0; n < N; n++) {
          double s = cc[n][0] * TC::SIDE[(k * NP) * NQ1 + jj];
#pragma unroll
          for (int i = 1; i < NP; i++) s = fma(cc[n][i], TC::SIDE[(k * NP + i) * NQ1 + jj], s);
          Umk[jj][n] = s;
        }
    }
    const double *bv =
        a.m.bnd_val + (size_t)(nb < 0 ? a.m.side[sid].w : 0) * (NQ1 * NMAX);
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      double *Um = Umk[jj];
      if (nb < 0) {
        // boundary side (this element is its left): wall ghost state from the
        // precomputed wall vector, or the precomputed Dirichlet state
        if (nb == -1) {
          wall_state<PHYS>(a.ph, Um, Un[jj], bv[jj * NMAX], bv[jj * NMAX + 1]);
        } else {
#pragma unroll
          for (int n = 0; n < N; n++) Un[jj][n] = bv[jj * NMAX + n];
        }
      }
      double F[NMAX], f;
      if constexpr (llf_only<PHYS>()) {
        // LLF is antisymmetric to the last bit, F(Ub, Ua, -n) = -F(Ua, Ub, n)
        // (every operation in it negates exactly), so each element takes
        // itself as the left state with its outward normal: the right
        // element's contribution equals the reference's +F(UL, UR, n) term
        // bit for bit, without selecting the states per point
        const double onx = role ? -nx : nx, ony = role ? -ny : ny;
        riemann<PHYS>(a.ph, Um, Un[jj], onx, ony, eps_me, eps_nb, F);
        f = -0.5 * len * TC::W1[jj];
      } else {
        double UL[NMAX], UR[NMAX];
#pragma unroll
        for (int n = 0; n < N; n++) {
          UL[n] = role ? Un[jj][n] : Um[n];
          UR[n] = role ? Um[n] : Un[jj][n];
        }
        const double eL = role ? eps_nb : eps_me, eR = role ? eps_me : eps_nb;
        PhysParams pv = a.ph;
        if constexpr (PHYS == PH_ADV) {
          // the velocity at the side point (the reference takes it at the
          // left element's point; this element's point jj is the same point)
          if (vary) {
            double x, y;
            side_point<NQ1>(Vv, TC::X1, k, jj, x, y);
            rothill_velocity(x, y, pv.ax, pv.ay);
          }
        }
        riemann<PHYS>(pv, UL, UR, nx, ny, eL, eR, F);
        f = (role ? 0.5 : -0.5) * len * TC::W1[jj];
      }
#pragma unroll
      for (int n = 0; n < N; n++) F[n] *= f;
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const double phi = TC::SIDE[(k * NP + i) * NQ1 + jj];
#pragma unroll
        for (int n = 0; n < N; n++) acc[n][i] = fma(F[n], phi, acc[n][i]);
      }
    }
    pin(acc);
    if (k < 2) {
      if constexpr (kPipe) {
#pragma unroll
        for (int n = 0; n < N; n++)
#pragma unroll
          for (int i = 0; i < NP; i++) cn[n][i] = cx[n][i];
      } else if constexpr (!kStream) {
        load_state(a.us, K, nbk[k + 1], cn);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);

  // ---- volume integral (eval_volume conserv_kernels.cu:1042-1081)
  if constexpr (kPark) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) c[n][i] = s_c[n * NP + i][threadIdx.x];
  }
#pragma unroll
  for (int j = 0; j < NQ; j++) {
    double U[NMAX];
#pragma unroll
    for (int n = 0; n < N; n++) {
      double s = c[n][0] * TC::B[j];
#pragma unroll
      for (int i = 1; i < NP; i++) s = fma(c[n][i], TC::B[i * NQ + j], s);
      U[n] = s;
    }
    double ga[NMAX], gb[NMAX];
    PhysParams pv = a.ph;
    if constexpr (PHYS == PH_ADV) {
      if (vary) {
        double x, y;
        elem_point(Vv, TC::R1[j], TC::R2[j], x, y);
        rothill_velocity(x, y, pv.ax, pv.ay);
      }
    }
    flux_contra<PHYS>(pv, U, xr, yr, xs, ys, ga, gb);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const double gx = TC::GX[i * NQ + j], gy = TC::GY[i * NQ + j];
#pragma unroll
      for (int n = 0; n < N; n++) acc[n][i] = fma(gx, ga[n], fma(gy, gb[n], acc[n][i]));
    }
  }
  pin(acc);
  __builtin_amdgcn_sched_barrier(0);

  // k = dt/J * residual (eval_rhs time_integrator.cu:26-101)
  const double sdt = dt / Jm;
  double sc[NMAX];
#pragma unroll
  for (int n = 0; n < N; n++) sc[n] = sdt;
  if constexpr (PHYS == PH_TE) {
    sc[0] = sdt * (1. / eps_me);
    sc[1] = sc[0];
    sc[2] = sdt * (1. / a.ph.mu);
  }
#pragma unroll
  for (int n = 0; n < N; n++)
#pragma unroll
    for (int i = 0; i < NP; i++) acc[n][i] *= sc[n];
  if constexpr (PHYS == PH_TE) {
    if (a.kout && act) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) gst(a.kout + (size_t)(n * NP + i) * K, e, acc[n][i]);
    }
  }

  // RK stage combination (rk_tempstorage / rk4 time_integrator.cu:107-130;
  // Shu-Osher forms); the branches are uniform.  The step base un was
  // requested during the last side.
  double (&o)[N][NP] = c;
  if constexpr (!kPipe) {
    if (use_un) load_state(a.un, K, ee, ub);
  }
  if (use_un && a.fix_un) fix_read<PHYS, P>(a.ph, flags, fl_me, ub);
  if (a.accmode == ACC_FINAL6) {
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const size_t q = (size_t)(n * NP + i) * K;
        o[n][i] = ub[n][i] + (gld(a.acc + q, ee) + acc[n][i] * (1. / 6.));
      }
  } else {
    if (a.accmode == ACC_SET6 && act) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++)
          gst(a.acc + (size_t)(n * NP + i) * K, e, acc[n][i] * (1. / 6.));
    } else if (a.accmode == ACC_ADD3 && act) {
      double ra[N][NP];  // all loads before the first store (which may alias them)
      load_state(a.acc, K, e, ra);
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++)
          gst(a.acc + (size_t)(n * NP + i) * K, e, ra[n][i] + acc[n][i] * (1. / 3.));
    }
    if (a.mode == M_US_CK) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) o[n][i] = o[n][i] + a.C * acc[n][i];
    } else if (a.mode == M_A_UN_B_US_CK) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) o[n][i] = (a.A * ub[n][i] + a.B * o[n][i]) + a.C * acc[n][i];
    } else if (a.mode == M_UN_CK) {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) o[n][i] = ub[n][i] + a.C * acc[n][i];
    } else {
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) o[n][i] = acc[n][i];
    }
  }
  // an intermediate stage state is only read as the next stage's input: fix
  // it once here; a step's final state is written as computed and fixed
  // where it is read next (fix_in / fix_un), as the reference fixes the
  // state it is about to evaluate
  if (a.fix_out) posfix<PHYS, P>(a.ph, o);
  if constexpr (has_positivity<PHYS>()) {
    // a step's final state: record whether it needs the fix when read
    if (a.last_stage && a.flags_out) {
      const bool b = needs_fix<PHYS, P>(a.ph, o);
      if (act) ((__attribute__((address_space(1))) int *)a.bad)[(unsigned)e & kIdxMask] = b ? 1 : 0;
    }
  }
  // step-to-step change before the store (out may alias un)
  // (the reference's norm covers the first K*N state entries: rows
  // n * n_p + i < N, time_integrator.cu:404-414)
  double dc = 0.;
  if (a.conv && act) {
#pragma unroll
    for (int r = 0; r < N && r < N * NP; r++)
      dc = fmax(dc, fabs(o[r / NP][r % NP] - gld(a.un + (size_t)r * K, ee)));
  }
  if (act) {
#pragma unroll
    for (int n = 0; n < N; n++)
#pragma unroll
      for (int i = 0; i < NP; i++) gst(a.out + (size_t)(n * NP + i) * K, e, o[n][i]);
  }
  if (a.last_stage) {
    double lam = 0.;
    if (a.cfl_mode && act) {
      double U[NMAX];
#pragma unroll
      for (int n = 0; n < N; n++) U[n] = o[n][0] * TC::B[0];
      PhysParams pv = a.ph;
      if constexpr (PHYS == PH_ADV) {
        // eval_global_lambda takes the velocity at quadrature point n_q / 2
        // (advection.cu:176-181)
        if (vary) {
          double x, y;
          elem_point(Vv, TC::R1[NQ / 2], TC::R2[NQ / 2], x, y);
          rothill_velocity(x, y, pv.ax, pv.ay);
        }
      }
      lam = ref_max_operand(cfl_speed<PHYS>(pv, U, eps_me), e);
    }
    block_partials<B>(lam, dc, a.conv != 0, a.part);
  }
}

void launch_stage(int phys, int p, const StageArgs &a0, hipStream_t s) {
  StageArgs a = a0;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    constexpr int B = stage_blk<ph, pp>();
    a.nblk = nblocks(a.m.K, B);
    hipLaunchKernelGGL((k_stage<ph, pp>), dim3(a.nblk), dim3(B), 0, s, a);
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

int stage_vgprs(int phys, int p) {
  int v = -1;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph = decltype(PH)::value, pp = decltype(PP)::value;
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&k_stage<ph, pp>)) ==
        hipSuccess)
      v = attr.numRegs;
  });
  return v;
}

// ------------------------------------------------------------ setup kernels

// boundary data of every boundary slot (eval_boundary conserv_kernels.cu:820-843)
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_bnd(const MeshDev m, const PhysParams ph,
                                             const DevState *st, double cstage, int use_state,
                                             double t0) {
  using TC = TabC<P>;
  constexpr int NQ1 = TC::NQ1;
  const int b = blockIdx.x * BLK + threadIdx.x;
  if (b >= m.nb_bnd) return;
  const double t = use_state ? st->t + cstage * st->dt : t0;
  const int K = m.K, e = m.bnd_elem[b], cd = m.bnd_code[b], k = cd & 3, code = cd >> 2;
  const double V0 = m.V[e], V1 = m.V[(size_t)K + e], V2 = m.V[2 * (size_t)K + e];
  const double V3 = m.V[3 * (size_t)K + e], V4 = m.V[4 * (size_t)K + e],
               V5 = m.V[5 * (size_t)K + e];
  const int sid = m.slot[(size_t)k * K + e].y >> 3;
  const double nx = m.snx[sid], ny = m.sny[sid];
  for (int jj = 0; jj < NQ1; jj++) {
    double ra, rb;
    if (k == 0) {
      ra = 0.5 + 0.5 * TC::X1[jj];
      rb = 0.;
    } else if (k == 1) {
      ra = (1. - TC::X1[jj]) / 2.;
      rb = (1. + TC::X1[jj]) / 2.;
    } else {
      ra = 0.;
      rb = 0.5 + 0.5 * TC::X1[NQ1 - 1 - jj];
    }
    const double X = V2 * ra + V4 * rb + V0 * (1 - ra - rb);
    const double Y = V3 * ra + V5 * rb + V1 * (1 - ra - rb);
    boundary_data<PHYS>(ph, X, Y, t, nx, ny, code, m.bnd_val + ((size_t)b * NQ1 + jj) * NMAX);
  }
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
  if (part) block_partials(lam, dc, cprev != nullptr, part);
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
                                               int *bad) {
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
}

void launch_flags(int phys, int p, int K, const PhysParams &ph, const double *c, int *bad,
                  hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    hipLaunchKernelGGL((k_flags<ph_, pp>), dim3(nblocks(K, BLK)), dim3(BLK), 0, s, K, ph, c, bad);
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
