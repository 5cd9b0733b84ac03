// dg_stage.h -- the per-element RK stage (device code shared by the
// one-mesh stage kernel in dg_kernels.hip and the ensemble kernel in
// dg_ensemble.hip): step-reduction partials, the element-local positivity
// fix and the stage body itself.  Everything here is a device function or a
// template; the tables are the compile-time TabC<P>.
#pragma once

#include "dg_device.h"
#include "dg_internal.h"
#include "dg_physics.h"
#include "dg_tabvals.h"

namespace dgk {

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
__device__ void block_partials(double lam, double dc, bool conv, unsigned long long *part,
                               int idx, int count) {
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
    part[idx] = kl;
    if (conv) part[count + idx] = kc;
  }
}

// boundary data of boundary slot b at time t, at the side's n_q1 points
// (eval_boundary conserv_kernels.cu:820-843)
template <int PHYS, int P>
__device__ __forceinline__ void bnd_slot(const MeshDev &m, const PhysParams &ph, double t, int b) {
  using TC = TabC<P>;
  constexpr int NQ1 = TC::NQ1;
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
  using T = TabC<P>;
  bool ok = true;
#pragma unroll 1
  for (int j = 0; j < NQ + 3 * NQ1; j++) {
    double U[NMAX];
#pragma unroll
    for (int n = 0; n < N; n++) {
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const double phi = j < NQ ? T::B[i * NQ + j]
                                  : T::SIDE[((j - NQ) / NQ1 * NP + i) * NQ1 + (j - NQ) % NQ1];
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

// The body runs one block of B elements: block pblock of a grid of nblk
// blocks (remapped so each XCD gets a contiguous element range), with the
// LDS for the side table and the parked coefficients passed in.  With ENS
// the mesh is an ensemble's merged mesh and a.ens[block] gives
// its instance (whose DevState it reads), the instance's first element and
// its end; elements from the end to the block's end are padding.
template <int PHYS, int P, int B, bool ENS = false>
__device__ __forceinline__ void stage_body(const StageArgs a, int pblock, int nblk,
                                           double *s_side, double (*s_c)[B]) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, NQ1 = TC::NQ1, N = nvars_of(PHYS);
  constexpr bool kPark = stage_park<PHYS, P>();
  constexpr bool kStream = stage_stream<PHYS, P>();
  // requests one side ahead only while a second neighbour state fits the
  // register budget (N * n_p <= 24 doubles); larger elements load at use
  constexpr bool kPipe = N * NP <= 24;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3 * NP * NQ1; q++) s_side[q] = TC::SIDE[q];
  }
  __syncthreads();
  const int lb = remap_block(pblock, nblk);
  int inst = 0, e0 = 0, eend = a.m.K;
  if constexpr (ENS) {  // block-uniform values: kept in scalar registers
    const int4 bi = a.ens[lb];
    inst = __builtin_amdgcn_readfirstlane(bi.x);
    e0 = __builtin_amdgcn_readfirstlane(bi.y);
    eend = __builtin_amdgcn_readfirstlane(bi.z);
  }
  DevState *st = a.st + inst;
  if (st->nan_flag | st->halt) return;
  const int K = a.m.K;
  const int e = lb * B + threadIdx.x;
  const bool act = e < eend;
  const int ee = act ? e : eend - 1;
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
          wall_state<PHYS>(a.ph, Um, Un[jj], bv[jj * NMAX], bv[jj * NMAX + 1], bv[jj * NMAX + 2]);
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
  // J = xr ys - xs yr, rounded as the host forms it (preval_jacobian
  // conserv_kernels.cu:452-473: two products and a difference, no fused
  // multiply-add), so it is not read from memory
  Jm = __dsub_rn(__dmul_rn(xr, ys), __dmul_rn(xs, yr));
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
    if (a.kout && act) {  // the E rows (the only ones the eps gradient reads)
#pragma unroll
      for (int n = 0; n < 2; n++)
#pragma unroll
        for (int i = 0; i < NP; i++) gst_nt(a.kout + (size_t)(n * NP + i) * K, e, acc[n][i]);
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
          gst_nt(a.acc + (size_t)(n * NP + i) * K, e, acc[n][i] * (1. / 6.));
    } else if (a.accmode == ACC_ADD3 && act) {
      double ra[N][NP];  // all loads before the first store (which may alias them)
      load_state(a.acc, K, e, ra);
#pragma unroll
      for (int n = 0; n < N; n++)
#pragma unroll
        for (int i = 0; i < NP; i++)
          gst_nt(a.acc + (size_t)(n * NP + i) * K, e, ra[n][i] + acc[n][i] * (1. / 3.));
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
  if (a.fix_out) {
    const bool f = needs_fix<PHYS, P>(a.ph, o);
    keep_means<PHYS, P>(f, o);
    if (f && act && a.fired) atomicOr(a.fired, 1);
  }
  if constexpr (has_positivity<PHYS>()) {
    // a step's final state: record whether it needs the fix when read
    if (a.last_stage && a.flags_out) {
      const bool b = needs_fix<PHYS, P>(a.ph, o);
      if (act) ((__attribute__((address_space(1))) int *)a.bad)[(unsigned)e & kIdxMask] = b ? 1 : 0;
      if (b && act && a.fired) atomicOr(a.fired, 1);
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
      for (int i = 0; i < NP; i++) gst_nt(a.out + (size_t)(n * NP + i) * K, e, o[n][i]);
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
      lam = ref_max_operand(cfl_speed<PHYS>(pv, U, eps_me), e - e0);
    }
    // an ensemble's partials are indexed by logical block (each instance's
    // blocks are contiguous there); one mesh's by the hardware block index,
    // which keeps the remapped index out of the scalar registers until here
    if constexpr (ENS)
      block_partials<B>(lam, dc, a.conv != 0, a.part, lb, nblk);
    else
      block_partials<B>(lam, dc, a.conv != 0, a.part, blockIdx.x, gridDim.x);
  }
}

// Step reduction, part 2 (one block): the max over the partials closes the
// step: dt of the next step from the CFL rule (time_integrator.cu:217) or
// kept, time bookkeeping (time_integrator.cu:226-231), the NaN guard
// (:219-223), the stop at t_end with the last dt shortened to land on it
// (:225-228) and the convergence stop (:403-417).  `initial` marks the call
// before the first step, which starts at t0.  pl / pc: the nblk partial keys
// of the CFL speed and of the step-to-step change (pc: nullptr without it).
__device__ __forceinline__ void finish_body(DevState *st, const unsigned long long *pl,
                                            const unsigned long long *pc, int nblk,
                                            const FinishArgs &f, double t0) {
  // maxima of the partials: each thread requests up to kU of them at once
  // (one round trip for up to kU * BLK blocks), then a shuffle reduction
  // per wave and one barrier
  constexpr int kU = 8, kW = BLK / 64;
  __shared__ unsigned long long s_l[kW], s_c[kW];
  unsigned long long kl = 0ull, kc = 0ull;
  const bool with_c = pc != nullptr;
  // thread 0 requests the step state together with its partials (one round
  // trip for both), and writes the fields back at the end
  DevState S;
  if (threadIdx.x == 0) S = *st;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += BLK * kU) {
    unsigned long long l[kU], c[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int b = b0 + u * BLK;
      l[u] = b < nblk ? pl[b] : 0ull;
      c[u] = (with_c && b < nblk) ? pc[b] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      kl = l[u] > kl ? l[u] : kl;
      kc = c[u] > kc ? c[u] : kc;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ol = __shfl_xor(kl, o), oc = __shfl_xor(kc, o);
    kl = ol > kl ? ol : kl;
    kc = oc > kc ? oc : kc;
  }
  if ((threadIdx.x & 63) == 0) {
    s_l[threadIdx.x >> 6] = kl;
    s_c[threadIdx.x >> 6] = kc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int w = 1; w < kW; w++) {
    s_l[0] = s_l[w] > s_l[0] ? s_l[w] : s_l[0];
    s_c[0] = s_c[w] > s_c[0] ? s_c[w] : s_c[0];
  }
  if (S.nan_flag | S.halt) return;
  const double lmax = ord_val(s_l[0]);
  double dtn;
  if (f.cfl_mode)
    dtn = f.cfl_num / lmax / (2. * f.p + 1.);
  else
    dtn = f.initial ? f.fixed_dt : S.dt;
  if (f.initial) {
    S.tdone = t0;
    S.dt_last = 0.;
    if (S.conv_tol >= 0.) {  // the reference starts from conv = 1 (time_integrator.cu:193)
      S.conv = 1.;
      if (!(1. > S.conv_tol)) S.halt = 2;
    }
  } else {
    S.tdone = S.t;
    S.dt_last = S.dt;
    S.steps += 1;
  }
  S.lam_max = lmax;
  if (dtn != dtn) S.nan_flag = 1;
  if (f.conv) {
    const double cv = ord_val(s_c[0]);
    S.conv = cv;
    if (!(cv > S.conv_tol)) S.halt = 2;
  }
  double tn = S.tdone + dtn;
  if (S.tdone >= S.t_end) {
    S.halt = 1;
  } else if (tn > S.t_end) {
    dtn = S.t_end - S.tdone;
    tn = S.t_end;
  }
  S.dt = dtn;
  S.t = tn;
  *st = S;
}

}  // namespace dgk
