// dg_stage_q.h -- the RK stage for large Euler elements (p >= 4) with four
// lanes per element, one per conserved variable.
//
// One thread per element needs all N * n_p coefficients of the element and
// of its accumulators in registers (Euler p = 4: 2 x 60 doubles), which
// leaves one wave per SIMD.  Here lane n of each group of four holds only
// variable n: its coefficients, its accumulators and its traces, and every
// contraction (interpolation to the points, projection onto the test
// functions) runs on the lane's own variable.  The flux is spread over the
// quad: the points are taken in groups of four and lane l evaluates the
// whole flux at point l of each group, so no flux is evaluated twice.  A 4 x
// 4 transpose within the quad through LDS brings every variable of point l to
// lane l, and another takes the flux components back to their variables'
// lanes.  The memory accesses stay coalesced per variable row.
#pragma once

#include "dg_stage.h"

namespace dgk {

// Euler p >= 4 (at p = 3 the one-thread kernel holds its coefficients in
// registers at two waves per SIMD and needs no exchanges)
template <int PHYS, int P>
__host__ __device__ constexpr bool stage_quad() {
  return PHYS == PH_EULER && P >= 4;
}
constexpr int kQuadBlk = 256;  // threads per block: 64 elements

// lane M's value in every lane of its quad
template <int M>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = M | (M << 2) | (M << 4) | (M << 6);
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), ctrl, 0xF,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), ctrl, 0xF,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ void quad_gather(double v, double *U) {
  U[0] = quad_bcast<0>(v);
  U[1] = quad_bcast<1>(v);
  U[2] = quad_bcast<2>(v);
  U[3] = quad_bcast<3>(v);
}
__device__ __forceinline__ double pick4(const double *F, int n) {
  return n == 0 ? F[0] : (n == 1 ? F[1] : (n == 2 ? F[2] : F[3]));
}

// 4 x 4 transpose within each quad through the thread's LDS row: lane l
// holds x[0..3] (row l); afterwards y[v] is row v's x[l].  A quad lies in
// one wave, so wavefront-scope ordering suffices: the first fence keeps the
// rows' earlier reads before the rewrite, the second makes the writes
// visible to the quad's other lanes before they read.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ void quad_transpose(double (*s_tr)[5], const double (&x)[4],
                                               double (&y)[4]) {
  const int t = threadIdx.x, l = t & 3, base = t & ~3;
  wave_sync();
#pragma unroll
  for (int j = 0; j < 4; j++) s_tr[t][j] = x[j];
  wave_sync();
#pragma unroll
  for (int v = 0; v < 4; v++) y[v] = s_tr[base + v][l];
}

// needs_fix of the one-thread kernel, each lane holding variable n's modes
template <int PHYS, int P>
__device__ __forceinline__ bool needs_fix_q(const PhysParams &ph, const double (&o)[np_of(P)]) {
  constexpr int NP = np_of(P), NQ = nq_of(P), NQ1 = P + 1;
  using T = TabC<P>;
  constexpr double margin = 1e-10;
  double d = 0.;
#pragma unroll
  for (int i = 1; i < NP; i++) d = fma(fabs(o[i]), phi_bound<P>(i), d);
  double mean[NMAX], dev[NMAX];
  quad_gather(o[0] * T::B[0], mean);
  quad_gather(d * (1. + margin), dev);
  bool sure = false;
  const double rmin = mean[0] - dev[0];
  const double rsc = fabs(mean[0]) + dev[0];
  if (rmin > margin * rsc) {
    const double mx = fabs(mean[1]) + dev[1], my = fabs(mean[2]) + dev[2];
    const double emin = mean[3] - dev[3];
    const double kin = 0.5 * (mx * mx + my * my) * frcp(rmin);
    sure = emin - kin > margin * (fabs(mean[3]) + dev[3] + kin);
  }
  if (sure) return false;
  bool ok = true;
#pragma unroll 1
  for (int j = 0; j < NQ + 3 * NQ1; j++) {
    double s = 0.;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const double phi = j < NQ ? T::B[i * NQ + j]
                                : T::SIDE[((j - NQ) / NQ1 * NP + i) * NQ1 + (j - NQ) % NQ1];
      s = fma(o[i], phi, s);
    }
    double U[NMAX];
    quad_gather(s, U);
    ok = ok && is_physical<PHYS>(ph, U);
  }
  return !ok;
}

// Lane n's rows of a [4][NP][K] state: mode i of variable n at element e is
// base_i[n * NP * K + e] with base_i = u + i * K, a uniform (scalar) base per
// mode and one 32-bit lane offset for all of them (the host keeps
// 4 * NP * K below 2^28), instead of a 64-bit address per load.
template <int NP>
__device__ __forceinline__ void load_rows(const double *u, int K, int off, double (&x)[NP]) {
#pragma unroll
  for (int i = 0; i < NP; i++) x[i] = gld(u + (size_t)i * K, off);
}
template <int NP>
__device__ __forceinline__ void store_rows_nt(double *u, int K, int off, const double (&x)[NP]) {
#pragma unroll
  for (int i = 0; i < NP; i++) gst_nt(u + (size_t)i * K, off, x[i]);
}

// One side point's flux for lane n: states Uo (own) and Ux (neighbour, or
// the ghost state built here on a boundary side) of every variable; the
// element takes itself as the left state with its outward normal (the LLF
// flux is antisymmetric to the last bit).  Returns the four components
// scaled by the point's surface weight -len/2 w_j.
template <int PHYS>
__device__ __forceinline__ void side_flux_q(const PhysParams &ph, int nb, const double *bvj,
                                            double onx, double ony, double f, double (&Uo)[4],
                                            double (&Ux)[4], double (&F)[4]) {
  if (nb < 0) {
    if (nb == -1) {
      wall_state<PHYS>(ph, Uo, Ux, bvj[0], bvj[1], bvj[2]);
    } else {
#pragma unroll
      for (int m = 0; m < NMAX; m++) Ux[m] = bvj[m];
    }
  }
  riemann<PHYS>(ph, Uo, Ux, onx, ony, 1., 1., F);
#pragma unroll
  for (int m = 0; m < 4; m++) F[m] *= f;
}

// Stored traces, laid out so that the 64 lanes of a wave (16 elements x 4
// variables) store 64 consecutive values: points in pairs,
// [side][pair][K][variable][2], 16 bytes per lane, and an odd last point in
// its own plane [side][K][variable] after them.
template <int P>
struct TrLayout {
  static constexpr int NQ1 = P + 1, NPR = NQ1 / 2;
  static __device__ __forceinline__ int pair(int K, int k, int q, int e, int n) {
    return ((k * NPR + q) * K + e) * 8 + n * 2;
  }
  static __device__ __forceinline__ int last(int K, int k, int e, int n) {
    return 24 * NPR * K + (k * K + e) * 4 + n;
  }
};

// Lane n's traces of its variable (coefficients o) on the element's three
// sides, for the neighbours to read: the same multiply-add chain, in the same
// order, as a reader that folds the modes itself.
template <int P>
__device__ __forceinline__ void store_traces(double *tr, int K, int e, int n,
                                             const double (&o)[np_of(P)]) {
  using TC = TabC<P>;
  using TL = TrLayout<P>;
  constexpr int NP = TC::NP, NQ1 = TC::NQ1;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double t[NQ1];
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) s = fma(o[i], TC::SIDE[(k * NP + i) * NQ1 + jj], s);
      t[jj] = s;
    }
#pragma unroll
    for (int q = 0; q < TL::NPR; q++) gst2_nt(tr, TL::pair(K, k, q, e, n), dbl2{t[2 * q], t[2 * q + 1]});
    if constexpr (NQ1 % 2 == 1) gst_nt(tr, TL::last(K, k, e, n), t[NQ1 - 1]);
  }
}

// the traces of a state that no stage wrote (the start of a run)
template <int PHYS, int P>
__global__ __launch_bounds__(kQuadBlk) void k_traces_q(const double *u, double *tr, int K) {
  constexpr int NP = np_of(P);
  const int e = blockIdx.x * (kQuadBlk / 4) + (threadIdx.x >> 2), n = threadIdx.x & 3;
  if (e >= K) return;
  double o[NP];
  load_rows(u, K, n * NP * K + e, o);
  store_traces<P>(tr, K, e, n, o);
}

// the volume integral on the matrix cores (v_mfma_f64_16x16x4f64): at p = 4
// the n_q = 16 points and n_p = 15 modes (padded to 16) make each
// contraction one 16 x 16 x 16 product per tile of 16 (element, variable)
// columns
template <int PHYS, int P>
__host__ __device__ constexpr bool volume_mfma_ok() {
  return stage_quad<PHYS, P>() && P == 4;
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// The MFMA variant is held to three waves per SIMD (164 registers, no
// spills): left free, it took 156 VGPRs + 16 AGPRs and ran two waves, 1.49
// against 1.36 ms per stage at p = 4 on 2,097,152 triangles.
// With ENS the mesh is an ensemble's order group (one launch for the group,
// 64-element logical blocks): a.ens[block] gives the block's instance, whose
// DevState it reads, the instance's first element and its end, as in the
// one-thread body; an ensemble runs the gather path (no stored traces).
template <int PHYS, int P, bool VM, bool TR, bool ENS = false>
__global__ __launch_bounds__(kQuadBlk, ((VM && P == 4) ? 3 : 1)) void k_stage_q(const StageArgs a) {
  using TC = TabC<P>;
  constexpr int NP = TC::NP, NQ = TC::NQ, NQ1 = TC::NQ1;
  constexpr int EB = kQuadBlk / 4, N = 4;
  constexpr bool kVM = VM && volume_mfma_ok<PHYS, P>();
  __shared__ double s_side[3 * NP * NQ1];
  __shared__ double s_tr[kQuadBlk][5];  // transpose rows (padded against bank conflicts)
  // with one leftover point per side (NQ1 = 4 k + 1) the three sides' last
  // points form one more group after the side loop: each lane's traces
  // there wait in its row, and the fluxes come back through the same rows
  constexpr bool kLast = NQ1 % 4 == 1;
  // MFMA volume: each wave stages one tile of 16 columns (16 modes each) at
  // a time in its own slab, which also holds its lanes' leftover-point rows
  // during the side phase (a wave never touches another wave's slab), and
  // its 16 elements' metric terms; s_a holds every lane's A operands
  constexpr int kSlab = 64 * 6;
  __shared__ double s_wv[kVM ? (kQuadBlk / 64) * kSlab : 1];
  __shared__ double s_met[kVM ? kQuadBlk : 1];
  __shared__ double s_a[kVM ? 3 * 4 * 64 : 1];
  __shared__ double s_lo_[(kLast && !kVM) ? kQuadBlk * 6 : 1];
  // thread t's six leftover-point values: [side][own, neighbour] or the flux
  auto lo_row = [&](int t) -> double * {
    if constexpr (kVM)
      return s_wv + (t >> 6) * kSlab + (t & 63) * 6;
    else
      return s_lo_ + t * 6;
  };
  for (int q = threadIdx.x; q < 3 * NP * NQ1; q += kQuadBlk) s_side[q] = TC::SIDE[q];
  if constexpr (kVM) {
    // A operands of step st for lane l (row l % 16, k l / 16): the basis at
    // point l % 16 of mode l / 16 + 4 st; the test-function gradients of
    // mode l % 16 at point l / 16 + 4 st (zero in the padding mode 15)
    for (int q = threadIdx.x; q < 4 * 64; q += kQuadBlk) {
      const int st = q >> 6, lr = q & 15, lk = (q & 63) >> 4, m = lk + 4 * st;
      s_a[q] = m < NP ? TC::B[m * NQ + lr] : 0.;
      s_a[256 + q] = lr < NP ? TC::GX[lr * NQ + m] : 0.;
      s_a[512 + q] = lr < NP ? TC::GY[lr * NQ + m] : 0.;
    }
  }
  __syncthreads();
  const int K = a.m.K;
  const int lb = remap_block(blockIdx.x, ENS ? (int)gridDim.x : a.nblk);
  int inst = 0, e0 = 0, eend = K;
  if constexpr (ENS) {  // block-uniform values: kept in scalar registers
    const int4 bi = a.ens[lb];
    inst = __builtin_amdgcn_readfirstlane(bi.x);
    e0 = __builtin_amdgcn_readfirstlane(bi.y);
    eend = __builtin_amdgcn_readfirstlane(bi.z);
  }
  DevState *st = a.st + inst;
  if (st->nan_flag | st->halt) return;
  const int n = threadIdx.x & 3;
  const int e = lb * EB + (threadIdx.x >> 2);
  const bool act = e < eend;
  const int ee = act ? e : eend - 1;
  const double dt = st->dt;
  const bool use_un = a.accmode == ACC_FINAL6 || a.mode == M_A_UN_B_US_CK || a.mode == M_UN_CK;

  int2 sl[3];
  int nbk[3];
#pragma unroll
  for (int k = 0; k < 3; k++) sl[k] = a.m.slot[(size_t)k * K + ee];
  const int rk = n * NP * K;  // lane n's row offset
  double c[NP];
  load_rows(a.us, K, rk + ee, c);
#pragma unroll
  for (int k = 0; k < 3; k++) nbk[k] = sl[k].x >= 0 ? sl[k].x : ee;
  // positivity on read: the stored states carry flags (the host computes
  // them for every input that is fixed on read)
  int fl_me = 0, fl_nb[3] = {0, 0, 0};
  fl_me = gld_int(a.bad, ee);
#pragma unroll
  for (int k = 0; k < 3; k++) fl_nb[k] = gld_int(a.bad, nbk[k]);
  if (a.fix_in && fl_me) {
#pragma unroll
    for (int i = 1; i < NP; i++) c[i] = 0.;
  }
  double acc[NP];
#pragma unroll
  for (int i = 0; i < NP; i++) acc[i] = 0.;

  // ---- surface integrals.  A side's data is requested at the side's
  // start, the own trace is formed while it arrives, and the neighbour's
  // coefficients are requested five at a time and folded into its trace
  // (three round trips per side; a boundary side requests the element's own
  // row as a stand-in and folds in nothing).  The offsets pass through an
  // opaque asm at those points and the trace is pinned after each batch:
  // without that the compiler requests every side's neighbour at the
  // kernel's start and holds them all in registers (245 VGPRs with the
  // first side's loads issued one at a time, against 147 here: three waves
  // per SIMD instead of two).
  double xr = 0., yr = 0., xs = 0., ys = 0.;
  double lnx = 0., lny = 0., llen = 0.;  // the leftover group's side (lane n's)
  // stored traces: side k's neighbour trace is requested before side k - 1's
  // flux (side 0's here); a boundary side requests the element's own as a
  // stand-in and uses none of it
  double ptv[NQ1];
  auto fetch_tr = [&](int kk) {
    if constexpr (TR) {
      using TL = TrLayout<P>;
      const int nbt = sl[kk].x, nbl = sl[kk].y & 3;
      int nbo = nbt >= 0 ? nbt : ee;
      asm volatile("" : "+v"(nbo));
#pragma unroll
      for (int q = 0; q < TL::NPR; q++) {
        const dbl2 v = gld2(a.tr_in, TL::pair(K, nbl, q, nbo, n));
        ptv[2 * q] = v[0];
        ptv[2 * q + 1] = v[1];
      }
      if constexpr (NQ1 % 2 == 1) ptv[NQ1 - 1] = gld(a.tr_in, TL::last(K, nbl, nbo, n));
    }
  };
  fetch_tr(0);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    __builtin_amdgcn_sched_barrier(0);
    const int nb = sl[k].x, nbls = sl[k].y & 3, role = (sl[k].y >> 2) & 1;
    int sid = sl[k].y >> 3, off = rk + nbk[k];
    asm volatile("" : "+v"(off), "+v"(sid));
    const double nx = gld(a.m.snx, sid), ny = gld(a.m.sny, sid), len = gld(a.m.slen, sid);
    if (k == 2) {
      int me = ee;
      asm volatile("" : "+v"(me));
      xr = gld(a.m.xr, me);
      yr = gld(a.m.yr, me);
      xs = gld(a.m.xs, me);
      ys = gld(a.m.ys, me);
      if constexpr (kLast) {
        // the leftover group's side for lane n (its normal and length are
        // requested here, a side ahead of their use); the choice among the
        // slots is made on opaque copies (see below)
        const int ks = n < 3 ? n : 2;
        int s0 = sl[0].y, s1 = sl[1].y, s2 = sl[2].y;
        asm volatile("" : "+v"(s0), "+v"(s1), "+v"(s2));
        int sidl = (ks == 0 ? s0 : (ks == 1 ? s1 : s2)) >> 3;
        asm volatile("" : "+v"(sidl));
        lnx = gld(a.m.snx, sidl);
        lny = gld(a.m.sny, sidl);
        llen = gld(a.m.slen, sidl);
      }
    }
    double Um[NQ1], Un[NQ1];
#pragma unroll
    for (int jj = 0; jj < NQ1; jj++) {
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) s = fma(c[i], TC::SIDE[(k * NP + i) * NQ1 + jj], s);
      Um[jj] = s;
      Un[jj] = 0.;
    }
    if constexpr (TR) {
      // the neighbour's trace on the shared side as its stage stored it
      // (points in reverse order); a neighbour whose stored state fails the
      // positivity test contributes its mean only, as in the gather below;
      // boundary sides keep zeros (their states come from the boundary data)
      if (nb >= 0) {
        if (a.fix_in && fl_nb[k]) {
          const double v0 = gld(a.us, off);
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++) Un[jj] = fma(v0, s_side[(nbls * NP) * NQ1 + NQ1 - 1 - jj], 0.);
        } else {
#pragma unroll
          for (int jj = 0; jj < NQ1; jj++) Un[jj] = ptv[NQ1 - 1 - jj];
        }
      }
      pin(Un);
      if (k < 2) fetch_tr(k + 1);
    } else {
      // a neighbour whose stored state fails the positivity test
      // contributes its means only
      const bool has = nb >= 0, hi = !(a.fix_in && fl_nb[k]);
#pragma unroll
      for (int i0 = 0; i0 < NP; i0 += 5) {
        int o2 = off;
        asm volatile("" : "+v"(o2));
        double vb[5];
        // the last batch may be partial (n_p = 21 at p = 5): only modes
        // below n_p are requested or used
#pragma unroll
        for (int b = 0; b < 5; b++)
          if (i0 + b < NP) vb[b] = gld(a.us + (size_t)(i0 + b) * K, o2);
#pragma unroll
        for (int b = 0; b < 5; b++) {
          const int i = i0 + b;
          if (i < NP) {
            const double v = has && (i == 0 || hi) ? vb[b] : 0.;
#pragma unroll
            for (int jj = 0; jj < NQ1; jj++)
              Un[jj] = fma(v, s_side[(nbls * NP + i) * NQ1 + NQ1 - 1 - jj], Un[jj]);
          }
        }
        pin(Un);
      }
    }
    const double *bv =
        a.m.bnd_val + (size_t)(nb < 0 ? a.m.side[sid].w : 0) * (NQ1 * NMAX);
    const double onx = role ? -nx : nx, ony = role ? -ny : ny;
    // groups of four points: lane n evaluates point jj0 + n
    constexpr int NG = NQ1 / 4;
#pragma unroll
    for (int g4 = 0; g4 < NG; g4++) {
      const int jj0 = 4 * g4, pt = jj0 + n;
      double x[4], Uo[4], Ux[4], F[4], Fn[4];
#pragma unroll
      for (int j = 0; j < 4; j++) x[j] = Um[jj0 + j];
      quad_transpose(s_tr, x, Uo);
#pragma unroll
      for (int j = 0; j < 4; j++) x[j] = Un[jj0 + j];
      quad_transpose(s_tr, x, Ux);
      double w1[4];
#pragma unroll
      for (int j = 0; j < 4; j++) w1[j] = TC::W1[jj0 + j];
      side_flux_q<PHYS>(a.ph, nb, bv + pt * NMAX, onx, ony, -0.5 * len * pick4(w1, n), Uo, Ux, F);
      quad_transpose(s_tr, F, Fn);
#pragma unroll
      for (int v = 0; v < 4; v++)
#pragma unroll
        for (int i = 0; i < NP; i++) acc[i] = fma(Fn[v], TC::SIDE[(k * NP + i) * NQ1 + jj0 + v], acc[i]);
    }
    if constexpr (kLast) {
      lo_row(threadIdx.x)[k * 2 + 0] = Um[NQ1 - 1];
      lo_row(threadIdx.x)[k * 2 + 1] = Un[NQ1 - 1];
    } else {
      // leftover points: every lane evaluates the whole flux, keeps its component
#pragma unroll
      for (int jj = 4 * NG; jj < NQ1; jj++) {
        double Uo[4], Ux[4], F[4];
        quad_gather(Um[jj], Uo);
        quad_gather(Un[jj], Ux);
        side_flux_q<PHYS>(a.ph, nb, bv + jj * NMAX, onx, ony, -0.5 * len * TC::W1[jj], Uo, Ux, F);
        const double Fv = pick4(F, n);
#pragma unroll
        for (int i = 0; i < NP; i++) acc[i] = fma(Fv, TC::SIDE[(k * NP + i) * NQ1 + jj], acc[i]);
      }
    }
    pin(acc);
  }
  if constexpr (kLast) {
    // the sides' last points: lane ks < 3 evaluates side ks's (lane 3
    // repeats side 2; its result is not read)
    constexpr int jl = NQ1 - 1;
    const int ks = n < 3 ? n : 2, base = threadIdx.x & ~3;
    wave_sync();
    double Uo[4], Ux[4], F[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      Uo[v] = lo_row(base + v)[ks * 2 + 0];
      Ux[v] = lo_row(base + v)[ks * 2 + 1];
    }
    // the per-lane choice among the three slots, made on copies that pass
    // through an opaque asm: a select among sl[] itself compiles to an
    // indexed load, which puts sl[] in scratch memory and turns every later
    // use of a slot into a memory round trip
    int s0x = sl[0].x, s0y = sl[0].y, s1x = sl[1].x, s1y = sl[1].y, s2x = sl[2].x, s2y = sl[2].y;
    asm volatile("" : "+v"(s0x), "+v"(s0y), "+v"(s1x), "+v"(s1y), "+v"(s2x), "+v"(s2y));
    const int nb = ks == 0 ? s0x : (ks == 1 ? s1x : s2x);
    const int sly = ks == 0 ? s0y : (ks == 1 ? s1y : s2y);
    const int role = (sly >> 2) & 1, sid = sly >> 3;
    const double *bv =
        a.m.bnd_val + (size_t)(nb < 0 ? a.m.side[sid].w : 0) * (NQ1 * NMAX);
    side_flux_q<PHYS>(a.ph, nb, bv + jl * NMAX, role ? -lnx : lnx, role ? -lny : lny,
                      -0.5 * llen * TC::W1[jl], Uo, Ux, F);
    wave_sync();
#pragma unroll
    for (int m = 0; m < 4; m++) lo_row(threadIdx.x)[m] = F[m];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double Fv = lo_row(base + k)[n];
#pragma unroll
      for (int i = 0; i < NP; i++) acc[i] = fma(Fv, TC::SIDE[(k * NP + i) * NQ1 + jl], acc[i]);
    }
    pin(acc);
  }
  __builtin_amdgcn_sched_barrier(0);

  // ---- volume integral: groups of four points, lane n evaluates point j0 + n.
  // The step base is requested first: it is read only by the combination.
  double ub[NP];
  int offu = rk + ee;
  asm volatile("" : "+v"(offu));
  if (use_un) load_rows(a.un, K, offu, ub);
  if constexpr (kVM) {
    // Matrix-core volume integral.  Wave columns: lane l holds column l =
    // (element l / 4 of the wave, variable l % 4); tile t is columns 16 t ..
    // 16 t + 15.  v_mfma_f64_16x16x4f64 operand layout: A[row l % 16][k l / 16],
    // B[k l / 16][col l % 16], D[row l / 16 + 4 r][col l % 16] (r = 0..3).
    // Interpolation U[q][col] = sum_i phi_i(q) c_i[col]: four k-steps over
    // the modes.  Projection acc_i[col] += sum_q GX[i][q] ga[q][col] +
    // GY[i][q] gb[q][col]: the D layout's points l / 16 + 4 r serve as
    // k-step r, so the flux outputs enter the second product where they lie.
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    double *sw = s_wv + w * kSlab;  // the tile's [16 columns][16 modes]
    double *sm = s_met + w * 64;    // [16 elements][xr, yr, xs, ys]
    wave_sync();  // the side phase's reads of the slab are done
    if (n == 0) {
      sm[(l >> 2) * 4 + 0] = xr;
      sm[(l >> 2) * 4 + 1] = yr;
      sm[(l >> 2) * 4 + 2] = xs;
      sm[(l >> 2) * 4 + 3] = ys;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (lk == t) {  // the tile's own lanes stage their coefficients
#pragma unroll
        for (int i = 0; i < NP; i++) sw[lr * 16 + i] = c[i];
        sw[lr * 16 + 15] = 0.;
      }
      wave_sync();
      dbl4 D = {0., 0., 0., 0.};
#pragma unroll
      for (int st = 0; st < 4; st++)
        D = __builtin_amdgcn_mfma_f64_16x16x4f64(s_a[st * 64 + l], sw[lr * 16 + lk + 4 * st], D, 0,
                                                 0, 0);
      // lane l: variable n of element 4 t + lr / 4 at points lk + 4 r; after
      // the transpose every variable at point lk + 4 n
      double x[4] = {D[0], D[1], D[2], D[3]}, U[4], ga[4], gb[4], gan[4], gbn[4];
      quad_transpose(s_tr, x, U);
      const int el = 4 * t + (lr >> 2);
      flux_contra<PHYS>(a.ph, U, sm[el * 4 + 0], sm[el * 4 + 1], sm[el * 4 + 2], sm[el * 4 + 3],
                        ga, gb);
      quad_transpose(s_tr, ga, gan);  // variable n's components at points lk + 4 r
      quad_transpose(s_tr, gb, gbn);
      // two independent chains (the GY and GX halves), four k-steps each
      dbl4 Ey = {0., 0., 0., 0.}, Ex = {0., 0., 0., 0.};
#pragma unroll
      for (int st = 0; st < 4; st++) {
        Ey = __builtin_amdgcn_mfma_f64_16x16x4f64(s_a[512 + st * 64 + l], gbn[st], Ey, 0, 0, 0);
        Ex = __builtin_amdgcn_mfma_f64_16x16x4f64(s_a[256 + st * 64 + l], gan[st], Ex, 0, 0, 0);
      }
      const dbl4 E = Ex + Ey;
      // lane l: modes lk + 4 r of column lr; the staged coefficients have
      // been read, so the slab takes the result, and the tile's own lanes
      // add their column to their accumulators
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; r++) sw[lr * 16 + lk + 4 * r] = E[r];
      wave_sync();
      if (lk == t) {
#pragma unroll
        for (int i = 0; i < NP; i++) acc[i] += sw[lr * 16 + i];
      }
      wave_sync();
    }
  } else {
  constexpr int NGV = NQ / 4;
#pragma unroll
  for (int g4 = 0; g4 < NGV; g4++) {
    const int j0 = 4 * g4;
    double xu[4], U[4], ga[4], gb[4], gan[4], gbn[4];
#pragma unroll
    for (int jl = 0; jl < 4; jl++) {
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NP; i++) s = fma(c[i], TC::B[i * NQ + j0 + jl], s);
      xu[jl] = s;
    }
    quad_transpose(s_tr, xu, U);
    flux_contra<PHYS>(a.ph, U, xr, yr, xs, ys, ga, gb);
    quad_transpose(s_tr, ga, gan);
    quad_transpose(s_tr, gb, gbn);
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int i = 0; i < NP; i++)
        acc[i] = fma(TC::GX[i * NQ + j0 + v], gan[v], fma(TC::GY[i * NQ + j0 + v], gbn[v], acc[i]));
  }
  // leftover points: every lane evaluates the flux and keeps its component
#pragma unroll
  for (int j = 4 * NGV; j < NQ; j++) {
    double s = 0.;
#pragma unroll
    for (int i = 0; i < NP; i++) s = fma(c[i], TC::B[i * NQ + j], s);
    double U[NMAX], ga[NMAX], gb[NMAX];
    quad_gather(s, U);
    flux_contra<PHYS>(a.ph, U, xr, yr, xs, ys, ga, gb);
    const double gan = pick4(ga, n), gbn = pick4(gb, n);
#pragma unroll
    for (int i = 0; i < NP; i++)
      acc[i] = fma(TC::GX[i * NQ + j], gan, fma(TC::GY[i * NQ + j], gbn, acc[i]));
  }
  }
  pin(acc);
  __builtin_amdgcn_sched_barrier(0);

  // k = dt/J * residual; J as the host forms it
  const double Jm = __dsub_rn(__dmul_rn(xr, ys), __dmul_rn(xs, yr));
  const double sdt = dt / Jm;
#pragma unroll
  for (int i = 0; i < NP; i++) acc[i] *= sdt;

  // RK stage combination, variable n's rows
  if (use_un && a.fix_un && fl_me) {
#pragma unroll
    for (int i = 1; i < NP; i++) ub[i] = 0.;
  }
  double(&o)[NP] = c;
  if (a.accmode == ACC_FINAL6) {
#pragma unroll
    for (int i = 0; i < NP; i++)
      o[i] = ub[i] + (gld(a.acc + (size_t)i * K, rk + ee) + acc[i] * (1. / 6.));
  } else {
    if (a.accmode == ACC_SET6 && act) {
#pragma unroll
      for (int i = 0; i < NP; i++) gst_nt(a.acc + (size_t)i * K, rk + e, acc[i] * (1. / 6.));
    } else if (a.accmode == ACC_ADD3 && act) {
      double ra[NP];  // all loads before the first store
#pragma unroll
      for (int i = 0; i < NP; i++) ra[i] = gld(a.acc + (size_t)i * K, rk + e);
#pragma unroll
      for (int i = 0; i < NP; i++) gst_nt(a.acc + (size_t)i * K, rk + e, ra[i] + acc[i] * (1. / 3.));
    }
    if (a.mode == M_US_CK) {
#pragma unroll
      for (int i = 0; i < NP; i++) o[i] = o[i] + a.C * acc[i];
    } else if (a.mode == M_A_UN_B_US_CK) {
#pragma unroll
      for (int i = 0; i < NP; i++) o[i] = (a.A * ub[i] + a.B * o[i]) + a.C * acc[i];
    } else if (a.mode == M_UN_CK) {
#pragma unroll
      for (int i = 0; i < NP; i++) o[i] = ub[i] + a.C * acc[i];
    } else {
#pragma unroll
      for (int i = 0; i < NP; i++) o[i] = acc[i];
    }
  }
  if (a.fix_out) {
    if (needs_fix_q<PHYS, P>(a.ph, o)) {
#pragma unroll
      for (int i = 1; i < NP; i++) o[i] = 0.;
      if (act && n == 0 && a.fired) atomicOr(a.fired, 1);
    }
  }
  if (a.last_stage && a.flags_out) {
    const bool b = needs_fix_q<PHYS, P>(a.ph, o);
    if (act && n == 0) ((__attribute__((address_space(1))) int *)a.bad)[(unsigned)e & kIdxMask] = b ? 1 : 0;
    if (b && act && n == 0 && a.fired) atomicOr(a.fired, 1);
  }
  // step-to-step change over rows r < N (variable 0, modes 0..N-1)
  double dc = 0.;
  if (a.conv && act && n == 0) {
#pragma unroll
    for (int i = 0; i < N && i < NP; i++)
      dc = fmax(dc, fabs(o[i] - gld(a.un + (size_t)i * K, ee)));
  }
  if (act) store_rows_nt(a.out, K, rk + e, o);
  if constexpr (TR) {
    if (a.tr_out && act) store_traces<P>(a.tr_out, K, e, n, o);
  }
  if (a.last_stage) {
    double lam = 0.;
    double U[NMAX];
    quad_gather(o[0] * TC::B[0], U);
    if (a.cfl_mode && act) lam = ref_max_operand(cfl_speed<PHYS>(a.ph, U, 1.), e - e0);
    // an ensemble's partials by logical block (each instance's blocks are
    // contiguous there), one mesh's by the hardware block index
    if constexpr (ENS)
      block_partials<kQuadBlk>(lam, dc, a.conv != 0, a.part, lb, gridDim.x);
    else
      block_partials<kQuadBlk>(lam, dc, a.conv != 0, a.part, blockIdx.x, gridDim.x);
  }
}

}  // namespace dgk
