// dg_ensemble.hip -- one launch per RK stage for an ensemble of independent
// meshes (no counterpart in the reference, which runs one mesh per process).
//
// The instances of an ensemble are merged by polynomial order into groups:
// group p holds the elements of every instance of order p, each instance's
// range padded to a multiple of 256 elements, with the instance's
// neighbour, side and boundary indices offset into the group's arrays.  A
// stage is ONE launch over all groups: blocks [gb[p], gb[p+1]) run the stage
// body of order p on group p's arrays.  Every block lies inside one
// instance, so its step-reduction partial belongs to that instance alone;
// k_finish_ens then closes the step of every instance (one block each) with
// the instance's own CFL dt.  The per-element arithmetic is the one-mesh
// stage kernel's, so an instance evolves as it would alone (up to how the
// compiler contracts a multiply-add in the merged kernel: a few ulps).
#include "dg_stage.h"
#include "dg_stage_q.h"

namespace dgk {

template <int PHYS, int P, int B>
__host__ __device__ constexpr int ens_lds_of() {
  return 3 * np_of(P) * (P + 1) + nvars_of(PHYS) * np_of(P) * B;
}
template <int PHYS, int B>
__host__ __device__ constexpr int ens_lds() {
  int m = 0;
  m = ens_lds_of<PHYS, 0, B>() > m ? ens_lds_of<PHYS, 0, B>() : m;
  m = ens_lds_of<PHYS, 1, B>() > m ? ens_lds_of<PHYS, 1, B>() : m;
  m = ens_lds_of<PHYS, 2, B>() > m ? ens_lds_of<PHYS, 2, B>() : m;
  m = ens_lds_of<PHYS, 3, B>() > m ? ens_lds_of<PHYS, 3, B>() : m;
  return m;
}

template <int PHYS, int P, int B>
__device__ __forceinline__ void ens_group(const EnsStageArgs &ea, double *lds) {
  constexpr int side = 3 * np_of(P) * (P + 1);
  stage_body<PHYS, P, B, true>(ea.a[P], blockIdx.x - ea.gb[P], ea.gb[P + 1] - ea.gb[P], lds,
                               reinterpret_cast<double(*)[B]>(lds + side));
}

// One RK stage of every instance.  The branch on the block's group is
// uniform per block; the LDS is the largest group's.
template <int PHYS, int B>
__global__ __launch_bounds__(B) void k_stage_ens(const EnsStageArgs ea) {
  __shared__ double lds[ens_lds<PHYS, B>()];
  const int b = blockIdx.x;
  if (b < ea.gb[1])
    ens_group<PHYS, 0, B>(ea, lds);
  else if (b < ea.gb[2])
    ens_group<PHYS, 1, B>(ea, lds);
  else if (b < ea.gb[3])
    ens_group<PHYS, 2, B>(ea, lds);
  else
    ens_group<PHYS, 3, B>(ea, lds);
}

// the orders and block sizes an ensemble group runs at (orders 4 and 5 in
// 64-element blocks; Euler there with the four-lane kernel)
template <int PHYS, int P, int B>
__host__ __device__ constexpr bool ens_order_ok() {
  return P <= 3 || B == 64;
}

// One RK stage of one group only (the split form: a launch per order, each
// with that order's register budget and LDS instead of the largest order's)
template <int PHYS, int P, int B>
__global__ __launch_bounds__(B) void k_stage_ens_g(const EnsStageArgs ea) {
  __shared__ double lds[ens_lds_of<PHYS, P, B>()];
  constexpr int side = 3 * np_of(P) * (P + 1);
  stage_body<PHYS, P, B, true>(ea.a[P], blockIdx.x, ea.gb[P + 1] - ea.gb[P], lds,
                               reinterpret_cast<double(*)[B]>(lds + side));
}

// CFL speed and positivity flags of the current state of one group, with
// the partials in 256-element blocks (each inside one instance); the NaN
// rule of the reference's max applies to each instance's own element 0
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_lambda_ens(const MeshDev m, const PhysParams ph,
                                                    const double *c, const int4 *blk,
                                                    unsigned long long *part, int *bad) {
  constexpr int NP = np_of(P), N = nvars_of(PHYS);
  using TC = TabC<P>;
  const int K = m.K;
  const int4 bi = blk[blockIdx.x];
  const int e0 = __builtin_amdgcn_readfirstlane(bi.y), eend = __builtin_amdgcn_readfirstlane(bi.z);
  const int e = blockIdx.x * BLK + threadIdx.x;
  double lam = 0.;
  if (e < eend) {
    double o[N][NP];
    for (int n = 0; n < N; n++)
      for (int i = 0; i < NP; i++) o[n][i] = c[(size_t)(n * NP + i) * K + e];
    if constexpr (has_positivity<PHYS>()) bad[e] = needs_fix<PHYS, P>(ph, o) ? 1 : 0;
    double U[NMAX];
    for (int n = 0; n < N; n++) U[n] = o[n][0] * TC::B[0];
    lam = ref_max_operand(cfl_speed<PHYS>(ph, U, PHYS == PH_TE ? m.eps[e] : 1.), e - e0);
  }
  block_partials<BLK>(lam, 0., false, part, blockIdx.x, gridDim.x);
}

// boundary data of a group's boundary slots at their own instances' stage
// times (time-dependent boundary problems): slot b belongs to instance
// inst[b], with its problem parameters ph[inst[b]] and its DevState
template <int PHYS, int P>
__global__ __launch_bounds__(BLK) void k_bnd_ens(const MeshDev m, const PhysParams *ph,
                                                 const int *inst, const DevState *st,
                                                 double cstage) {
  const int b = blockIdx.x * BLK + threadIdx.x;
  if (b >= m.nb_bnd) return;
  const int i = inst[b];
  const DevState *s = st + i;
  bnd_slot<PHYS, P>(m, ph[i], s->t + cstage * s->dt, b);
}

void launch_bnd_ens(int phys, int p, const MeshDev &m, const PhysParams *ph, const int *inst,
                    const DevState *st, double cstage, hipStream_t s) {
  if (m.nb_bnd <= 0) return;
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    if constexpr (ens_order_ok<ph_, pp, 64>())
      hipLaunchKernelGGL((k_bnd_ens<ph_, pp>), dim3(nblocks(m.nb_bnd, BLK)), dim3(BLK), 0, s, m,
                         ph, inst, st, cstage);
  });
}

// closes the step of every instance: block i reduces instance i's partials
// (blocks of parts.bs[g] elements of its group g) and updates its DevState
__global__ __launch_bounds__(BLK) void k_finish_ens(DevState *st, const EnsInst *inst,
                                                    EnsParts parts, FinishArgs f) {
  const EnsInst in = inst[blockIdx.x];
  const int bs = parts.bs[in.g];
  const int b0 = in.e0 / bs, b1 = (in.e1 + bs - 1) / bs;
  FinishArgs fi = f;
  fi.cfl_num = in.cfl_num;
  fi.p = in.p;
  DevState *s = st + blockIdx.x;
  finish_body(s, parts.part[in.g] + b0, nullptr, b1 - b0, fi, s->tdone);
}

void launch_stage_ens(int phys, int B, const EnsStageArgs &ea, int nblocks_total,
                      hipStream_t s) {
  auto go = [&](auto PH, auto BB) {
    constexpr int ph = decltype(PH)::value, bb = decltype(BB)::value;
    hipLaunchKernelGGL((k_stage_ens<ph, bb>), dim3(nblocks_total), dim3(bb), 0, s, ea);
  };
  auto by_b = [&](auto PH) {
    if (B == 64)
      go(PH, std::integral_constant<int, 64>{});
    else if (B == 128)
      go(PH, std::integral_constant<int, 128>{});
    else
      go(PH, std::integral_constant<int, 256>{});
  };
  switch (phys) {
    case PH_ADV: by_b(std::integral_constant<int, PH_ADV>{}); break;
    case PH_EULER: by_b(std::integral_constant<int, PH_EULER>{}); break;
    case PH_SWE: by_b(std::integral_constant<int, PH_SWE>{}); break;
    case PH_TM: by_b(std::integral_constant<int, PH_TM>{}); break;
    default: by_b(std::integral_constant<int, PH_TE>{}); break;
  }
}

// the split form of one stage: one launch per non-empty order group, group
// p in blocks of bs[p] elements
void launch_stage_ens_split(int phys, const int *bs, const EnsStageArgs &ea, hipStream_t s) {
  auto by_b = [&](auto PH) {
    constexpr int ph = decltype(PH)::value;
    auto one = [&](auto PP) {
      constexpr int pp = decltype(PP)::value;
      const int n = ea.gb[pp + 1] - ea.gb[pp];
      if (n <= 0) return;
      // orders 4 and 5 run in 64-element blocks only (their parked
      // coefficients fill the LDS of larger blocks); no Euler instance has them
      auto go = [&](auto BB) {
        constexpr int bb = decltype(BB)::value;
        if constexpr (stage_quad<ph, pp>()) {
          // large Euler elements: four lanes per element, 64 elements per block
          if constexpr (bb == kQuadBlk / 4)
            hipLaunchKernelGGL((k_stage_q<ph, pp, false, false, true>), dim3(n), dim3(kQuadBlk),
                               0, s, ea.a[pp]);
        } else if constexpr (ens_order_ok<ph, pp, bb>()) {
          hipLaunchKernelGGL((k_stage_ens_g<ph, pp, bb>), dim3(n), dim3(bb), 0, s, ea);
        }
      };
      if (bs[pp] == 64)
        go(std::integral_constant<int, 64>{});
      else if (bs[pp] == 128)
        go(std::integral_constant<int, 128>{});
      else
        go(std::integral_constant<int, 256>{});
    };
    one(std::integral_constant<int, 0>{});
    one(std::integral_constant<int, 1>{});
    one(std::integral_constant<int, 2>{});
    one(std::integral_constant<int, 3>{});
    one(std::integral_constant<int, 4>{});
    one(std::integral_constant<int, 5>{});
  };
  switch (phys) {
    case PH_ADV: by_b(std::integral_constant<int, PH_ADV>{}); break;
    case PH_EULER: by_b(std::integral_constant<int, PH_EULER>{}); break;
    case PH_SWE: by_b(std::integral_constant<int, PH_SWE>{}); break;
    case PH_TM: by_b(std::integral_constant<int, PH_TM>{}); break;
    default: by_b(std::integral_constant<int, PH_TE>{}); break;
  }
}

void launch_lambda_ens(int phys, int p, const MeshDev &m, const PhysParams &ph, const double *c,
                       const int4 *blk256, unsigned long long *part, int *bad, hipStream_t s) {
  dispatch(phys, p, [&](auto PH, auto PP) {
    constexpr int ph_ = decltype(PH)::value, pp = decltype(PP)::value;
    if constexpr (ens_order_ok<ph_, pp, 64>())
      hipLaunchKernelGGL((k_lambda_ens<ph_, pp>), dim3(m.K / BLK), dim3(BLK), 0, s, m, ph, c,
                         blk256, part, bad);
  });
}

void launch_finish_ens(DevState *st, const EnsInst *inst, int ninst, const EnsParts &parts,
                       const FinishArgs &f, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_ens, dim3(ninst), dim3(BLK), 0, s, st, inst, parts, f);
}

int ens_vgprs(int phys, int B) {
  int v = -1;
  auto get = [&](const void *fn) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) v = attr.numRegs;
  };
  auto by_b = [&](auto PH) {
    constexpr int ph = decltype(PH)::value;
    if (B == 64)
      get(reinterpret_cast<const void *>(&k_stage_ens<ph, 64>));
    else if (B == 128)
      get(reinterpret_cast<const void *>(&k_stage_ens<ph, 128>));
    else
      get(reinterpret_cast<const void *>(&k_stage_ens<ph, 256>));
  };
  switch (phys) {
    case PH_ADV: by_b(std::integral_constant<int, PH_ADV>{}); break;
    case PH_EULER: by_b(std::integral_constant<int, PH_EULER>{}); break;
    case PH_SWE: by_b(std::integral_constant<int, PH_SWE>{}); break;
    case PH_TM: by_b(std::integral_constant<int, PH_TM>{}); break;
    default: by_b(std::integral_constant<int, PH_TE>{}); break;
  }
  return v;
}

}  // namespace dgk
