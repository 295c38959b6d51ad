// gpe_pair.hip -- fused stage-pair kernels (FP32 forms) of one tile variant
// (see fused.hpp).  Compiled once per GPE_PVID.
#include "fused.hpp"
#include "launch.h"

#ifndef GPE_PVID
#error "GPE_PVID (pair tile variant) must be defined"
#endif
#define GPE_REAL float
#define GPE_PNAME f32
#define GPE_CAT3(a, b, c) a##b##c
#define GPE_PFN(p, v) GPE_CAT3(launch_pair_, p, _v##v)
#define GPE_PFN2(p, v) GPE_PFN(p, v)
#define GPE_PBFN(p, v) GPE_CAT3(launch_pair_batched_, p, _v##v)
#define GPE_PBFN2(p, v) GPE_PBFN(p, v)

namespace gpe {

namespace {
constexpr PairTile kT = kPairTiles[GPE_PVID];

template <int R, int PAIR, bool COMP>
hipError_t launch_one(const StageArgs &A, const StageArgs &B, hipStream_t s) {
  hipLaunchKernelGGL((pair_kernel<GPE_REAL, R, kT.nw, kT.p, kT.d, PAIR, COMP>), dim3(B.nblocks), dim3(64 * kT.nw), 0, s,
                     A, B);
  return hipGetLastError();
}

template <int R, int PAIR, bool COMP>
hipError_t launch_one_b(const PairBatchLaunch &L, hipStream_t s) {
  hipLaunchKernelGGL((pair_kernel_batched<GPE_REAL, R, kT.nw, kT.p, kT.d, PAIR, COMP>), dim3(L.grid), dim3(64 * kT.nw), 0,
                     s, L.instA, L.instB, L.work, L.nwork, L.active, L.pn, L.pe);
  return hipGetLastError();
}

// RK4 stages 1-2 do not depend on the compensated form (as the one-stage
// kernels of those kinds)
template <int R, bool COMP>
hipError_t launch_r(int pair, const StageArgs &A, const StageArgs &B, hipStream_t s) {
  switch (pair) {
    case P_RK4_12: return launch_one<R, P_RK4_12, false>(A, B, s);
    case P_RK4_34: return launch_one<R, P_RK4_34, COMP>(A, B, s);
    case P_CK_12: return launch_one<R, P_CK_12, COMP>(A, B, s);
    case P_CK_34: return launch_one<R, P_CK_34, COMP>(A, B, s);
    case P_CK_56: return launch_one<R, P_CK_56, COMP>(A, B, s);
    default: return hipErrorInvalidValue;
  }
}

template <int R, bool COMP>
hipError_t launch_rb(int pair, const PairBatchLaunch &L, hipStream_t s) {
  switch (pair) {
    case P_RK4_12: return launch_one_b<R, P_RK4_12, false>(L, s);
    case P_RK4_34: return launch_one_b<R, P_RK4_34, COMP>(L, s);
    case P_CK_12: return launch_one_b<R, P_CK_12, COMP>(L, s);
    case P_CK_34: return launch_one_b<R, P_CK_34, COMP>(L, s);
    case P_CK_56: return launch_one_b<R, P_CK_56, COMP>(L, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

hipError_t GPE_PFN2(GPE_PNAME, GPE_PVID)(int pair, int radius, bool comp, const StageArgs &A, const StageArgs &B,
                                         hipStream_t s) {
  if (radius == 1) return comp ? launch_r<1, true>(pair, A, B, s) : launch_r<1, false>(pair, A, B, s);
  if (radius == 2) return comp ? launch_r<2, true>(pair, A, B, s) : launch_r<2, false>(pair, A, B, s);
  return hipErrorInvalidValue;
}

hipError_t GPE_PBFN2(GPE_PNAME, GPE_PVID)(int pair, int radius, bool comp, const PairBatchLaunch &L, hipStream_t s) {
  if (radius == 1) return comp ? launch_rb<1, true>(pair, L, s) : launch_rb<1, false>(pair, L, s);
  if (radius == 2) return comp ? launch_rb<2, true>(pair, L, s) : launch_rb<2, false>(pair, L, s);
  return hipErrorInvalidValue;
}

}  // namespace gpe
