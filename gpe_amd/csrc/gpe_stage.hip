// gpe_stage.hip -- stage kernels of one precision and one tile variant.
// Compiled once per (GPE_F64 or not, GPE_VID); the FP64 build uses
// -ffp-contract=off so every operation is one IEEE rounding, in the
// reference's evaluation order.
#include "launch.h"
#include "stencil.hpp"

#ifndef GPE_VID
#error "GPE_VID (tile variant) must be defined"
#endif
#ifdef GPE_F64
#define GPE_REAL double
#define GPE_PNAME f64
#else
#define GPE_REAL float
#define GPE_PNAME f32
#endif
#define GPE_CAT3(a, b, c) a##b##c
#define GPE_FN(p, v) GPE_CAT3(launch_stage_, p, _v##v)
#define GPE_FN2(p, v) GPE_FN(p, v)
#define GPE_BFN(p, v) GPE_CAT3(launch_batched_, p, _v##v)
#define GPE_BFN2(p, v) GPE_BFN(p, v)

namespace gpe {

namespace {
constexpr Tile kT = kTiles[GPE_VID];

template <int R, int KIND, bool COMP>
hipError_t launch_one(const StageArgs &A, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel<GPE_REAL, R, kT.nw, kT.p, KIND, COMP>), dim3(A.nblocks), dim3(64 * kT.nw), 0, s,
                     A);
  return hipGetLastError();
}

template <int R, int KIND, bool COMP>
hipError_t launch_one_b(const BatchLaunch &B, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel_batched<GPE_REAL, R, kT.nw, kT.p, KIND, COMP>), dim3(B.grid), dim3(64 * kT.nw), 0,
                     s, B.insts, B.work, B.nwork, B.active, B.pn, B.pe);
  return hipGetLastError();
}

template <int R, bool COMP>
hipError_t launch_rb(int kind, const BatchLaunch &B, hipStream_t s) {
  switch (kind) {
    case K_EULER: return launch_one_b<R, K_EULER, COMP>(B, s);
    case K_RK2_1: return launch_one_b<R, K_RK2_1, false>(B, s);
    case K_RK2_2: return launch_one_b<R, K_RK2_2, COMP>(B, s);
    case K_RK4_1: return launch_one_b<R, K_RK4_1, false>(B, s);
    case K_RK4_2: return launch_one_b<R, K_RK4_2, false>(B, s);
    case K_RK4_3: return launch_one_b<R, K_RK4_3, false>(B, s);
    case K_RK4_4: return launch_one_b<R, K_RK4_4, COMP>(B, s);
    case K_CK_1: return launch_one_b<R, K_CK_1, COMP>(B, s);
    case K_CK_2: return launch_one_b<R, K_CK_2, COMP>(B, s);
    case K_CK_3: return launch_one_b<R, K_CK_3, COMP>(B, s);
    case K_CK_4: return launch_one_b<R, K_CK_4, COMP>(B, s);
    case K_CK_5: return launch_one_b<R, K_CK_5, COMP>(B, s);
    case K_CK_6: return launch_one_b<R, K_CK_6, COMP>(B, s);
    default: return hipErrorInvalidValue;
  }
}

template <int R, bool COMP>
hipError_t launch_r(int kind, const StageArgs &A, hipStream_t s) {
  switch (kind) {
    case K_RHS: return launch_one<R, K_RHS, false>(A, s);
    case K_EULER: return launch_one<R, K_EULER, COMP>(A, s);
    case K_RK2_1: return launch_one<R, K_RK2_1, false>(A, s);
    case K_RK2_2: return launch_one<R, K_RK2_2, COMP>(A, s);
    case K_RK4_1: return launch_one<R, K_RK4_1, false>(A, s);
    case K_RK4_2: return launch_one<R, K_RK4_2, false>(A, s);
    case K_RK4_3: return launch_one<R, K_RK4_3, false>(A, s);
    case K_RK4_4: return launch_one<R, K_RK4_4, COMP>(A, s);
    case K_CK_1: return launch_one<R, K_CK_1, COMP>(A, s);
    case K_CK_2: return launch_one<R, K_CK_2, COMP>(A, s);
    case K_CK_3: return launch_one<R, K_CK_3, COMP>(A, s);
    case K_CK_4: return launch_one<R, K_CK_4, COMP>(A, s);
    case K_CK_5: return launch_one<R, K_CK_5, COMP>(A, s);
    case K_CK_6: return launch_one<R, K_CK_6, COMP>(A, s);
    case K_ENERGY: return launch_one<R, K_ENERGY, false>(A, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

hipError_t GPE_FN2(GPE_PNAME, GPE_VID)(int kind, int radius, bool comp, const StageArgs &A, hipStream_t s) {
#ifdef GPE_F64
  (void)comp;
  if (radius == 1) return launch_r<1, false>(kind, A, s);
  if (radius == 2) return launch_r<2, false>(kind, A, s);
#else
  if (radius == 1) return comp ? launch_r<1, true>(kind, A, s) : launch_r<1, false>(kind, A, s);
  if (radius == 2) return comp ? launch_r<2, true>(kind, A, s) : launch_r<2, false>(kind, A, s);
#endif
  return hipErrorInvalidValue;
}

hipError_t GPE_BFN2(GPE_PNAME, GPE_VID)(int kind, int radius, bool comp, const BatchLaunch &B, hipStream_t s) {
#ifdef GPE_F64
  (void)comp;
  if (radius == 1) return launch_rb<1, false>(kind, B, s);
  if (radius == 2) return launch_rb<2, false>(kind, B, s);
#else
  if (radius == 1) return comp ? launch_rb<1, true>(kind, B, s) : launch_rb<1, false>(kind, B, s);
  if (radius == 2) return comp ? launch_rb<2, true>(kind, B, s) : launch_rb<2, false>(kind, B, s);
#endif
  return hipErrorInvalidValue;
}

}  // namespace gpe
