// gpe_fp32.hip -- FP32 (complex64) instantiations: plain FP32 and the
// compensated form (FP32 hi/lo pair for psi, FP64 final update, FP64
// reductions).
#include "launch.h"
#include "pointwise.hpp"

namespace gpe {

template <int R, int KIND, bool COMP>
static hipError_t launch_one(const StageArgs &A, hipStream_t s) {
  constexpr int NW = 4, P = 4;
  hipLaunchKernelGGL((stage_kernel<float, R, NW, P, KIND, COMP>), dim3(A.nblocks), dim3(64 * NW), 0, s, A);
  return hipGetLastError();
}

template <int R, bool COMP>
static hipError_t launch_r(int kind, const StageArgs &A, hipStream_t s) {
  switch (kind) {
    case K_RHS: return launch_one<R, K_RHS, false>(A, s);
    case K_EULER: return launch_one<R, K_EULER, COMP>(A, s);
    case K_RK2_1: return launch_one<R, K_RK2_1, false>(A, s);
    case K_RK2_2: return launch_one<R, K_RK2_2, COMP>(A, s);
    case K_RK4_1: return launch_one<R, K_RK4_1, false>(A, s);
    case K_RK4_2: return launch_one<R, K_RK4_2, false>(A, s);
    case K_RK4_3: return launch_one<R, K_RK4_3, false>(A, s);
    case K_RK4_4: return launch_one<R, K_RK4_4, COMP>(A, s);
    case K_CK_1: return launch_one<R, K_CK_1, false>(A, s);
    case K_CK_2: return launch_one<R, K_CK_2, false>(A, s);
    case K_CK_3: return launch_one<R, K_CK_3, false>(A, s);
    case K_CK_4: return launch_one<R, K_CK_4, false>(A, s);
    case K_CK_5: return launch_one<R, K_CK_5, false>(A, s);
    case K_CK_6: return launch_one<R, K_CK_6, COMP>(A, s);
    case K_ENERGY: return launch_one<R, K_ENERGY, false>(A, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_f32(int kind, int radius, bool comp, const StageArgs &A, hipStream_t s) {
  if (radius == 1) return comp ? launch_r<1, true>(kind, A, s) : launch_r<1, false>(kind, A, s);
  if (radius == 2) return comp ? launch_r<2, true>(kind, A, s) : launch_r<2, false>(kind, A, s);
  return hipErrorInvalidValue;
}

hipError_t launch_norm_f32(const void *psi, const void *lo, int nx, int ny, int nz, int weighted, double w1,
                           double w2, double w3, double *part, hipStream_t s) {
  if (lo)
    hipLaunchKernelGGL((norm_kernel<float, true>), dim3(PW_BLOCKS), dim3(256), 0, s, psi, lo, nx, ny, nz, weighted,
                       w1, w2, w3, part);
  else
    hipLaunchKernelGGL((norm_kernel<float, false>), dim3(PW_BLOCKS), dim3(256), 0, s, psi, nullptr, nx, ny, nz,
                       weighted, w1, w2, w3, part);
  return hipGetLastError();
}

hipError_t launch_renorm_f32(void *psi, void *lo, size_t n, double a, double b, hipStream_t s) {
  if (lo)
    hipLaunchKernelGGL(renorm_kernel_f32<true>, dim3(PW_BLOCKS), dim3(256), 0, s, psi, lo, n, a, b);
  else
    hipLaunchKernelGGL(renorm_kernel_f32<false>, dim3(PW_BLOCKS), dim3(256), 0, s, psi, nullptr, n, a, b);
  return hipGetLastError();
}

}  // namespace gpe
