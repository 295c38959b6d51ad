// gpe_fp64.hip -- FP64 (complex128) instantiations: the parity path.
// Built with -ffp-contract=off: every operation is one IEEE rounding, in the
// reference's evaluation order.
#include "launch.h"
#include "pointwise.hpp"

namespace gpe {

Tile tile_shape(int prec_bytes, int radius, int kind) {
  // FP64 stage inputs that carry many k-fields keep fewer rows per wave to
  // stay within the register budget.
  if (prec_bytes == 16) {
    if (kind == K_CK_6 || kind == K_CK_5) return Tile{8, 2};
    return Tile{4, 4};
  }
  (void)radius;
  return Tile{4, 4};
}

template <int R, int KIND>
static hipError_t launch_one(const StageArgs &A, hipStream_t s) {
  constexpr int NW = (KIND == K_CK_6 || KIND == K_CK_5) ? 8 : 4;
  constexpr int P = (KIND == K_CK_6 || KIND == K_CK_5) ? 2 : 4;
  hipLaunchKernelGGL((stage_kernel<double, R, NW, P, KIND, false>), dim3(A.nblocks), dim3(64 * NW), 0, s, A);
  return hipGetLastError();
}

template <int R>
static hipError_t launch_r(int kind, const StageArgs &A, hipStream_t s) {
  switch (kind) {
    case K_RHS: return launch_one<R, K_RHS>(A, s);
    case K_EULER: return launch_one<R, K_EULER>(A, s);
    case K_RK2_1: return launch_one<R, K_RK2_1>(A, s);
    case K_RK2_2: return launch_one<R, K_RK2_2>(A, s);
    case K_RK4_1: return launch_one<R, K_RK4_1>(A, s);
    case K_RK4_2: return launch_one<R, K_RK4_2>(A, s);
    case K_RK4_3: return launch_one<R, K_RK4_3>(A, s);
    case K_RK4_4: return launch_one<R, K_RK4_4>(A, s);
    case K_CK_1: return launch_one<R, K_CK_1>(A, s);
    case K_CK_2: return launch_one<R, K_CK_2>(A, s);
    case K_CK_3: return launch_one<R, K_CK_3>(A, s);
    case K_CK_4: return launch_one<R, K_CK_4>(A, s);
    case K_CK_5: return launch_one<R, K_CK_5>(A, s);
    case K_CK_6: return launch_one<R, K_CK_6>(A, s);
    case K_ENERGY: return launch_one<R, K_ENERGY>(A, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_f64(int kind, int radius, const StageArgs &A, hipStream_t s) {
  if (radius == 1) return launch_r<1>(kind, A, s);
  if (radius == 2) return launch_r<2>(kind, A, s);
  return hipErrorInvalidValue;
}

hipError_t launch_reduce(const double *pn, const double *pe, const double *pen, int n, double scale_norm,
                         double eps, double scale_energy, double *out, double *hist, hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(256), 0, s, pn, pe, pen, n, scale_norm, eps, scale_energy,
                     out, hist);
  return hipGetLastError();
}

hipError_t launch_norm_f64(const void *psi, int nx, int ny, int nz, int weighted, double w1, double w2, double w3,
                           double *part, hipStream_t s) {
  hipLaunchKernelGGL((norm_kernel<double, false>), dim3(PW_BLOCKS), dim3(256), 0, s, psi, nullptr, nx, ny, nz,
                     weighted, w1, w2, w3, part);
  return hipGetLastError();
}

hipError_t launch_renorm_f64(void *psi, size_t n, double a, DivC b, hipStream_t s) {
  hipLaunchKernelGGL(renorm_kernel_f64, dim3(PW_BLOCKS), dim3(256), 0, s, psi, n, a, b);
  return hipGetLastError();
}

}  // namespace gpe
