// gpe_misc.hip -- reductions and pointwise kernels, and the dispatch over
// the stage-kernel variants.  Built with -ffp-contract=off.
#include "launch.h"
#include "pointwise.hpp"

namespace gpe {

hipError_t launch_stage(int prec_bytes, int vid, int kind, int radius, bool comp, const StageArgs &A,
                        hipStream_t s) {
  if (prec_bytes == 16) {
    switch (vid) {
      case 0: return launch_stage_f64_v0(kind, radius, comp, A, s);
      case 1: return launch_stage_f64_v1(kind, radius, comp, A, s);
      case 2: return launch_stage_f64_v2(kind, radius, comp, A, s);
      case 3: return launch_stage_f64_v3(kind, radius, comp, A, s);
      case 4: return launch_stage_f64_v4(kind, radius, comp, A, s);
    }
  } else {
    switch (vid) {
      case 0: return launch_stage_f32_v0(kind, radius, comp, A, s);
      case 1: return launch_stage_f32_v1(kind, radius, comp, A, s);
      case 2: return launch_stage_f32_v2(kind, radius, comp, A, s);
      case 3: return launch_stage_f32_v3(kind, radius, comp, A, s);
      case 4: return launch_stage_f32_v4(kind, radius, comp, A, s);
    }
  }
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

hipError_t launch_renorm_f64(void *psi, size_t n, double a, DivC b, hipStream_t s) {
  hipLaunchKernelGGL(renorm_kernel_f64, dim3(PW_BLOCKS), dim3(256), 0, s, psi, n, a, b);
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
