// host_common.cpp -- host helpers shared by the single-simulation and the
// batch entry points: the launch descriptor's constant part and field
// conversions between host arrays and device storage (error reporting is
// in errors.cpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "host_common.h"

namespace gpe {
// Division by a real constant in the reference's form (see stage_args.h).
DivC make_divc(double s) {
  DivC d{};
  int e = ilogb(s);
  d.e = e;
  d.cs = ldexp(s, -e);
  d.den = d.cs * d.cs + 0.0 * 0.0;
  return d;
}
}  // namespace gpe

namespace gpe {

void fill_base(const gpe_params *params, int prec, StageArgs &A) {
  const int R = params->order / 2;

  A.nx = params->nx;
  A.ny = params->ny;
  A.nz = params->nz;
  A.periodic = params->bcs == 1;
  A.eqn = params->eqn;
  A.use_diss = params->diss_amp != 0.0;
  A.use_urhs = params->urhs != 0.0;
  A.scale_real = params->real_time ? 1 : 0;
  A.xl = params->xl;
  A.yl = params->yl;
  A.zl = params->zl;
  A.dx = params->dx;
  A.dy = params->dy;
  A.dz = params->dz;
  A.g = params->g;
  A.mu = params->mu;
  A.ox2 = params->omx * params->omx;
  A.oy2 = params->omy * params->omy;
  A.oz2 = params->omz * params->omz;
  A.urhs = params->urhs;
  A.diss_amp = params->diss_amp;
  // a = x(nx-3) etc. (solve.f90:397-399)
  A.a_diss = params->xl + (double)(params->nx - 3) * params->dx;
  A.b_diss = params->yl + (double)(params->ny - 3) * params->dy;
  A.c_diss = params->zl + (double)(params->nz - 3) * params->dz;
  const double dx2 = params->dx * params->dx, dy2 = params->dy * params->dy, dz2 = params->dz * params->dz;
  const double sxx = R == 1 ? dx2 : 12.0 * dx2;
  const double syy = R == 1 ? dy2 : 12.0 * dy2;
  const double szz = R == 1 ? dz2 : 12.0 * dz2;
  const double sx = R == 1 ? 2.0 * params->dx : 12.0 * params->dx;
  if (prec == GPE_FP32) {
    // the reference built with pr = single forms these in single precision
    const float fdx = (float)params->dx, fdy = (float)params->dy, fdz = (float)params->dz;
    const float fx2 = fdx * fdx, fy2 = fdy * fdy, fz2 = fdz * fdz;
    A.inv_xx = 1.0 / (double)(R == 1 ? fx2 : 12.0f * fx2);
    A.inv_yy = 1.0 / (double)(R == 1 ? fy2 : 12.0f * fy2);
    A.inv_zz = 1.0 / (double)(R == 1 ? fz2 : 12.0f * fz2);
    A.inv_x = 1.0 / (double)(R == 1 ? 2.0f * fdx : 12.0f * fdx);
  } else {
    A.inv_xx = 1.0 / sxx;
    A.inv_yy = 1.0 / syy;
    A.inv_zz = 1.0 / szz;
    A.inv_x = 1.0 / sx;
  }
  A.d_xx = make_divc(sxx);
  A.d_yy = make_divc(syy);
  A.d_zz = make_divc(szz);
  A.d_x = make_divc(sx);
  A.d6 = make_divc(6.0);
  A.d3 = make_divc(3.0);
  A.w1 = 3.0 / 8.0;
  A.w2 = 7.0 / 6.0;
  A.w3 = 23.0 / 24.0;
  }

}  // namespace gpe

namespace gpe {

// Host array -> device storage.  src_pitch_planes: elements between
// z-planes of the host array; row_off: element offset of the interior's
// (0, 0, 0).  FP32 forms from complex128 split each value into hi = fl32(v)
// and lo = fl32(v - hi); complex64 into a compensated form sets lo = 0.
int field_upload(int prec, size_t plane, int nz, const void *src, int host_type, size_t src_pitch_planes,
                 size_t row_off, void *hi, void *lo) {
  const size_t n = plane * (size_t)nz;
  const bool contiguous = src_pitch_planes == plane && row_off == 0;
  const bool comp = prec == GPE_FP32C;
  if (host_type != GPE_HOST_C128 && host_type != GPE_HOST_C64) return fail(GPE_EINVAL, "unsupported host type");
  if (prec == GPE_FP64) {
    if (host_type != GPE_HOST_C128) return fail(GPE_EINVAL, "FP64 storage takes complex128 host data");
    if (contiguous) {
      GPE_HIPCHK(hipMemcpy(hi, src, n * 16, hipMemcpyHostToDevice));
    } else {
      GPE_HIPCHK(hipMemcpy2D(hi, plane * 16, (const char *)src + row_off * 16, src_pitch_planes * 16, plane * 16,
                             nz, hipMemcpyHostToDevice));
    }
    return GPE_OK;
  }
  if (host_type == GPE_HOST_C64) {
    if (contiguous) {
      GPE_HIPCHK(hipMemcpy(hi, src, n * 8, hipMemcpyHostToDevice));
    } else {
      GPE_HIPCHK(hipMemcpy2D(hi, plane * 8, (const char *)src + row_off * 8, src_pitch_planes * 8, plane * 8, nz,
                             hipMemcpyHostToDevice));
    }
    if (comp) GPE_HIPCHK(hipMemset(lo, 0, n * 8));
    return GPE_OK;
  }
  std::vector<float> h(2 * plane), l(2 * plane);
  for (int k = 0; k < nz; ++k) {
    const double *s = (const double *)src + 2 * (row_off + (size_t)k * src_pitch_planes);
    for (size_t q = 0; q < 2 * plane; ++q) {
      const float v = (float)s[q];
      h[q] = v;
      l[q] = (float)(s[q] - (double)v);
    }
    GPE_HIPCHK(hipMemcpy((char *)hi + (size_t)k * plane * 8, h.data(), plane * 8, hipMemcpyHostToDevice));
    if (comp) GPE_HIPCHK(hipMemcpy((char *)lo + (size_t)k * plane * 8, l.data(), plane * 8, hipMemcpyHostToDevice));
  }
  return GPE_OK;
}

// Device storage -> contiguous host array (complex128 gets hi + lo in
// double; complex64 gets fl32(hi + lo)).  The caller orders this after the
// kernels that wrote the field.
int field_download(int prec, size_t plane, int nz, const void *hi, const void *lo, void *dst, int host_type) {
  const size_t n = plane * (size_t)nz;
  const bool comp = prec == GPE_FP32C;
  if (host_type != GPE_HOST_C128 && host_type != GPE_HOST_C64) return fail(GPE_EINVAL, "unsupported host type");
  if (prec == GPE_FP64) {
    if (host_type != GPE_HOST_C128) return fail(GPE_EINVAL, "FP64 storage returns complex128 host data");
    GPE_HIPCHK(hipMemcpy(dst, hi, n * 16, hipMemcpyDeviceToHost));
    return GPE_OK;
  }
  if (host_type == GPE_HOST_C64 && !comp) {
    GPE_HIPCHK(hipMemcpy(dst, hi, n * 8, hipMemcpyDeviceToHost));
    return GPE_OK;
  }
  std::vector<float> h(2 * plane), l(2 * plane, 0.f);
  for (int k = 0; k < nz; ++k) {
    GPE_HIPCHK(hipMemcpy(h.data(), (const char *)hi + (size_t)k * plane * 8, plane * 8, hipMemcpyDeviceToHost));
    if (comp)
      GPE_HIPCHK(hipMemcpy(l.data(), (const char *)lo + (size_t)k * plane * 8, plane * 8, hipMemcpyDeviceToHost));
    if (host_type == GPE_HOST_C128) {
      double *d = (double *)dst + 2 * (size_t)k * plane;
      for (size_t q = 0; q < 2 * plane; ++q) d[q] = (double)h[q] + (double)l[q];
    } else {
      float *d = (float *)dst + 2 * (size_t)k * plane;
      for (size_t q = 0; q < 2 * plane; ++q) d[q] = (float)((double)h[q] + (double)l[q]);
    }
  }
  return GPE_OK;
}

}  // namespace gpe
