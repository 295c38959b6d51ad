// pointwise.hpp -- whole-field reductions and updates with no stencil.
#pragma once
#include "stencil.hpp"

namespace gpe {

// get_norm (weighted = 1: extended closed rule per axis, variables.f90:554-576)
// or mass (weighted = 0: plain sum, variables.f90:608-628) of |psi|^2.  Block b
// sums a contiguous slice of sites into part[b] (fixed order, so the result
// does not depend on scheduling).
template <typename T, bool COMP>
__global__ void __launch_bounds__(256) norm_kernel(const void *psi, const void *lo, int nx, int ny, int nz,
                                                   int weighted, double w1, double w2, double w3, double *part) {
  __shared__ double red[4];
  const size_t n = (size_t)nx * ny * nz;
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = (size_t)blockIdx.x * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  double acc = 0.0;
  for (size_t s = beg + threadIdx.x; s < end; s += blockDim.x) {
    double m2;
    if constexpr (std::is_same<T, double>::value) {
      double m = cabs_(ldc<double>(psi, s));
      m2 = m * m;  // abs(in_var)**2
    } else {
      cx<double> v = todbl(ldc<T>(psi, s));
      if constexpr (COMP) v = v + todbl(ldc<T>(lo, s));
      m2 = v.re * v.re + v.im * v.im;
    }
    if (weighted) {
      const int i = (int)(s % nx);
      const size_t r = s / nx;
      const int j = (int)(r % ny);
      const int k = (int)(r / ny);
      m2 *= nweight(i, nx, w1, w2, w3) * nweight(j, ny, w1, w2, w3) * nweight(k, nz, w1, w2, w3);
    }
    acc += m2;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Sums of |psi|^2, |psi|^4 and V |psi|^2 (V the trap of ic.f90:713-731 at
// x = xl + i dx, ...), in FP64 whatever the storage form: the terms of the
// eqn-4 Hamiltonian besides the kinetic one.  Fixed-order block partials,
// as norm_kernel.
template <typename T, bool COMP>
__global__ void __launch_bounds__(256)
    moments_kernel(const void *psi, const void *lo, int nx, int ny, int nz, double xl, double yl, double zl,
                   double dx, double dy, double dz, double ox2, double oy2, double oz2, double *p2, double *p4,
                   double *pv) {
  __shared__ double red[3][4];
  const size_t n = (size_t)nx * ny * nz;
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = (size_t)blockIdx.x * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  double a2 = 0.0, a4 = 0.0, av = 0.0;
  for (size_t s = beg + threadIdx.x; s < end; s += blockDim.x) {
    cx<double> v = todbl(ldc<T>(psi, s));
    if constexpr (COMP) v = v + todbl(ldc<T>(lo, s));
    const double m2 = v.re * v.re + v.im * v.im;
    const int i = (int)(s % nx);
    const size_t r = s / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const double x = xl + (double)i * dx, y = yl + (double)j * dy, z = zl + (double)k * dz;
    const double V = 0.5 * ((ox2 * (x * x) + oy2 * (y * y)) + oz2 * (z * z));
    a2 += m2;
    a4 += m2 * m2;
    av += V * m2;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a2 += __shfl_xor(a2, o);
    a4 += __shfl_xor(a4, o);
    av += __shfl_xor(av, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a2;
    red[1][threadIdx.x >> 6] = a4;
    red[2][threadIdx.x >> 6] = av;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    p2[blockIdx.x] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    p4[blockIdx.x] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    pv[blockIdx.x] = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
  }
}

// renormalise (variables.f90:715-726): psi = (a*psi)/b, a = sqrt(nn),
// b = sqrt(norm), with the reference's complex / real division.
static __global__ void __launch_bounds__(256) renorm_kernel_f64(void *psi, size_t n, double a, DivC b) {
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    cx<double> v = rmul(a, ldc<double>(psi, s));
    stc<double>(psi, s, rdiv(v, b, 0.0));
  }
}

template <bool COMP>
__global__ void __launch_bounds__(256) renorm_kernel_f32(void *psi, void *lo, size_t n, double a, double b) {
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    cx<double> v = todbl(ldc<float>(psi, s));
    if constexpr (COMP) v = v + todbl(ldc<float>(lo, s));
    v = rmul(a, v);
    v = mk(v.re / b, v.im / b);
    cx<float> hi = mk((float)v.re, (float)v.im);
    stc<float>(psi, s, hi);
    if constexpr (COMP) stc<float>(lo, s, mk((float)(v.re - (double)hi.re), (float)(v.im - (double)hi.im)));
  }
}

// imprint_vortex_line (variables.f90:1166-1180): psi = abs(psi) *
// vortex_line(vl), vortex_line = exp(i theta) (times amp(r) unless the
// line imprints its phase only, ic.f90:332-353), theta and r of get_theta /
// get_r (ic.f90:735-889) for a line along `dir` displaced by
// (a1 sin, a2 cos)(2 pi (c - c0)/(ll + safe)).  T is the arithmetic type:
// double for FP64 and the compensated form (on hi + lo), float for FP32.
template <typename T, int PREC>
__global__ void __launch_bounds__(256) imprint_kernel(void *psi, void *lo, LineArgs L) {
  const size_t n = (size_t)L.nx * L.ny * L.nz;
  const T twopi = (T)2.0 * (T)3.14159265358979323846;  // 2.0_pr*pi (parameters.f90:124)
  const T x0 = (T)L.x0, y0 = (T)L.y0, z0 = (T)L.z0, a1 = (T)L.a1, a2 = (T)L.a2, scal = (T)L.scal;
  const T w1 = (T)L.ll + (T)L.safe1, w2 = (T)L.ll + (T)L.safe2, sgn = (T)L.sgn;
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(s % L.nx);
    const size_t rr = s / L.nx;
    const int j = (int)(rr % L.ny);
    const int k = (int)(rr / L.ny);
    const T x = (T)L.xl + (T)i * (T)L.dx;
    const T y = (T)L.yl + (T)j * (T)L.dy;
    const T z = (T)L.zl + (T)k * (T)L.dz;
    T A, B;
    if (L.dir == 0) {  // 'x': atan2(z.., y..)
      const T t = x - x0;
      A = y - y0 - a1 * sin(twopi * t / w1);
      B = z - z0 - a2 * cos(twopi * t / w2);
    } else if (L.dir == 1) {  // 'y': atan2(x.., z..)
      const T t = y - y0;
      A = z - z0 - a1 * sin(twopi * t / w1);
      B = x - x0 - a2 * cos(twopi * t / w2);
    } else {  // 'z': atan2(y.., x..)
      const T t = z - z0;
      A = x - x0 - a1 * sin(twopi * t / w1);
      B = y - y0 - a2 * cos(twopi * t / w2);
    }
    const T theta = sgn * atan2(scal * B, scal * A);
    T vr = cos(theta), vi = sin(theta);
    if (!L.phase_only) {
      const T r = sqrt(scal * (A * A) + scal * (B * B));
      const T am = (T)1.0 - exp((T)-0.7 * pow(r, (T)1.15));  // amp (ic.f90:314-328)
      vr = am * vr;
      vi = am * vi;
    }
    if constexpr (PREC == 0) {
      const double m = cabs_(ldc<double>(psi, s));
      stc<double>(psi, s, mk(m * vr, m * vi));
    } else if constexpr (PREC == 1) {
      const float m = cabs_(ldc<float>(psi, s));
      stc<float>(psi, s, mk(m * vr, m * vi));
    } else {
      const cx<double> v = todbl(ldc<float>(psi, s)) + todbl(ldc<float>(lo, s));
      const double m = refm::hypot_ref(v.re, v.im);
      const cx<double> w = mk(m * vr, m * vi);
      const cx<float> hi = mk((float)w.re, (float)w.im);
      stc<float>(psi, s, hi);
      stc<float>(lo, s, mk((float)(w.re - (double)hi.re), (float)(w.im - (double)hi.im)));
    }
  }
}

}  // namespace gpe
