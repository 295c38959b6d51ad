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

}  // namespace gpe
