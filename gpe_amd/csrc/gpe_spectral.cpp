// gpe_spectral.cpp -- FFT-side diagnostics on the device: the unitary
// forward transform of psi (fft, ic.f90:910-939), the condensate and
// temperature of condensed_particles (io.f90:304-395), the occupation
// spectrum and integral distribution of spectrum (io.f90:623-716) and the
// low-pass filtered field of filtered_surface (io.f90:524-619).
//
// Spectral layout: a(i, j, k) at i + nx (j + ny k), wavenumber indices
// 0 .. n-1 along each axis in natural order, x fastest, as the single-rank
// z-pencil output the reference's loops walk (xs1:xe1 = 0:nx-1, ...).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gpe_ic.h"
#include "refmath.h"
#include "gpe_spectral.h"

namespace gpe {
namespace {

constexpr int kSpecThreads = 256;
constexpr int kNShells = 7;  // nshells, io.f90:630

// k^2 of the reference's index wrap along one axis (io.f90:642-659):
// i <= (n-1)/2 + 1 counts as i, above it as n - i.
__host__ __device__ inline long long wrap2(int i, int n) {
  const int n1 = n - 1;
  const long long v = i <= n1 / 2 + 1 ? i : n1 - i + 1;
  return v * v;
}

__global__ void to_c128(int prec, const void *psi, const void *lo, double2 *out, size_t n) {
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    double2 v;
    if (prec == 0) {
      v = static_cast<const double2 *>(psi)[s];
    } else {
      const float2 h = static_cast<const float2 *>(psi)[s];
      v = make_double2((double)h.x, (double)h.y);
      if (lo) {
        const float2 l = static_cast<const float2 *>(lo)[s];
        v.x += (double)l.x;
        v.y += (double)l.y;
      }
    }
    out[s] = v;
  }
}

// out_var = out_var / sqrt(real(nx*ny*nz)) (ic.f90:933)
__global__ void div_real(double2 *a, size_t n, double d) {
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    double2 v = a[s];
    v.x = v.x / d;
    v.y = v.y / d;
    a[s] = v;
  }
}

// abs(a)**2 with abs the complex modulus (the host libm's hypot, restated)
__device__ inline double mod2(double2 v) {
  const double m = gpe::refm::hypot_ref(v.x, v.y);
  return m * m;
}

// One block per z plane: shell sums and counts (eta, nharm) and the
// per-radius sums whose prefix sums are F (io.f90:661-681).  Partials per
// plane: [eta 0..6 | bins 0..nb-1] doubles and [nharm 0..6] counts.
__global__ void __launch_bounds__(kSpecThreads)
    spectrum_planes(const double2 *a, int nx, int ny, int nz, int nb, double *part, long long *cnt) {
  extern __shared__ double sh[];
  double *eta = sh, *bins = sh + kNShells;
  __shared__ int nh[kNShells];
  const int k = blockIdx.x;
  for (int q = threadIdx.x; q < kNShells + nb; q += blockDim.x) sh[q] = 0.0;
  if (threadIdx.x < kNShells) nh[threadIdx.x] = 0;
  __syncthreads();
  const long long kk2 = wrap2(k, nz);
  const double pi = 3.14159265358979323846;
  for (int q = threadIdx.x; q < nx * ny; q += blockDim.x) {
    const int j = q / nx, i = q - j * nx;
    const long long k2 = wrap2(i, nx) + wrap2(j, ny) + kk2;
    if (k2 == 0) continue;
    const double v = mod2(a[(size_t)k * nx * ny + q]);
    // F(kk) takes every mode with sqrt(k2) <= kk: bin ceil(sqrt(k2))
    long long b = (long long)sqrt((double)k2);
    while (b * b < k2) ++b;
    while (b > 1 && (b - 1) * (b - 1) >= k2) --b;
    if (b <= nb) atomicAdd(&bins[b - 1], v);
    // shell m: m-1 <= |log2 (|k| / (2 pi))| < m
    const double log2k = log(0.5 * sqrt((double)k2) / pi) / log(2.0);
    const double al = fabs(log2k);
    for (int m = 1; m <= kNShells; ++m) {
      if (al < (double)m && al >= (double)(m - 1)) {
        atomicAdd(&eta[m - 1], v);
        atomicAdd(&nh[m - 1], 1);
        break;
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kNShells + nb; q += blockDim.x) part[(size_t)k * (kNShells + nb) + q] = sh[q];
  if (threadIdx.x < kNShells) cnt[(size_t)k * kNShells + threadIdx.x] = nh[threadIdx.x];
}

// One block per z plane: sum of (k2 + rho0) / (k4 + 2 rho0 k2) over the
// grid of k = -kc + i dk (io.f90:349-366); cx, cy, cz hold each axis's
// (-2 cos(2 k d) + 32 cos(k d) - 30) / d^2, formed on the host.
__global__ void __launch_bounds__(kSpecThreads)
    condensed_planes(const double *cx, const double *cy, const double *cz, int nx, int ny, double rho0,
                     double *part) {
  __shared__ double red[kSpecThreads];
  const int k = blockIdx.x;
  const double ck = cz[k];
  double acc = 0.0;
  for (int q = threadIdx.x; q < nx * ny; q += blockDim.x) {
    const int j = q / nx, i = q - j * nx;
    const double k2 = -(1.0 / 12.0) * ((cx[i] + cy[j]) + ck);
    if (k2 == 0.0) continue;
    const double k4 = k2 * k2;
    acc += (k2 + rho0) / (k4 + 2.0 * rho0 * k2);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kSpecThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[k] = red[0];
}

// a = a * max(1 - m k2 / kc^2, 0) (io.f90:570)
__global__ void lowpass(double2 *a, int nx, int ny, int nz, double m, double kc) {
  const size_t n = (size_t)nx * ny * nz;
  const double kc2 = kc * kc;
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(s % nx);
    const size_t r = s / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const long long k2 = wrap2(i, nx) + wrap2(j, ny) + wrap2(k, nz);
    const double f = fmax(1.0 - m * (double)k2 / kc2, 0.0);
    double2 v = a[s];
    v.x = v.x * f;
    v.y = v.y * f;
    a[s] = v;
  }
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + kSpecThreads - 1) / kSpecThreads, 65536); }

// on an error, let queued copies into host vectors finish before returning
#define SP_CHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (void)hipStreamSynchronize(s);                                                   \
      return fail(GPE_EHIP, "%s: %s", #call, hipGetErrorString(e_));                   \
    }                                                                                  \
  } while (0)

}  // namespace

int spectral_forward(int nx, int ny, int nz, int prec, const void *psi, const void *lo, double2 *a, hipStream_t s) {
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(to_c128, dim3(grid_for(n)), dim3(kSpecThreads), 0, s, prec, psi, lo, a, n);
  SP_CHK(hipGetLastError());
  if (int rc = fft_z2z(nx, ny, nz, a, -1, s)) return rc;
  hipLaunchKernelGGL(div_real, dim3(grid_for(n)), dim3(kSpecThreads), 0, s, a, n, std::sqrt((double)n));
  SP_CHK(hipGetLastError());
  return GPE_OK;
}

int spectral_spectrum(int nx, int ny, int nz, const double2 *a, int nb, double *part, long long *cnt,
                      long long *nharm, double *eta, double *bins, hipStream_t s) {
  const size_t shm = sizeof(double) * (kNShells + (size_t)nb);
  if (shm > 60 * 1024) return fail(GPE_EINVAL, "spectrum: %d radius bins do not fit one workgroup", nb);
  hipLaunchKernelGGL(spectrum_planes, dim3(nz), dim3(kSpecThreads), shm, s, a, nx, ny, nz, nb, part, cnt);
  SP_CHK(hipGetLastError());
  std::vector<double> hp((size_t)nz * (kNShells + nb));
  std::vector<long long> hc((size_t)nz * kNShells);
  SP_CHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
  SP_CHK(hipMemcpyAsync(hc.data(), cnt, sizeof(long long) * hc.size(), hipMemcpyDeviceToHost, s));
  SP_CHK(hipStreamSynchronize(s));
  for (int m = 0; m < kNShells; ++m) {
    nharm[m] = 0;
    eta[m] = 0.0;
  }
  for (int b = 0; b < nb; ++b) bins[b] = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double *p = hp.data() + (size_t)k * (kNShells + nb);
    for (int m = 0; m < kNShells; ++m) {
      eta[m] += p[m];
      nharm[m] += hc[(size_t)k * kNShells + m];
    }
    for (int b = 0; b < nb; ++b) bins[b] += p[kNShells + b];
  }
  return GPE_OK;
}

int spectral_condensed_sum(int nx, int ny, int nz, double dx, double dy, double dz, double rho0, double *dev,
                           double *tot, hipStream_t s) {
  // kc = pi, dk = 2 kc / nx1 on every axis (io.f90:321-322, 352-357)
  const double kc = 3.14159265358979323846, dk = 2.0 * kc / (double)(nx - 1);
  std::vector<double> h((size_t)nx + ny + nz);
  auto axis = [&](double *c, int n, double d) {
    const double d2 = d * d;
    for (int i = 0; i < n; ++i) {
      const double kv = -kc + (double)i * dk;
      c[i] = ((-2.0 * std::cos(2.0 * kv * d) + 32.0 * std::cos(kv * d)) - 30.0) / d2;
    }
  };
  axis(h.data(), nx, dx);
  axis(h.data() + nx, ny, dy);
  axis(h.data() + nx + ny, nz, dz);
  double *dc = dev, *part = dev + h.size();
  SP_CHK(hipMemcpyAsync(dc, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(condensed_planes, dim3(nz), dim3(kSpecThreads), 0, s, dc, dc + nx, dc + nx + ny, nx, ny, rho0,
                     part);
  SP_CHK(hipGetLastError());
  std::vector<double> hp(nz);
  SP_CHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * nz, hipMemcpyDeviceToHost, s));
  SP_CHK(hipStreamSynchronize(s));
  double acc = 0.0;
  for (double v : hp) acc += v;
  *tot = acc;
  return GPE_OK;
}

int spectral_filter(int nx, int ny, int nz, double2 *a, double m, double kc, hipStream_t s) {
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(lowpass, dim3(grid_for(n)), dim3(kSpecThreads), 0, s, a, nx, ny, nz, m, kc);
  SP_CHK(hipGetLastError());
  if (int rc = fft_z2z(nx, ny, nz, a, +1, s)) return rc;
  hipLaunchKernelGGL(div_real, dim3(grid_for(n)), dim3(kSpecThreads), 0, s, a, n, std::sqrt((double)n));
  SP_CHK(hipGetLastError());
  return GPE_OK;
}

size_t spectral_scratch_doubles(int nx, int ny, int nz) {
  const int nb = nx / 2;
  // spectrum partials and counts (counts as doubles' worth of space), or
  // the condensed sum's axis tables and plane partials
  const size_t spec = (size_t)nz * (kNShells + nb) + (size_t)nz * kNShells;
  const size_t cond = (size_t)nx + ny + nz + nz;
  return std::max(spec, cond);
}

}  // namespace gpe
