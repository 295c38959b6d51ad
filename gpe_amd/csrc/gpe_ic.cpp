// gpe_ic.cpp -- the random-phase initial state built on the device.
//
// The reference builds random_phase (ic.f90:195-274) on the host: modes
// inside a sphere of wavenumbers get a fixed amplitude and a random phase,
// then an inverse FFT (2DECOMP&FFT over FFTW, ic.f90:910-939) gives psi.
// This package's version of that state (the 512^3 random-phase field) is:
// every mode k = 2 pi m / (n d) (m as numpy's fftfreq orders it) with
// |k| <= kmax gets amplitude 1 and phase 2 pi u, u from SplitMix64(seed) in
// the order of the modes' flat index; inverse FFT; rescale so <|psi|^2> = 1.
// The mode list is tiny (about 3.5e4 modes at 512^3) and is formed on the
// host in exactly the order and arithmetic of the host generator
// (gpe_amd/ic.py random_phase); the device scatters it, runs hipFFT's
// inverse Z2Z, sums |psi|^2 with fixed-order block partials, and scales and
// converts into the storage form in one pass.  Compiled as HIP.
#include "gpe_ic.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace gpe {

namespace {

constexpr int kIcBlocks = 1024;

uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 2 pi * fftfreq(n, d)[i] as numpy forms it: (i or i - n) * (1 / (n d)),
// times 2 pi
double kvec(int i, int n, double d) {
  const int m = i < (n + 1) / 2 ? i : i - n;
  const double val = 1.0 / ((double)n * d);
  return (2.0 * M_PI) * ((double)m * val);
}

__global__ void ic_scatter(double2 *f, const long long *idx, const double2 *v, int nm) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nm) f[idx[q]] = v[q];
}

__global__ void __launch_bounds__(256) ic_sumsq(const double2 *f, size_t n, double *part) {
  double acc = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double2 v = f[q];
    acc += v.x * v.x + v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// psi = f * s in the storage form: prec 0 complex128 (in place when out ==
// f), 1 complex64, 2 complex64 hi + lo
__global__ void ic_store(const double2 *f, size_t n, double s, int prec, void *out, void *lo) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double2 v = f[q];
    const double re = v.x * s, im = v.y * s;
    if (prec == GPE_FP64) {
      static_cast<double2 *>(out)[q] = make_double2(re, im);
    } else {
      const float hr = (float)re, hi = (float)im;
      static_cast<float2 *>(out)[q] = make_float2(hr, hi);
      if (prec == GPE_FP32C)
        static_cast<float2 *>(lo)[q] = make_float2((float)(re - (double)hr), (float)(im - (double)hi));
    }
  }
}

// Inverse Z2Z plans by grid, created on first use (or ahead of it by
// gpe_random_phase_prepare, from any thread) and kept for the life of the
// process: creating a plan costs far more than running it, and a served
// request stream brings the same few grid sizes again and again.  The
// mutex also serialises set-stream + execute on a shared plan.
std::mutex g_plan_mu;
std::map<std::tuple<int, int, int>, hipfftHandle> g_plans;

const char *fft_err(hipfftResult r);

// A cached plan for the (nz, ny, nx) grid; call with g_plan_mu held.
int plan_for(int nx, int ny, int nz, hipfftHandle *out) {
  const auto key = std::make_tuple(nx, ny, nz);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    *out = it->second;
    return GPE_OK;
  }
  hipfftHandle plan = 0;
  // (nz, ny, nx) with x fastest: hipFFT's first size is the slowest
  const hipfftResult fr = hipfftPlan3d(&plan, nz, ny, nx, HIPFFT_Z2Z);
  if (fr != HIPFFT_SUCCESS) return fail(GPE_EHIP, "hipfftPlan3d: %s", fft_err(fr));
  g_plans[key] = plan;
  *out = plan;
  return GPE_OK;
}

const char *fft_err(hipfftResult r) {
  switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    default: return "hipFFT error";
  }
}

}  // namespace

int random_phase_modes(int nx, int ny, int nz, double dx, double dy, double dz, double kmax, uint64_t seed,
                       std::vector<long long> &idx, std::vector<double> &re, std::vector<double> &im) {
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || !(dy > 0) || !(dz > 0) || !(kmax >= 0))
    return fail(GPE_EINVAL, "random phase: bad grid or kmax");
  idx.clear();
  re.clear();
  im.clear();
  std::vector<int> mx, my, mz;
  std::vector<double> kx, ky, kz;
  for (int i = 0; i < nx; ++i)
    if (std::fabs(kvec(i, nx, dx)) <= kmax) mx.push_back(i), kx.push_back(kvec(i, nx, dx));
  for (int j = 0; j < ny; ++j)
    if (std::fabs(kvec(j, ny, dy)) <= kmax) my.push_back(j), ky.push_back(kvec(j, ny, dy));
  for (int k = 0; k < nz; ++k)
    if (std::fabs(kvec(k, nz, dz)) <= kmax) mz.push_back(k), kz.push_back(kvec(k, nz, dz));
  const double k2max = kmax * kmax;
  // ascending flat index: k, then j, then i (each list is ascending)
  for (size_t c = 0; c < mz.size(); ++c)
    for (size_t b = 0; b < my.size(); ++b)
      for (size_t a = 0; a < mx.size(); ++a) {
        const double K2 = (kz[c] * kz[c] + ky[b] * ky[b]) + kx[a] * kx[a];
        if (K2 <= k2max) idx.push_back(((long long)mz[c] * ny + my[b]) * nx + mx[a]);
      }
  re.resize(idx.size());
  im.resize(idx.size());
  for (size_t q = 0; q < idx.size(); ++q) {
    const double u = (double)(splitmix64(seed, q + 1) >> 11) * (1.0 / 9007199254740992.0);
    const double ph = (2.0 * M_PI) * u;
    re[q] = std::cos(ph);
    im[q] = std::sin(ph);
  }
  return GPE_OK;
}

int random_phase_device(int nx, int ny, int nz, double dx, double dy, double dz, double kmax, uint64_t seed,
                        int prec, void *out, void *lo, hipStream_t s) {
  std::vector<long long> idx;
  std::vector<double> re, im;
  int rc = random_phase_modes(nx, ny, nz, dx, dy, dz, kmax, seed, idx, re, im);
  if (rc) return rc;
  if (idx.empty()) return fail(GPE_EINVAL, "random phase: no mode with |k| <= kmax");
  const size_t n = (size_t)nx * ny * nz;
  const int nm = (int)idx.size();
  std::vector<double2> v(nm);
  for (int q = 0; q < nm; ++q) v[q] = make_double2(re[q], im[q]);
  // complex128 work field: the output itself for FP64, else a temporary
  double2 *f = nullptr;
  bool own = false;
  if (prec == GPE_FP64) {
    f = static_cast<double2 *>(out);
  } else {
    if (hipMalloc((void **)&f, n * sizeof(double2)) != hipSuccess)
      return fail(GPE_ENOMEM, "random phase: %zu-byte work field", n * sizeof(double2));
    own = true;
  }
  long long *d_idx = nullptr;
  double2 *d_v = nullptr;
  double *d_part = nullptr;
  std::vector<double> part(kIcBlocks);
  hipfftResult fr;
  double sum = 0.0;
#define IC_CHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      rc = fail(GPE_EHIP, "random phase: %s: %s", #call, hipGetErrorString(e_));       \
      goto done;                                                                       \
    }                                                                                  \
  } while (0)
  IC_CHK(hipMalloc((void **)&d_idx, sizeof(long long) * nm));
  IC_CHK(hipMalloc((void **)&d_v, sizeof(double2) * nm));
  IC_CHK(hipMalloc((void **)&d_part, sizeof(double) * kIcBlocks));
  IC_CHK(hipMemcpyAsync(d_idx, idx.data(), sizeof(long long) * nm, hipMemcpyHostToDevice, s));
  IC_CHK(hipMemcpyAsync(d_v, v.data(), sizeof(double2) * nm, hipMemcpyHostToDevice, s));
  IC_CHK(hipMemsetAsync(f, 0, n * sizeof(double2), s));
  hipLaunchKernelGGL(ic_scatter, dim3((nm + 255) / 256), dim3(256), 0, s, f, d_idx, d_v, nm);
  IC_CHK(hipGetLastError());
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    hipfftHandle plan = 0;
    if ((rc = plan_for(nx, ny, nz, &plan))) goto done;
    if ((fr = hipfftSetStream(plan, s)) != HIPFFT_SUCCESS ||
        (fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)f, (hipfftDoubleComplex *)f, HIPFFT_BACKWARD)) !=
            HIPFFT_SUCCESS) {
      rc = fail(GPE_EHIP, "random phase: inverse FFT: %s", fft_err(fr));
      goto done;
    }
  }
  hipLaunchKernelGGL(ic_sumsq, dim3(kIcBlocks), dim3(256), 0, s, f, n, d_part);
  IC_CHK(hipGetLastError());
  IC_CHK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * kIcBlocks, hipMemcpyDeviceToHost, s));
  IC_CHK(hipStreamSynchronize(s));
  for (double p : part) sum += p;
  if (!(sum > 0.0)) {
    rc = fail(GPE_EINVAL, "random phase: zero field");
    goto done;
  }
  // the inverse FFT is unnormalised; <|psi|^2> = 1 fixes the scale
  hipLaunchKernelGGL(ic_store, dim3(kIcBlocks), dim3(256), 0, s, f, n, 1.0 / std::sqrt(sum / (double)n), prec, out,
                     lo);
  IC_CHK(hipGetLastError());
  IC_CHK(hipStreamSynchronize(s));
done:
#undef IC_CHK
  // on an early exit, let queued copies and kernels finish before their
  // buffers (host vectors included) go away
  if (rc) (void)hipStreamSynchronize(s);
  if (d_idx) (void)hipFree(d_idx);
  if (d_v) (void)hipFree(d_v);
  if (d_part) (void)hipFree(d_part);
  if (own && f) (void)hipFree(f);
  return rc;
}

int fft_z2z(int nx, int ny, int nz, void *f, int direction, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  hipfftHandle plan = 0;
  if (int rc = plan_for(nx, ny, nz, &plan)) return rc;
  hipfftResult fr = hipfftSetStream(plan, s);
  if (fr == HIPFFT_SUCCESS)
    fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex *)f, (hipfftDoubleComplex *)f,
                       direction < 0 ? HIPFFT_FORWARD : HIPFFT_BACKWARD);
  if (fr != HIPFFT_SUCCESS) return fail(GPE_EHIP, "3-D FFT: %s", fft_err(fr));
  return GPE_OK;
}

}  // namespace gpe

extern "C" int gpe_random_phase_modes(int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                                      double kmax, uint64_t seed, int64_t *flat_index, double *phasor, int64_t cap,
                                      int64_t *count) {
  if (!count || cap < 0 || (cap > 0 && (!flat_index || !phasor))) return gpe::fail(GPE_EINVAL, "bad output buffer");
  std::vector<long long> idx;
  std::vector<double> re, im;
  const int rc = gpe::random_phase_modes(nx, ny, nz, dx, dy, dz, kmax, seed, idx, re, im);
  if (rc) return rc;
  *count = (int64_t)idx.size();
  for (int64_t q = 0; q < std::min<int64_t>(cap, (int64_t)idx.size()); ++q) {
    flat_index[q] = idx[q];
    phasor[2 * q] = re[q];
    phasor[2 * q + 1] = im[q];
  }
  return GPE_OK;
}

extern "C" int gpe_random_phase_prepare(int32_t nx, int32_t ny, int32_t nz) {
  if (nx < 1 || ny < 1 || nz < 1) return gpe::fail(GPE_EINVAL, "random phase: bad grid");
  std::lock_guard<std::mutex> lk(gpe::g_plan_mu);
  hipfftHandle plan = 0;
  return gpe::plan_for(nx, ny, nz, &plan);
}
