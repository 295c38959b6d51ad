// gpe_noalloc.cpp -- energy() of a caller's device field without any device
// allocation (gpe_energy_noalloc in include/gpe_hip.h).
//
// The stage kernel of energy() (kind K_ENERGY: the reference's
// sum Re(-lap psi psi*) + 1/2 |psi|^4 times dV, variables.f90:580-604) writes
// one partial sum per workgroup, which reduce_partials adds in a fixed
// order.  gpe_sim_energy keeps those partials in its context's buffer; this
// entry point keeps them in a file-scope device array instead, so a caller
// that cannot allocate (code under HIP graph capture, a serving loop, a
// harness that forbids hipMalloc) can call it on any field it holds.
//
// The array lives once per GPU for the life of the process.  Its device
// address is looked up with hipGetSymbolAddress once per device, with that
// device current, and kept; host code uses only that address (the array's
// name in host code is not a device address).  Calls that overlap on one GPU
// share the array and race; calls on one stream are ordered.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gpe_hip.h"
#include "host_common.h"
#include "launch.h"

using namespace gpe;

namespace {

// Partial sums (one per workgroup) and the reduction's three outputs.
constexpr int kMaxPartials = 65536;
constexpr int kOut = kMaxPartials;
constexpr int kScratchDoubles = kMaxPartials + 4;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_ptr(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(GPE_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(GPE_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    double *p = nullptr;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(GPE_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return GPE_OK;
}

}  // namespace

extern "C" int gpe_energy_noalloc(const gpe_params *params, const void *psi, double *energy, void *stream) {
  if (!params || !psi || !energy) return fail(GPE_EINVAL, "NULL argument");
  if (params->nx < 5 || params->ny < 5 || params->nz < 5) return fail(GPE_EINVAL, "grid must be at least 5 points per axis");
  if (params->bcs != 1 && params->bcs != 2) return fail(GPE_EINVAL, "bcs must be 1 (periodic) or 2 (reflective)");
  if (params->order != 2 && params->order != 4) return fail(GPE_EINVAL, "order must be 2 or 4");
  if (params->precision < GPE_FP64 || params->precision > GPE_FP32C) return fail(GPE_EINVAL, "unknown precision");
  if (!(params->dx > 0 && params->dy > 0 && params->dz > 0)) return fail(GPE_EINVAL, "grid spacing must be positive");
  const int prec = params->precision;
  StageArgs A{};
  fill_base(params, prec, A);
  // the launch configuration of a new context (gpe_ctx_create), so the
  // partial sums and their order equal gpe_sim_energy's
  const int vid = kDefaultVid;
  const int nb = stage_grid(params->nx, params->ny, params->nz, vid, 0, A);
  if (nb > kMaxPartials)
    return fail(GPE_EINVAL, "grid %dx%dx%d needs %d workgroups; the no-allocation scratch holds %d partial sums",
                params->nx, params->ny, params->nz, nb, kMaxPartials);
  double *sp = nullptr;
  if (int rc = scratch_ptr(&sp)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  A.psi = psi;
  A.part_norm = nullptr;
  A.part_emax = nullptr;
  A.part_energy = sp;
  hipError_t e = launch_stage(prec == GPE_FP64 ? 16 : 8, vid, K_ENERGY, params->order / 2, prec == GPE_FP32C, A, s);
  if (e != hipSuccess) return fail(GPE_EHIP, "energy launch: %s", hipGetErrorString(e));
  const double dv = params->dx * params->dy * params->dz;
  e = launch_reduce(nullptr, nullptr, sp, nb, dv, params->eps, dv, sp + kOut, nullptr, s);
  if (e != hipSuccess) return fail(GPE_EHIP, "reduce launch: %s", hipGetErrorString(e));
  e = hipMemcpyAsync(energy, sp + kOut + 2, sizeof(double), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return fail(GPE_EHIP, "result copy: %s", hipGetErrorString(e));
  return GPE_OK;
}
