// host_common.h -- host helpers shared by the single-simulation and batch
// entry points.
#pragma once
#include "../../include/gpe_hip.h"
#include "stage_args.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace gpe {
// Record the message of a failed call (read back by gpe_last_error) and
// return code.
int fail(int code, const char *fmt, ...);
#define GPE_HIPCHK(call)                                                                             \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return ::gpe::fail(GPE_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

DivC make_divc(double s);
// Geometry, coefficients and constant divisors of one simulation (no field
// pointers, no dt).
void fill_base(const gpe_params *params, int prec, StageArgs &A);
// Field conversions between host arrays and device storage (hi, and lo for
// the compensated form).
int field_upload(int prec, size_t plane, int nz, const void *src, int host_type, size_t src_pitch_planes,
                 size_t row_off, void *hi, void *lo);
int field_download(int prec, size_t plane, int nz, const void *hi, const void *lo, void *dst, int host_type);
extern thread_local std::string g_err;

// A roctx range for the lifetime of the object (rocprofv3 --marker-trace
// shows one per solver() call, rk45 attempt, stage launch and batch run;
// without a profiler attached a push/pop costs a few tens of ns).
struct Range {
  explicit Range(const char *name);
  ~Range();
  Range(const Range &) = delete;
  Range &operator=(const Range &) = delete;
};

// One stored launch configuration: stage-kind name, tile variant, z-chunk.
struct TuneItem {
  std::string name;
  int variant = 0, zchunk = 0;
};
// The entry stored under key in the launch-configuration table at path
// (tune_table.cpp).  A missing file or key leaves found false and returns
// GPE_OK; an unreadable or malformed file returns an error.
int read_tune_table(const char *path, const char *key, std::vector<TuneItem> &out, bool &found);
}  // namespace gpe
