// ranges.cpp -- roctx ranges around the library's host-side phases (solver()
// calls, rk45 attempts, stage launches, batch runs), for rocprofv3
// --marker-trace timelines.
#include <rocprofiler-sdk-roctx/roctx.h>

#include "host_common.h"

namespace gpe {

Range::Range(const char *name) { roctxRangePush(name); }
Range::~Range() { roctxRangePop(); }

}  // namespace gpe
