// launch.h -- host-side view of the kernels in gpe_fp64.hip / gpe_fp32.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "stage_args.h"

namespace gpe {

// Tile shape of a stage kernel: NW waves, P rows per wave (TY = NW*P rows
// of 64 x-sites).  The launchers and the host grid computation both use
// this table.
struct Tile {
  int nw, p;
};
Tile tile_shape(int prec_bytes, int radius, int kind);

// Launch stage kernel `kind` (radius 1 or 2) over A.nblocks workgroups.
hipError_t launch_stage_f64(int kind, int radius, const StageArgs &A, hipStream_t s);
hipError_t launch_stage_f32(int kind, int radius, bool comp, const StageArgs &A, hipStream_t s);

// Reduce per-block partials to out[0..2] (see reduce_partials).
hipError_t launch_reduce(const double *pn, const double *pe, const double *pen, int n, double scale_norm,
                         double eps, double scale_energy, double *out, double *hist, hipStream_t s);

// Pointwise helpers over a whole field.  Partials: one double per block.
constexpr int PW_BLOCKS = 1024;
hipError_t launch_norm_f64(const void *psi, int nx, int ny, int nz, int weighted, double w1, double w2,
                           double w3, double *part, hipStream_t s);
hipError_t launch_norm_f32(const void *psi, const void *lo, int nx, int ny, int nz, int weighted, double w1,
                           double w2, double w3, double *part, hipStream_t s);
hipError_t launch_renorm_f64(void *psi, size_t n, double a, DivC b, hipStream_t s);
hipError_t launch_renorm_f32(void *psi, void *lo, size_t n, double a, double b, hipStream_t s);

}  // namespace gpe
