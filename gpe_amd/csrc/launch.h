// launch.h -- host-side view of the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "stage_args.h"

namespace gpe {

// Tile shape of a stage kernel: NW waves, P rows per wave, so TY = NW*P rows
// of 64 x-sites.  Each variant is compiled into its own translation unit
// (gpe_stage.hip with -DGPE_VID) and the host picks one per stage kind
// (defaults below, or the autotuner's choice).
struct Tile {
  int nw, p;
};
constexpr int N_TILE_VARIANTS = 5;
constexpr Tile kTiles[N_TILE_VARIANTS] = {{4, 4}, {8, 2}, {16, 1}, {4, 2}, {8, 1}};

// FP64 Cash-Karp stages do not run the (4,4) and (8,2) tiles.  An earlier
// build gave wrong 4th-order FP64 RKF45 fields with exactly these two
// shapes (every other shape, scheme and precision matched the oracle), and
// no fault in the source was found.  They are the two shapes whose FP64
// 4th-order last-stage instantiation is allocated the full 256 VGPRs and
// still spills to scratch (484 and 292 B per lane in the single-simulation
// kernel, 624 and 372 B batched); (16,1) spills too, but under a 128-VGPR
// budget, and passes.  The FP64 Cash-Karp default is (8,1), which needs no
// scratch in any stage.
constexpr int kVidF64Ck = 4;
constexpr bool variant_allowed(bool fp64, int kind, int vid) {
  return vid >= 0 && vid < N_TILE_VARIANTS && !(fp64 && kind >= K_CK_1 && kind <= K_CK_6 && vid <= 1);
}

// Tile variant of a new context's stage kinds (the (8,2) tile), where allowed.
constexpr int kDefaultVid = 1;

// Tiling of one stage sweep over an nx x ny x nz grid with tile variant vid
// and z-chunk zc (0: automatic): sets A.ntx, A.nty, A.zc, A.nzc and
// A.nblocks and returns the number of workgroups.  Automatic chunks give
// enough workgroups to fill 256 CUs several times over, but stay long
// enough that re-reading 2R planes per chunk stays small: at least 32
// planes unless that leaves fewer than 1024 workgroups (small grids, where
// the z march's latency, not its traffic, sets the time).
inline int stage_grid(int nx, int ny, int nz, int vid, int zc, StageArgs &A) {
  const Tile t = kTiles[vid];
  const int TY = t.nw * t.p;
  A.ntx = (nx + 63) / 64;
  A.nty = (ny + TY - 1) / TY;
  const int tiles2d = A.ntx * A.nty;
  if (zc <= 0) {
    int nzc = (2048 + tiles2d - 1) / tiles2d;
    zc = (nz + nzc - 1) / nzc;
    if (zc < 32 && tiles2d * ((nz + 31) / 32) >= 1024) zc = 32;
  }
  if (zc < 8) zc = 8;
  if (zc > nz) zc = nz;
  A.zc = zc;
  A.nzc = (nz + zc - 1) / zc;
  A.nblocks = tiles2d * A.nzc;
  return A.nblocks;
}

// Launch stage kernel `kind` (radius 1 or 2) over A.nblocks workgroups with
// tile variant `vid`.  comp selects the compensated FP32 form.
hipError_t launch_stage(int prec_bytes, int vid, int kind, int radius, bool comp, const StageArgs &A,
                        hipStream_t s);

// Fused stage pairs (fused.hpp): pair kinds P_RK4_12 .. P_CK_56 and their
// tile variants.
constexpr int N_PAIR_KINDS = 5;
constexpr int N_PAIR_VARIANTS = 3;
// waves, rows per wave, planes of prefetch
struct PairTile {
  int nw, p, d;
};
constexpr PairTile kPairTiles[N_PAIR_VARIANTS] = {{8, 2, 2}, {16, 1, 2}, {16, 1, 4}};
// FP64 (the parity path) always runs the stages one by one: its fused
// kernels would spill ~1 KB per lane to scratch, gain nothing, and a
// batched 4th-order FP64 build gave results different from the
// stage-by-stage sweeps for reasons not found in the source.
constexpr int N_PAIR_VARIANTS_F64 = 0;
// A: stage A's descriptor, B: stage B's (B carries the grid: ntx/nty over
// the (64-2R) x (TY-2R) interiors, zc, nblocks).
hipError_t launch_pair(int prec_bytes, int vid, int pair, int radius, bool comp, const StageArgs &A,
                       const StageArgs &B, hipStream_t s);

// Batched launch over a work table (see stage_kernel_batched).
struct BatchLaunch {
  const StageArgs *insts;  // this kind's StageArgs of every instance slot
  const WorkItem *work;
  int nwork;
  const int *active;
  double *pn, *pe;  // partial sums of the final stage (null otherwise)
  int grid;
};
hipError_t launch_stage_batched(int prec_bytes, int vid, int kind, int radius, bool comp, const BatchLaunch &B,
                                hipStream_t s);
struct PairBatchLaunch {
  const StageArgs *instA, *instB;  // stage A's / B's StageArgs of every instance slot
  const WorkItem *work;
  int nwork;
  const int *active;
  double *pn, *pe;
  int grid;
};
hipError_t launch_pair_batched(int prec_bytes, int vid, int pair, int radius, bool comp, const PairBatchLaunch &L,
                               hipStream_t s);

// Per-instance step control after the final stage of a batched iteration.
hipError_t launch_batch_control(InstCtl *ctl, StageArgs *args, int cap, int *active, const int *range,
                                const double *pn, const double *pe, hipStream_t s);

// Reduce per-block partials to out[0..2] (see reduce_partials).
hipError_t launch_reduce(const double *pn, const double *pe, const double *pen, int n, double scale_norm,
                         double eps, double scale_energy, double *out, double *hist, hipStream_t s);

// Pointwise helpers over a whole field.  Partials: one double per block.
constexpr int PW_BLOCKS = 1024;
hipError_t launch_norm_f64(const void *psi, int nx, int ny, int nz, int weighted, double w1, double w2,
                           double w3, double *part, hipStream_t s);
// Block partials (PW_BLOCKS each) of sum |psi|^2, |psi|^4, V |psi|^2 (prec 0: FP64 storage).
hipError_t launch_moments(int prec, const void *psi, const void *lo, const StageArgs &A, double *p2, double *p4,
                          double *pv, hipStream_t s);
hipError_t launch_norm_f32(const void *psi, const void *lo, int nx, int ny, int nz, int weighted, double w1,
                           double w2, double w3, double *part, hipStream_t s);
hipError_t launch_renorm_f64(void *psi, size_t n, double a, DivC b, hipStream_t s);
hipError_t launch_renorm_f32(void *psi, void *lo, size_t n, double a, double b, hipStream_t s);

// Vortex-line length cells (diag.hpp): prec GPE_FP64 / GPE_FP32 / GPE_FP32C
// (lo = low word for the compensated form); nblocks partial sums into part.
hipError_t launch_linelength(int prec, const void *psi, const void *lo, int nx, int ny, int nz, double *part,
                             int nblocks, LinePoint *writers, LinePoint *singles, int *counts, int cap,
                             hipStream_t s);

// Vortex-core points (diag.hpp): which = 0 zeros, 1 re_zeros, 2 im_zeros;
// appends up to cap points to out and counts all of them in *count.
hipError_t launch_zeros(int prec, const void *psi, const void *lo, int nx, int ny, int nz, const double *xyzl,
                        const double *dxyz, int which, ZeroPoint *out, int *count, int cap, int nblocks,
                        hipStream_t s);

// imprint_vortex_line over the whole field (pointwise.hpp); prec as
// GPE_FP64 / GPE_FP32 / GPE_FP32C (lo = low word of the compensated form).
hipError_t launch_imprint(int prec, void *psi, void *lo, const LineArgs &L, hipStream_t s);

// HBM read / write / copy bandwidth (GB/s) of float4 arrays of `bytes`.
hipError_t probe_bandwidth(size_t bytes, int reps, double *gbs);

// per-variant entry points (one definition per translation unit)
#define GPE_DECL_VARIANT(P, V) \
  hipError_t launch_stage_##P##_v##V(int kind, int radius, bool comp, const StageArgs &A, hipStream_t s);
GPE_DECL_VARIANT(f64, 0)
GPE_DECL_VARIANT(f64, 1)
GPE_DECL_VARIANT(f64, 2)
GPE_DECL_VARIANT(f64, 3)
GPE_DECL_VARIANT(f64, 4)
GPE_DECL_VARIANT(f32, 0)
GPE_DECL_VARIANT(f32, 1)
GPE_DECL_VARIANT(f32, 2)
GPE_DECL_VARIANT(f32, 3)
GPE_DECL_VARIANT(f32, 4)
#undef GPE_DECL_VARIANT
#define GPE_DECL_BVARIANT(P, V)                                                                            \
  hipError_t launch_batched_##P##_v##V(int kind, int radius, bool comp, const BatchLaunch &B, hipStream_t s);
GPE_DECL_BVARIANT(f64, 0)
GPE_DECL_BVARIANT(f64, 1)
GPE_DECL_BVARIANT(f64, 2)
GPE_DECL_BVARIANT(f64, 3)
GPE_DECL_BVARIANT(f64, 4)
GPE_DECL_BVARIANT(f32, 0)
GPE_DECL_BVARIANT(f32, 1)
GPE_DECL_BVARIANT(f32, 2)
GPE_DECL_BVARIANT(f32, 3)
GPE_DECL_BVARIANT(f32, 4)
#undef GPE_DECL_BVARIANT
#define GPE_DECL_PVARIANT(P, V)                                                                              \
  hipError_t launch_pair_##P##_v##V(int pair, int radius, bool comp, const StageArgs &A, const StageArgs &B,  \
                                    hipStream_t s);                                                         \
  hipError_t launch_pair_batched_##P##_v##V(int pair, int radius, bool comp, const PairBatchLaunch &L,       \
                                            hipStream_t s);
GPE_DECL_PVARIANT(f32, 0)
GPE_DECL_PVARIANT(f32, 1)
GPE_DECL_PVARIANT(f32, 2)
#undef GPE_DECL_PVARIANT

}  // namespace gpe
