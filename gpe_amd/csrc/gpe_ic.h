// gpe_ic.h -- the random-phase initial state built on the device (gpe_ic.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "host_common.h"

namespace gpe {
// Modes of the random-phase field: flat indices (ascending) of every mode
// k = 2 pi m/(n d) with |k| <= kmax, and their unit phasors (cos, sin of
// 2 pi u, u from SplitMix64(seed) in that order).
int random_phase_modes(int nx, int ny, int nz, double dx, double dy, double dz, double kmax, uint64_t seed,
                       std::vector<long long> &idx, std::vector<double> &re, std::vector<double> &im);
// The normalised field (<|psi|^2> = 1) in storage form prec into out (and
// lo for GPE_FP32C), on stream s; returns when it is written.
int random_phase_device(int nx, int ny, int nz, double dx, double dy, double dz, double kmax, uint64_t seed,
                        int prec, void *out, void *lo, hipStream_t s);
// In-place unnormalised 3-D DFT of a complex128 (nx, ny, nz) field, x
// fastest, with the cached plan of the grid: direction -1 forward
// (exp(-i k.x)), +1 backward; queued on stream s.
int fft_z2z(int nx, int ny, int nz, void *f, int direction, hipStream_t s);
}  // namespace gpe
