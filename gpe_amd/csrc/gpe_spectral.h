// gpe_spectral.h -- FFT-side diagnostics on the device (gpe_spectral.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "host_common.h"

namespace gpe {
// a = the unitary forward DFT of the field (storage form prec; lo for the
// compensated form), as the reference's fft(.., 'forward') leaves it.
int spectral_forward(int nx, int ny, int nz, int prec, const void *psi, const void *lo, double2 *a, hipStream_t s);
// Occupation per octave shell (sums eta[7] and counts nharm[7]) and the
// per-radius sums bins[nb] of |a|^2 (F(kk) is their prefix sum).  part and
// cnt are device scratch of nz (7 + nb) doubles and nz * 7 counts.
int spectral_spectrum(int nx, int ny, int nz, const double2 *a, int nb, double *part, long long *cnt,
                      long long *nharm, double *eta, double *bins, hipStream_t s);
// The k-space sum of condensed_particles for condensate density rho0; dev
// is device scratch of nx + ny + 2 nz doubles.
int spectral_condensed_sum(int nx, int ny, int nz, double dx, double dy, double dz, double rho0, double *dev,
                           double *tot, hipStream_t s);
// a <- unitary inverse DFT of a * max(1 - m k^2 / kc^2, 0), in place.
int spectral_filter(int nx, int ny, int nz, double2 *a, double m, double kc, hipStream_t s);
// Device scratch (in doubles) the calls above need besides the field.
size_t spectral_scratch_doubles(int nx, int ny, int nz);
}  // namespace gpe
