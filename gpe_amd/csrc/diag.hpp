// diag.hpp -- vortex-line length of a resident field (linelength,
// variables.f90:896-1162).
//
// Every interior cell (corners i..i+1, j..j+1, k..k+1, 1 <= i,j,k <= n-2)
// is independent except for one thing: the reference keeps its point arrays
// from cell to cell, so a cell that finds only one zero point measures the
// segment to the SECOND point of the most recent earlier cell (x-fastest
// order) that stored one.  The kernel therefore does the per-cell work in
// parallel -- faces, edge crossings, the intersection of the Re/Im zero
// lines, the special face/edge cases and every segment with two fresh
// points -- and lists the rest: cells that stored a second point (with that
// point) and cells that found exactly one (with it).  The host matches each
// one-point cell with the latest earlier second point and adds those
// segments.  Arithmetic is in T (double for FP64 and the compensated form,
// float for plain FP32), one rounding per operation, in the reference's
// order.
#pragma once
#include <hip/hip_runtime.h>

#include "stage_args.h"

namespace gpe {

template <typename T>
struct LLFace {
  // Edge crossings of one component w[1..5] (w[5] = w[1]) of a face in face
  // coordinates; returns the reference's count index.
  __device__ static int crossings(const T w[6], T *xs, T *ys) {
    int l = 1;
#pragma unroll
    for (int q = 1; q <= 4; ++q) {
      if (w[q] == (T)0) {
        // a zero corner is recorded without advancing the index
        xs[l] = (q == 2 || q == 3) ? (T)1 : (T)0;
        ys[l] = (q >= 3) ? (T)1 : (T)0;
      } else if (w[q] * w[q + 1] < (T)0) {
        if (q == 1) {
          xs[l] = fabs(w[1] / (w[1] - w[2]));
          ys[l] = (T)0;
        } else if (q == 2) {
          xs[l] = (T)1;
          ys[l] = fabs(w[2] / (w[2] - w[3]));
        } else if (q == 3) {
          xs[l] = fabs(w[4] / (w[3] - w[4]));
          ys[l] = (T)1;
        } else {
          xs[l] = (T)0;
          ys[l] = fabs(w[5] / (w[4] - w[5]));
        }
        if (l == 1 || xs[l] != xs[l - 1] || ys[l] != ys[l - 1]) ++l;
      }
    }
    return l;
  }
};

template <typename T>
__device__ __forceinline__ bool ll_in01(T v) {
  return v == (T)0 || v == (T)1;
}

template <typename T>
__device__ __forceinline__ T ll_segment(T x1, T y1, T z1, T x2, T y2, T z2) {
  T a = (T)1;
  const bool ex = x1 == x2 && ll_in01(x1);
  const bool ey = y1 == y2 && ll_in01(y1);
  const bool ez = z1 == z2 && ll_in01(z1);
  if (ex || ey) {  // the reference tests y twice and z never here
    a = a * (T)0.5;
    if ((ex && ey) || (ey && ez)) a = a * (T)0.5;
  }
  return a * sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
}

// S: stored real type (double or float); T: arithmetic type; lo: low word
// of the compensated form (or null).
template <typename S, typename T>
__global__ void __launch_bounds__(256)
    linelength_cells(const S *__restrict__ f, const S *__restrict__ lo, int nx, int ny, int nz,
                     double *__restrict__ part, LinePoint *__restrict__ writers, LinePoint *__restrict__ singles,
                     int *__restrict__ counts, int cap) {
  const long long cx = nx - 2, cy = ny - 2, cz = nz - 2;
  const long long ncell = cx * cy * cz;
  double acc = 0.0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(c % cx);
    const long long r = c / cx;
    const int j = 1 + (int)(r % cy);
    const int k = 1 + (int)(r / cy);
    // the 8 corners, (di, dj, dk) = bits (0, 1, 2) of the index
    T cu[8], cv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const size_t s = ((size_t)(k + ((q >> 2) & 1)) * ny + (j + ((q >> 1) & 1))) * nx + (i + (q & 1));
      if (lo) {
        cu[q] = (T)((double)f[2 * s] + (double)lo[2 * s]);
        cv[q] = (T)((double)f[2 * s + 1] + (double)lo[2 * s + 1]);
      } else {
        cu[q] = (T)f[2 * s];
        cv[q] = (T)f[2 * s + 1];
      }
    }
    // corners u(1..4) of faces 1..6 (front k, back k+1, right i+1, left i,
    // top j+1, bottom j) as indices into the cube's 8 corners
    constexpr int face[6][4] = {{0, 1, 3, 2}, {4, 5, 7, 6}, {1, 5, 7, 3}, {0, 4, 6, 2}, {2, 3, 7, 6}, {0, 1, 5, 4}};
    int m = 1;
    bool done = false;
    T px1 = 0, py1 = 0, pz1 = 0, px2 = 0, py2 = 0, pz2 = 0;
    for (int n = 0; n < 6 && !done; ++n) {
      T u[6], v[6];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        u[q + 1] = cu[face[n][q]];
        v[q + 1] = cv[face[n][q]];
      }
      u[5] = u[1];
      v[5] = v[1];
      const bool uz = u[1] == 0 && u[2] == 0 && u[3] == 0 && u[4] == 0;
      const bool vz = v[1] == 0 && v[2] == 0 && v[3] == 0 && v[4] == 0;
      const bool vs = v[1] * v[2] < 0 || v[2] * v[3] < 0 || v[3] * v[4] < 0 || v[4] * v[1] < 0;
      const bool us = u[1] * u[2] < 0 || u[2] * u[3] < 0 || u[3] * u[4] < 0 || u[4] * u[1] < 0;
      if ((uz && vs) || (vz && us)) {  // a face of zeros
        acc += 0.5;
        done = true;
        break;
      }
#pragma unroll
      for (int q = 1; q <= 4; ++q)
        if (!done && u[q] == 0 && u[q + 1] == 0 && v[q] == 0 && v[q + 1] == 0) {  // an edge of zeros
          acc += 0.25;
          done = true;
        }
      if (done) break;
      T xu[7], yu[7], xv[7], yv[7];
      const int lu = LLFace<T>::crossings(u, xu, yu);
      const int lv = LLFace<T>::crossings(v, xv, yv);
      if (lu > 2 && lv > 2) {
        T den = xv[2] * (yu[1] - yu[2]) + xv[1] * (yu[2] - yu[1]) + (xu[1] - xu[2]) * (yv[1] - yv[2]);
        if (den == (T)0) den = den + (T)0.0000001;
        const T xx = (xu[1] * (xv[2] * (yv[1] - yu[2]) + xv[1] * (yu[2] - yv[2])) +
                      xu[2] * (xv[2] * (yu[1] - yv[1]) + xv[1] * (yv[2] - yu[1]))) / den;
        const T yy = (xv[2] * (yu[1] - yu[2]) * yv[1] + xu[1] * yu[2] * yv[1] - xv[1] * yu[1] * yv[2] -
                      xu[1] * yu[2] * yv[2] + xv[1] * yu[2] * yv[2] + xu[2] * yu[1] * (yv[2] - yv[1])) / den;
        if (xx >= 0 && xx <= 1 && yy >= 0 && yy <= 1) {
          T qx, qy, qz;
          switch (n) {
            case 0: qx = xx; qy = yy; qz = 0; break;
            case 1: qx = xx; qy = yy; qz = 1; break;
            case 2: qx = 1; qy = yy; qz = xx; break;
            case 3: qx = 0; qy = yy; qz = xx; break;
            case 4: qx = xx; qy = 1; qz = yy; break;
            default: qx = xx; qy = 0; qz = yy; break;
          }
          if (m == 1) {
            px1 = qx;
            py1 = qy;
            pz1 = qz;
          } else if (m == 2) {
            px2 = qx;
            py2 = qy;
            pz2 = qz;
          }
          ++m;
        }
      }
    }
    if (m >= 3) {
      // this cell stored a second point: later one-point cells may use it
      const int w = atomicAdd(&counts[0], 1);
      if (w < cap) writers[w] = LinePoint{c, (double)px2, (double)py2, (double)pz2};
      if (!done) acc += (double)ll_segment<T>(px1, py1, pz1, px2, py2, pz2);
    } else if (m == 2 && !done) {
      const int w = atomicAdd(&counts[1], 1);
      if (w < cap) singles[w] = LinePoint{c, (double)px1, (double)py1, (double)pz1};
    }
  }
  // deterministic block sum: xor-shuffle within the wave, waves in order
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

}  // namespace gpe
