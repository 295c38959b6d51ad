// diag.hpp -- vortex diagnostics of a resident field: the vortex-line
// length (linelength, variables.f90:896-1162) and the vortex-core points
// (get_zeros, get_extra_zeros and get_re_im_zeros, io.f90:886-1393).
//
// Line length.
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

// ---- vortex-core points -----------------------------------------------------
// One thread per interior lower corner (i, j, k), 1 <= i,j,k <= n-2, as the
// reference's loops; every test of one cell is independent of the others,
// so hits are appended in any order (one atomic per point) and the host
// sorts them by (section, cell) into the reference's file order.
//
// which = 0 (the zeros file): sections 0..2 are the edges in x, y, z on
// which Re and Im both change sign, or vanish at the first corner
// (io.f90:924-931), the point being where the linear interpolant of Re
// vanishes; sections 3..5 are the xy, xz and yz faces around which Re and Im
// change sign on alternate edges (io.f90:1039-1098 and its xz/yz copies),
// the point being the crossing of the line through the two crossings of one
// component with the line through the two crossings of the other, formed
// with the reference's slopes.  which = 1 / 2 (re_zeros / im_zeros): edges
// on which that component vanishes at the first corner or changes sign; in z
// the reference joins the two tests with .and. (io.f90:1304, 1375), kept as
// written.  Arithmetic in T, one rounding per operation, in the reference's
// order (the translation unit is built without contraction).

template <typename T>
__device__ __forceinline__ T zr_interp(T w0, T w1, T a0, T a1, T d) {
  return -(w1 * a0 / d) + w0 * a1 / d;
}

template <typename T>
__device__ __forceinline__ bool zr_edge(T r0, T i0, T r1, T i1) {
  return (r0 == (T)0 && i0 == (T)0) || (r0 == (T)0 && i0 * i1 < (T)0) || (i0 == (T)0 && r0 * r1 < (T)0) ||
         (r0 * r1 < (T)0 && i0 * i1 < (T)0);
}

// corners u[1..4], v[1..4] around a face (a0,b0), (a1,b0), (a1,b1), (a0,b1):
// u changes sign on edges 1-2 and 3-4 only, v on 2-3 and 4-1 only
template <typename T>
__device__ __forceinline__ bool zr_face_pattern(const T *u, const T *v) {
  return (u[1] * u[2] < (T)0 && v[1] * v[2] >= (T)0) && (v[2] * v[3] < (T)0 && u[2] * u[3] >= (T)0) &&
         (u[3] * u[4] < (T)0 && v[3] * v[4] >= (T)0) && (v[4] * v[1] < (T)0 && u[4] * u[1] >= (T)0);
}

template <typename T>
__device__ __forceinline__ void zr_face_point(const T *u, const T *v, T a0, T a1, T b0, T b1, T &ap, T &bp) {
  const T z1 = zr_interp<T>(u[1], u[2], a0, a1, u[1] - u[2]);
  const T z2 = zr_interp<T>(v[2], v[3], b0, b1, v[2] - v[3]);
  const T z3 = zr_interp<T>(u[3], u[4], a1, a0, u[3] - u[4]);
  const T z4 = zr_interp<T>(v[4], v[1], b1, b0, v[4] - v[1]);
  const T m1 = (b0 - b1) / (z1 - z3);
  const T m2 = (z2 - z4) / (a1 - a0);
  const T a = (z4 - a0 * m2 - b0 + z1 * m1) / (m1 - m2);
  ap = a;
  bp = a * m1 + b0 - z1 * m1;
}

template <typename S, typename T>
__global__ void __launch_bounds__(256)
    zeros_cells(const S *__restrict__ f, const S *__restrict__ lo, int nx, int ny, int nz, double xl, double yl,
                double zl, double dxd, double dyd, double dzd, int which, ZeroPoint *__restrict__ out,
                int *__restrict__ count, int cap) {
  const long long cx = nx - 2, cy = ny - 2, cz = nz - 2;
  const long long ncell = cx * cy * cz;
  const T txl = (T)xl, tyl = (T)yl, tzl = (T)zl, dx = (T)dxd, dy = (T)dyd, dz = (T)dzd;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(c % cx);
    const long long r = c / cx;
    const int j = 1 + (int)(r % cy);
    const int k = 1 + (int)(r / cy);
    // corners (di, dj, dk) = bits (0, 1, 2) of q
    T re[8], im[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const size_t s = ((size_t)(k + ((q >> 2) & 1)) * ny + (j + ((q >> 1) & 1))) * nx + (i + (q & 1));
      if (lo) {
        re[q] = (T)((double)f[2 * s] + (double)lo[2 * s]);
        im[q] = (T)((double)f[2 * s + 1] + (double)lo[2 * s + 1]);
      } else {
        re[q] = (T)f[2 * s];
        im[q] = (T)f[2 * s + 1];
      }
    }
    // x(i) = xl + real(i)*dx (ic.f90:53-66)
    const T x0 = txl + (T)i * dx, x1 = txl + (T)(i + 1) * dx;
    const T y0 = tyl + (T)j * dy, y1 = tyl + (T)(j + 1) * dy;
    const T z0 = tzl + (T)k * dz, z1 = tzl + (T)(k + 1) * dz;
    const long long cell = ((long long)k * ny + j) * nx + i;
    const int nsec = which == 0 ? 6 : 3;
    for (int sec = 0; sec < nsec; ++sec) {
      T px = x0, py = y0, pz = z0;
      bool hit = false;
      if (sec < 3) {
        const int q1 = 1 << sec;  // the edge's far corner: +x, +y or +z
        T w0, w1;
        bool cond;
        if (which == 0) {
          w0 = re[0];
          w1 = re[q1];
          cond = zr_edge<T>(re[0], im[0], re[q1], im[q1]);
        } else {
          w0 = which == 1 ? re[0] : im[0];
          w1 = which == 1 ? re[q1] : im[q1];
          cond = sec < 2 ? (w0 == (T)0 || w0 * w1 < (T)0) : (w0 == (T)0 && w0 * w1 < (T)0);
        }
        if (cond) {
          const T d = w0 - w1;
          if (d != (T)0) {
            hit = true;
            if (sec == 0) px = zr_interp<T>(w0, w1, x0, x1, d);
            else if (sec == 1) py = zr_interp<T>(w0, w1, y0, y1, d);
            else pz = zr_interp<T>(w0, w1, z0, z1, d);
          }
        }
      } else {
        // c1..c4 as corner indices, and the face axes (a, b)
        int c2, c3, c4;
        T a0, a1, b0, b1;
        if (sec == 3) {         // i,j,k --> i+1,j,k --> i+1,j+1,k --> i,j+1,k
          c2 = 1; c3 = 3; c4 = 2;
          a0 = x0; a1 = x1; b0 = y0; b1 = y1;
        } else if (sec == 4) {  // i,j,k --> i+1,j,k --> i+1,j,k+1 --> i,j,k+1
          c2 = 1; c3 = 5; c4 = 4;
          a0 = x0; a1 = x1; b0 = z0; b1 = z1;
        } else {                // i,j,k --> i,j,k+1 --> i,j+1,k+1 --> i,j+1,k
          c2 = 4; c3 = 6; c4 = 2;
          a0 = z0; a1 = z1; b0 = y0; b1 = y1;
        }
        const T u[5] = {(T)0, re[0], re[c2], re[c3], re[c4]};
        const T v[5] = {(T)0, im[0], im[c2], im[c3], im[c4]};
        T ap = 0, bp = 0;
        if (zr_face_pattern<T>(u, v)) {
          zr_face_point<T>(u, v, a0, a1, b0, b1, ap, bp);
          hit = true;
        } else if (zr_face_pattern<T>(v, u)) {
          zr_face_point<T>(v, u, a0, a1, b0, b1, ap, bp);
          hit = true;
        }
        if (hit) {
          if (sec == 3) { px = ap; py = bp; }
          else if (sec == 4) { px = ap; pz = bp; }
          else { pz = ap; py = bp; }
        }
      }
      if (hit) {
        const int w = atomicAdd(count, 1);
        if (w < cap) out[w] = ZeroPoint{sec, 0, cell, (double)px, (double)py, (double)pz};
      }
    }
  }
}

}  // namespace gpe
