// stencil.hpp -- the fused stage kernel of the GP time-stepping path (gfx950).
//
// One workgroup = NW wave64s owns a TX x TY tile of the xy plane (TX = 64:
// one x-site per lane, so a row is one coalesced wave access; TY = NW*P:
// each wave owns P rows) and marches z over a chunk of planes.  For every
// plane it
//   * loads the new plane k+R of the stage input u = psi + sum coef*k_m,
//     reconstructed in registers from psi and the previous stages' k-fields
//     (so only one field is staged, whatever the stage), into a register
//     queue of the 2R+1 z-neighbours (R = order/2),
//   * writes plane k (queue centre) and its x/y ghost cells into a
//     double-buffered LDS tile (the ghost cells were loaded R planes earlier,
//     together with the neighbouring tile's load of the same rows),
//   * evaluates the Laplacian (x and y from LDS, z from the queue), the GP
//     right-hand side and the stage outputs for plane k,
//   * keeps the pointwise partial sums the final combination needs in a
//     small register FIFO (computed when the plane was loaded), so no field
//     is read twice.
// Boundaries are index maps applied at load time: periodic wrap (the
// one-rank halo exchange of variables.f90:203-375 and derivs.f90:286-301)
// or the mirror about the wall point (derivs.f90:302-313, 365-394).
//
// FP64 kernels are compiled with -ffp-contract=off and evaluate every
// expression in the reference's order with its complex semantics (complex
// division by a real in compiler-rt's logb-scaled form, abs() as hypot), so
// they reproduce the reference arithmetic operation by operation.  FP32
// kernels multiply by reciprocals of the constant divisors and are also
// built without contraction (one rounding per operation, so every
// instantiation of a stage gives the same bits); their final psi update and
// all reductions are FP64 (with an FP32 hi/lo pair when compensated).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "refmath.h"
#include "stage_args.h"

namespace gpe {

template <typename T>
using vec2_t = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;

template <typename T>
struct cx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) {
  cx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  T r1 = a.re * b.re, r2 = a.im * b.im, i1 = a.re * b.im, i2 = a.im * b.re;
  return mk(r1 - r2, i1 + i2);
}
template <typename T>
__device__ __forceinline__ cx<T> rmul(T s, cx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T>
__device__ __forceinline__ cx<double> todbl(cx<T> a) { return mk((double)a.re, (double)a.im); }

template <typename T>
__device__ __forceinline__ cx<T> ldc(const void *p, size_t i) {
  vec2_t<T> v = static_cast<const vec2_t<T> *>(p)[i];
  return mk<T>(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void stc(void *p, size_t i, cx<T> a) {
#ifdef GPE_NT_STORE
  // streaming outputs are not re-read in the sweep: keep them out of L2
  typedef T nv2 __attribute__((ext_vector_type(2)));
  nv2 v = {a.re, a.im};
  __builtin_nontemporal_store(v, static_cast<nv2 *>(p) + i);
#else
  vec2_t<T> v;
  v.x = a.re;
  v.y = a.im;
  static_cast<vec2_t<T> *>(p)[i] = v;
#endif
}

// complex / real constant
__device__ __forceinline__ cx<double> rdiv(cx<double> a, const DivC &d, double) {
  double x = (a.re * d.cs + a.im * 0.0) / d.den;
  double y = (a.im * d.cs - a.re * 0.0) / d.den;
  return mk(ldexp(x, -d.e), ldexp(y, -d.e));
}
__device__ __forceinline__ cx<float> rdiv(cx<float> a, const DivC &, double inv) {
  float f = (float)inv;
  return mk(a.re * f, a.im * f);
}

// complex / complex (rk45 error), compiler-rt's logb-scaled quotient
__device__ __forceinline__ cx<double> cdiv(cx<double> a, cx<double> b) {
  double c = b.re, d = b.im;
  double m = fmax(fabs(c), fabs(d));
  int e = 0;
  if (m != 0.0 && isfinite(m)) {
    e = ilogb(m);
    c = ldexp(c, -e);
    d = ldexp(d, -e);
  }
  double cc = c * c, dd = d * d;
  double den = cc + dd;
  double ac = a.re * c, bd = a.im * d, bc = a.im * c, ad = a.re * d;
  return mk(ldexp((ac + bd) / den, -e), ldexp((bc - ad) / den, -e));
}

// the same in single precision (__divsc3): the scaling keeps |b|^2 from
// underflowing when the rk45 error scale is as small as its 1e-30 floor
__device__ __forceinline__ cx<float> cdiv(cx<float> a, cx<float> b) {
  float c = b.re, d = b.im;
  float m = fmaxf(fabsf(c), fabsf(d));
  int e = 0;
  if (m != 0.0f && isfinite(m)) {
    e = ilogbf(m);
    c = ldexpf(c, -e);
    d = ldexpf(d, -e);
  }
  float cc = c * c, dd = d * d;
  float den = cc + dd;
  float ac = a.re * c, bd = a.im * d, bc = a.im * c, ad = a.re * d;
  return mk(ldexpf((ac + bd) / den, -e), ldexpf((bc - ad) / den, -e));
}

// abs() of a complex: hypot, as flang emits it (no underflow of |a|^2); in
// FP64 the host libm's hypot restated (refmath.h), so |.| rounds as in the
// reference's binary
__device__ __forceinline__ double cabs_(cx<double> a) { return refm::hypot_ref(a.re, a.im); }
__device__ __forceinline__ float cabs_(cx<float> a) { return hypotf(a.re, a.im); }

// y + b*x.  The compensated FP32 form (FMA) rounds once per component with an
// explicit fused multiply-add, so every instantiation of a stage (single,
// batched, fused pair) gives the same bits; otherwise the product and the sum
// round separately, in the reference's order.
template <bool FMA, typename T>
__device__ __forceinline__ cx<T> axpy(T b, cx<T> x, cx<T> y) {
  if constexpr (FMA && std::is_same<T, float>::value) {
    return mk(fmaf(b, x.re, y.re), fmaf(b, x.im, y.im));
  } else {
    return y + rmul(b, x);
  }
}

// |a| of an FP32 value through FP64 (no underflow of |a|^2) and the
// hardware square root (about 1 ulp): for the compensated form's error scale,
// which only sets the tolerance.
__device__ __forceinline__ float mag_fast(cx<float> a) {
  const double r = (double)a.re, i = (double)a.im;
  return (float)__builtin_amdgcn_sqrt(fma(r, r, i * i));
}

// |err/scale|^2 in FP64 (no underflow at the 1e-30 floor of the scale): the
// compensated form's rk45 error measure; the maximum's square root is taken
// once per thread.
__device__ __forceinline__ double err_ratio2(cx<float> e, cx<float> sc) {
  const double er = (double)e.re, ei = (double)e.im, sr = (double)sc.re, si = (double)sc.im;
  return fma(er, er, ei * ei) / fma(sr, sr, si * si);
}

// Index map for a ghost index i in [-R, n-1+R].
__device__ __forceinline__ int bcmap(int i, int n, int periodic) {
  if (periodic) return i < 0 ? i + n : (i >= n ? i - n : i);
  return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i);
}

// Extended closed-rule weight of index i on an axis of n points
// (integrate_x/y/z, variables.f90:742-748, 773-778, 818-823).
__device__ __forceinline__ double nweight(int i, int n, double w1, double w2, double w3) {
  if (i == 0 || i == n - 1) return w1;
  if (i == 1 || i == n - 2) return w2;
  if (i == 2 || i == n - 3) return w3;
  return 1.0;
}

// ---------------------------------------------------------------------------
// Laplacian in the reference's evaluation order (derivs.f90:164-181,
// 209-224, 251-266; variables.f90:68).
//
// FP32 forms the same sums from differences with the centre first:
// (f1 - f0) + (f-1 - f0) and 16[(f1 - f0) + (f-1 - f0)] - [(f2 - f0) + (f-2 - f0)].
// For a smooth field neighbouring values are within a factor 2, so each
// difference is exact (Sterbenz), and the sum no longer cancels ~64|f|
// down to ~theta^2 |f|: the FP32 Laplacian of a well-resolved field keeps
// ~1e-6 relative accuracy instead of ~1e-3, which keeps the rk45 error
// estimate (itself a ~1e-6 fraction of the stage values) above rounding
// noise.
template <typename T, int R>
__device__ __forceinline__ cx<T> d2(cx<T> m2, cx<T> m1, cx<T> c, cx<T> p1, cx<T> p2, const DivC &d,
                                    double inv) {
  if constexpr (std::is_same<T, double>::value) {
    if constexpr (R == 1) {
      cx<T> num = (p1 - rmul((T)2.0, c)) + m1;
      return rdiv(num, d, inv);
    } else {
      cx<T> num = (((mk(-p2.re, -p2.im) + rmul((T)16.0, p1)) - rmul((T)30.0, c)) + rmul((T)16.0, m1)) - m2;
      return rdiv(num, d, inv);
    }
  } else {
    if constexpr (R == 1) {
      return rdiv((p1 - c) + (m1 - c), d, inv);
    } else {
      cx<T> near = (p1 - c) + (m1 - c);
      cx<T> far = (p2 - c) + (m2 - c);
      return rdiv(rmul((T)16.0, near) - far, d, inv);
    }
  }
}

// deriv_x (derivs.f90:41-57)
template <typename T, int R>
__device__ __forceinline__ cx<T> d1(cx<T> m2, cx<T> m1, cx<T> p1, cx<T> p2, const DivC &d, double inv) {
  if constexpr (R == 1) {
    return rdiv(p1 - m1, d, inv);
  } else {
    cx<T> num = ((mk(-p2.re, -p2.im) + rmul((T)8.0, p1)) - rmul((T)8.0, m1)) + m2;
    return rdiv(num, d, inv);
  }
}

// get_rhs at one site (solve.f90:414-437).
template <typename T>
__device__ __forceinline__ cx<T> rhs_site(const StageArgs &A, cx<T> u, cx<T> lap, cx<T> dpx, int i, int j,
                                          int k) {
  const T x = (T)A.xl + (T)i * (T)A.dx;
  const T y = (T)A.yl + (T)j * (T)A.dy;
  const T z = (T)A.zl + (T)k * (T)A.dz;
  if (A.eqn == 4) {
    T V = (T)0.5 * (((T)A.ox2 * (x * x) + (T)A.oy2 * (y * y)) + (T)A.oz2 * (z * z));
    T s;
    if constexpr (std::is_same<T, double>::value) {
      cx<T> u2 = cmul(u, u);  // abs(in_var**2)
      s = cabs_(u2);
    } else {
      s = u.re * u.re + u.im * u.im;
    }
    cx<T> t1 = rmul((T)-0.5, lap);
    cx<T> t2 = rmul(V, u);
    cx<T> t3 = rmul((T)A.g, rmul(s, u));
    cx<T> t4 = rmul((T)A.mu, u);
    cx<T> in = ((t1 + t2) + t3) - t4;
    return mk(in.im, -in.re);  // -eye*(...)
  }
  T m = cabs_(u);
  T m2 = m * m;
  cx<T> r;
  if (A.eqn == 2) {
    cx<T> X = lap - rmul(m2, u);
    r = mk(-X.im, X.re);  // eye*(...)
  } else {
    T diss = (T)0.0;
    if (A.use_diss) {
      T tx = (T)1.0 + tanh(x - (T)A.a_diss) * tanh(x + (T)A.a_diss);
      T ty = (T)1.0 + tanh(y - (T)A.b_diss) * tanh(y + (T)A.b_diss);
      T tz = (T)1.0 + tanh(z - (T)A.c_diss) * tanh(z + (T)A.c_diss);
      diss = (T)A.diss_amp * ((tx + ty) + tz);
    }
    cx<T> X = lap + rmul((T)1.0 - m2, u);
    r = cmul(mk((T)0.5 * diss, (T)0.5), X);  // 0.5*(eye+diss)*(...)
  }
  if (A.use_urhs) r = r + rmul((T)A.urhs, dpx);
  if (A.eqn == 3) {
    T rr = sqrt((x * x + y * y) + z * z);
    T sph = (T)0.5 * ((T)1.0 + tanh((rr - (T)10.0) - (T)2.0));  // sphere(), ic.f90:640-662
    r = mk(r.re * sph, r.im * sph);
  }
  return r;
}

// ---------------------------------------------------------------------------
template <int KIND>
struct KindInfo {
  static constexpr int nin =
      (KIND == K_RK2_2 || KIND == K_RK4_2 || KIND == K_RK4_3 || KIND == K_RK4_4) ? 1
      : (KIND >= K_CK_2 && KIND <= K_CK_6)                                           ? (KIND - K_CK_1)
                                                                                     : 0;
  static constexpr bool complex_coef = (KIND == K_RK2_2 || KIND == K_RK4_2 || KIND == K_RK4_3 || KIND == K_RK4_4);
  static constexpr bool final_ = (KIND == K_EULER || KIND == K_RK2_2 || KIND == K_RK4_4 || KIND == K_CK_6);
  static constexpr bool is_ck = (KIND >= K_CK_1 && KIND <= K_CK_6);
  static constexpr bool has_acc = (KIND == K_RK4_4);
};

// Pointwise values the output plane needs from the fields read for its stage
// input (complex T per point), kept in a register FIFO of R+1 planes.
template <int KIND, bool COMP>
struct FifoInfo {
  static constexpr int nf = KIND == K_RK2_2   ? 1   // psi
                            : KIND == K_RK4_2 ? 1   // kk1/6
                            : KIND == K_RK4_4 ? 2   // psi, acc + kk3/3
                            : KIND == K_CK_6  ? (COMP ? 4 : 2)  // psi_hi, sum c kk, sum dc kk, scale | psi + sum c kk, sum dc kk
                                              : 0;
};

// The compensated form forms the rk45 error scale |psi| + dt_step |f(psi)|
// + 1e-30 in the last stage, from psi and kk1 = dt f(psi) it reads anyway
// (|f| = |kk1| / |dt|), instead of storing it in stage 1 and reading it
// back: one field pass less per attempt.  FP64 and plain FP32 store it as
// the reference computes it (solve.f90:176-178).
__device__ __forceinline__ cx<float> scale_from_kk1(const StageArgs &A, cx<float> ps, cx<float> kk1) {
  const float af = mag_fast(kk1) * (float)A.inv_absdt, au = mag_fast(ps);
  return mk(fmaf((float)A.sdtr, af, au) + 1e-30f, (float)A.sdti * af);
}

// Pointwise fields read only at the output plane: the rk45 error scale and
// the compensated low word.
template <int KIND, bool COMP>
struct OutInfo {
  static constexpr int no = ((KIND == K_CK_6 && !COMP) ? 1 : 0) + ((COMP && KindInfo<KIND>::final_) ? 1 : 0);
};

// Stage kinds reading at most this many fields per site prefetch two planes.
#ifndef GPE_PREFETCH2_MAX
#define GPE_PREFETCH2_MAX 2
#endif

// Workgroup barrier that waits only for this wave's LDS traffic: the global
// loads issued ahead for the next plane stay in flight across it.
#ifdef GPE_PLAIN_BARRIER
__device__ __forceinline__ void lds_barrier() { __syncthreads(); }
#else
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
#endif

// One tile (64 x TY columns, planes z0 .. z1-1) of one stage sweep.  The per-block
// partial sums go to slot `slot` of the partial arrays pn / pe / pen (any
// of them may be null).
template <typename T, int R, int NW, int P, int KIND, bool COMP>
__device__ __forceinline__ void stage_tile(const StageArgs &A, const int tx, const int ty, const int z0,
                                           const int z1, const int slot, double *pn, double *pe, double *pen) {
  constexpr int TX = 64, TY = NW * P, LX = TX + 2 * R, LY = TY + 2 * R;
  constexpr int NQ = 2 * R + 1;
  constexpr int NT = 64 * NW;
  using KI = KindInfo<KIND>;
  constexpr int NIN = KI::nin;
  constexpr int NINA = NIN > 0 ? NIN : 1;
  constexpr int NF = FifoInfo<KIND, COMP>::nf;
  constexpr int NFA = NF > 0 ? NF : 1;
  constexpr int NO = OutInfo<KIND, COMP>::no;
  constexpr int NOA = NO > 0 ? NO : 1;
  constexpr bool IS_DBL = std::is_same<T, double>::value;
  constexpr int NYH = 2 * R * TX, NXH = 2 * R * TY, NH = NYH + NXH;
  constexpr int HPT = (NH + NT - 1) / NT;
  using V = vec2_t<T>;

  __shared__ V lds[2][LY * LX];
  __shared__ double red[3][NW];

  struct Raw {
    cx<T> ps;
    cx<T> kk[NINA];
    cx<T> ac;
  };

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  const int tx0 = tx * TX, ty0 = ty * TY;

  const int nx = A.nx, ny = A.ny, nz = A.nz, per = A.periodic;
  const size_t plane = (size_t)nx * (size_t)ny;

  // Own points: column gi, rows gj[p].  Loads use the boundary-mapped offset
  // (always in range); stores happen only for points inside the grid.
  const int gi = tx0 + lane;
  const bool act_x = gi < nx;
  const int gim = bcmap(min(gi, nx - 1 + R), nx, per);
  int gj[P];
  int roff[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    gj[p] = ty0 + wave * P + p;
    const int gjm = bcmap(min(gj[p], ny - 1 + R), ny, per);
    roff[p] = gjm * nx + gim;
  }
  // Ghost cells this thread fills: fixed for the whole z march.
  int hl[HPT], hoff[HPT];
#pragma unroll
  for (int i = 0; i < HPT; ++i) {
    const int h = threadIdx.x + i * NT;
    hl[i] = -1;
    hoff[i] = 0;
    if (h < NH) {
      int lr, lc;
      if (h < NYH) {
        const int rs = h / TX, c = h - rs * TX;
        lr = rs < R ? rs : TY + rs;
        lc = c + R;
      } else {
        const int hh = h - NYH;
        const int r = hh / (2 * R), cs = hh - r * (2 * R);
        lr = r + R;
        lc = cs < R ? cs : TX + cs;
      }
      hl[i] = lr * LX + lc;
      const int hgi = bcmap(min(tx0 + lc - R, nx - 1 + R), nx, per);
      const int hgj = bcmap(min(ty0 + lr - R, ny - 1 + R), ny, per);
      hoff[i] = hgj * nx + hgi;
    }
  }

  const cx<T> H = KIND == K_RK4_4 ? mk((T)A.dtr, (T)A.dti) : mk((T)A.hr, (T)A.hi);
  // get_norm weights of this thread's column and rows (constant over the
  // z march); the product wx*wy*wz is formed in the same order as before
  const double wx = nweight(gi, nx, A.w1, A.w2, A.w3);
  double wy[P];
#pragma unroll
  for (int p = 0; p < P; ++p) wy[p] = nweight(gj[p], ny, A.w1, A.w2, A.w3);

  auto load_raw = [&](size_t s, Raw &r, const bool full) {
    r.ps = ldc<T>(A.psi, s);
    if constexpr (NIN > 0) {
#pragma unroll
      for (int m = 0; m < NIN; ++m) r.kk[m] = ldc<T>(A.kin[m], s);
    }
    if constexpr (KI::has_acc) {
      if (full) r.ac = ldc<T>(A.acc_in, s);
    }
  };
  // stage input u = psi + sum coef * k
  auto make_u = [&](const Raw &r) -> cx<T> {
    if constexpr (NIN == 0) {
      return r.ps;
    } else if constexpr (KI::complex_coef) {
      return r.ps + cmul(H, r.kk[0]);
    } else {
      cx<T> u = r.ps;
#pragma unroll
      for (int m = 0; m < NIN; ++m) u = axpy<COMP>((T)ck_b(KIND - K_CK_1 + 1, m), r.kk[m], u);
      return u;
    }
  };
  auto make_F = [&](const Raw &r, cx<T> *F) {
    if constexpr (KIND == K_RK2_2) {
      F[0] = r.ps;
    } else if constexpr (KIND == K_RK4_2) {
      F[0] = rdiv(r.kk[0], A.d6, 1.0 / 6.0);  // kk1/6
    } else if constexpr (KIND == K_RK4_4) {
      F[0] = r.ps;
      F[1] = r.ac + rdiv(r.kk[0], A.d3, 1.0 / 3.0);  // (kk1/6 + kk2/3) + kk3/3
    } else if constexpr (KIND == K_CK_6) {
      cx<T> pe = rmul((T)ck_dc(1), r.kk[0]);
#pragma unroll
      for (int m = 1; m < 5; ++m) pe = axpy<COMP>((T)ck_dc(m + 1), r.kk[m], pe);
      if constexpr (COMP) {
        cx<T> pd = rmul((T)ck_c(1), r.kk[0]);  // c2 = c5 = 0
        pd = axpy<COMP>((T)ck_c(3), r.kk[2], pd);
        pd = axpy<COMP>((T)ck_c(4), r.kk[3], pd);
        F[0] = r.ps;
        F[1] = pd;
        F[2] = pe;
        F[3] = scale_from_kk1(A, r.ps, r.kk[0]);
      } else {
        cx<T> pn = r.ps;
#pragma unroll
        for (int m = 0; m < 5; ++m) pn = pn + rmul((T)ck_c(m + 1), r.kk[m]);
        F[0] = pn;
        F[1] = pe;
      }
    }
  };
  auto load_out = [&](size_t s, cx<T> *o) {
    int i = 0;
    if constexpr (KIND == K_CK_6 && !COMP) {
      // real time: the scale is real (its imaginary part is 0*|f|), stored as one word
      if (A.scale_real)
        o[i++] = mk(static_cast<const T *>(A.scale_in)[s], (T)0);
      else
        o[i++] = ldc<T>(A.scale_in, s);
    }
    if constexpr (COMP && KI::final_) {
      // the low word is read at the tile's own sites only (no neighbour
      // reads it as a ghost): stream it past L2
      typedef T nv2 __attribute__((ext_vector_type(2)));
      const nv2 v = __builtin_nontemporal_load(static_cast<const nv2 *>(A.lo) + s);
      o[i++] = mk<T>(v.x, v.y);
    }
    (void)i;
  };

  cx<T> q[NQ][P];
  cx<T> F[R + 1][P][NFA];
  // Ghost-cell stage inputs ride a FIFO of R+1 planes too: they are loaded
  // with plane k+R, in the same iteration the neighbouring tile loads those
  // rows as its own interior, so the second read of each row finds it in L2.
  cx<T> G[R + 1][HPT];

  // Prime the queue with planes z0-R .. z0+R-1 (and the FIFOs with z0 ..).
#pragma unroll
  for (int m = 0; m < 2 * R; ++m) {
    const size_t zoff = (size_t)bcmap(z0 - R + m, nz, per) * plane;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      Raw r;
      load_raw(zoff + roff[p], r, m >= R);
      q[m][p] = make_u(r);
      if constexpr (NF > 0) {
        if (m >= R) make_F(r, F[m - R][p]);
      }
    }
    if (m >= R) {
#pragma unroll
      for (int i = 0; i < HPT; ++i) {
        if (hl[i] >= 0) {
          Raw r;
          load_raw(zoff + hoff[i], r, false);
          G[m - R][i] = make_u(r);
        } else {
          G[m - R][i] = mk((T)0, (T)0);
        }
      }
    }
  }

  // Loads in flight: interior and ghosts of plane k+R (two planes ahead
  // for the stages that read few fields per site, whose single plane of
  // prefetch would leave HBM idle between planes), output-time fields of
  // plane k.  The single-precision fourth Cash-Karp stage with one row per
  // wave also takes two planes: it fits them in 147 VGPRs and its 512^3
  // sweep ran 1.27 -> 1.03 ms (with two rows per wave two planes were
  // slower, 1.08 -> 1.29 ms).
  constexpr int NREAD = NIN + (KI::has_acc ? 1 : 0) + 1;
  constexpr bool kD2Ck4 = KIND == K_CK_4 && P == 1 && !IS_DBL;
  constexpr int D = (NREAD <= GPE_PREFETCH2_MAX || kD2Ck4) ? 2 : 1;
  Raw rn[D][P];
  Raw hn[D][HPT];
  cx<T> on[P][NOA];
  auto issue_raw = [&](int k, Raw *rb, Raw *hb) {
    const size_t zi = (size_t)bcmap(k + R, nz, per) * plane;
#pragma unroll
    for (int p = 0; p < P; ++p) load_raw(zi + roff[p], rb[p], true);
#pragma unroll
    for (int i = 0; i < HPT; ++i)
      if (hl[i] >= 0) load_raw(zi + hoff[i], hb[i], false);
  };
  auto issue_out = [&](int k) {
    if constexpr (NO > 0) {
#pragma unroll
      for (int p = 0; p < P; ++p) load_out((size_t)k * plane + roff[p], on[p]);
    }
  };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (z0 + d < z1) issue_raw(z0 + d, rn[d], hn[d]);
  issue_out(z0);

  double acc_norm = 0.0, acc_emax = 0.0, acc_en = 0.0;

  // one plane of the march; rb/hb hold plane k+R and are refilled with
  // plane k+D+R
  auto step = [&](const int k, Raw *rb, Raw *hb) {
    V *Ld = lds[(k - z0) & 1];
    // consume what was prefetched for plane k
#pragma unroll
    for (int p = 0; p < P; ++p) {
      q[2 * R][p] = make_u(rb[p]);
      if constexpr (NF > 0) make_F(rb[p], F[R][p]);
    }
#pragma unroll
    for (int i = 0; i < HPT; ++i) G[R][i] = hl[i] >= 0 ? make_u(hb[i]) : mk((T)0, (T)0);
    cx<T> oc[P][NOA];
    if constexpr (NO > 0) {
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int o = 0; o < NO; ++o) oc[p][o] = on[p][o];
    }
    // prefetch while plane k is computed
    if (k + D < z1) issue_raw(k + D, rb, hb);
    if (k + 1 < z1) issue_out(k + 1);
    // stage input of plane k into LDS: ghosts and interior
#pragma unroll
    for (int i = 0; i < HPT; ++i) {
      if (hl[i] >= 0) {
        V v;
        v.x = G[0][i].re;
        v.y = G[0][i].im;
        Ld[hl[i]] = v;
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      V v;
      v.x = q[R][p].re;
      v.y = q[R][p].im;
      Ld[(wave * P + p + R) * LX + lane + R] = v;
    }
    lds_barrier();

#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int c = (wave * P + p + R) * LX + lane + R;
      auto LD = [&](int off) {
        V v = Ld[c + off];
        return mk<T>(v.x, v.y);
      };
      const cx<T> c0 = q[R][p];
      cx<T> xm1 = LD(-1), xp1 = LD(1), ym1 = LD(-LX), yp1 = LD(LX);
      cx<T> xm2 = xm1, xp2 = xp1, ym2 = ym1, yp2 = yp1;
      const cx<T> zm2 = q[0][p], zp2 = q[2 * R][p];
      if constexpr (R == 2) {
        xm2 = LD(-2);
        xp2 = LD(2);
        ym2 = LD(-2 * LX);
        yp2 = LD(2 * LX);
      }
      const cx<T> zm1 = q[R - 1][p], zp1 = q[R + 1][p];
      const cx<T> dxx = d2<T, R>(xm2, xm1, c0, xp1, xp2, A.d_xx, A.inv_xx);
      const cx<T> dyy = d2<T, R>(ym2, ym1, c0, yp1, yp2, A.d_yy, A.inv_yy);
      const cx<T> dzz = d2<T, R>(zm2, zm1, c0, zp1, zp2, A.d_zz, A.inv_zz);
      const cx<T> lap = (dxx + dyy) + dzz;
      const bool act = act_x && (gj[p] < ny);
      const size_t s = (size_t)k * plane + (size_t)gj[p] * nx + gi;

      if constexpr (KIND == K_ENERGY) {
        if (act) {
          cx<T> pr = cmul(mk(-lap.re, -lap.im), mk(c0.re, -c0.im));
          T m = cabs_(c0);
          T m2 = m * m;
          acc_en += (double)(pr.re + (T)0.5 * (m2 * m2));
        }
      } else {
        cx<T> dpx = mk((T)0, (T)0);
        if (A.use_urhs) dpx = d1<T, R>(xm2, xm1, xp1, xp2, A.d_x, A.inv_x);
        const cx<T> f = rhs_site<T>(A, c0, lap, dpx, gi, gj[p], k);
        if (act) {
          if constexpr (KIND == K_RHS || KIND == K_RK2_1 || KIND == K_RK4_1 || KIND == K_RK4_3) {
            stc<T>(A.kout, s, f);
          } else if constexpr (KIND == K_RK4_2) {
            stc<T>(A.kout, s, f);
            stc<T>(A.acc_out, s, F[0][p][0] + rdiv(f, A.d3, 1.0 / 3.0));
          } else if constexpr (KI::is_ck) {
            cx<T> kk;  // kk = dt*f
            if constexpr (IS_DBL || !COMP) {
              kk = cmul(mk((T)A.dtr, (T)A.dti), f);
            } else {
              // FP64 dt: no systematic bias from dt rounded to FP32
              cx<double> kd = cmul(mk(A.dtr, A.dti), todbl(f));
              kk = mk((T)kd.re, (T)kd.im);
            }
            if constexpr (KIND == K_CK_1) {
              stc<T>(A.kout, s, kk);
              if constexpr (!COMP) {
                if (A.write_scale) {
                  // psi_scal = abs(in) + dt*abs(f) + 1e-30 (solve.f90:176-178)
                  T af = cabs_(f), au = cabs_(c0);
                  cx<T> sc = mk((au + (T)A.sdtr * af) + (T)1e-30, (T)A.sdti * af);
                  if (A.scale_real)
                    static_cast<T *>(A.scale_out)[s] = sc.re;
                  else
                    stc<T>(A.scale_out, s, sc);
                }
              }
            } else if constexpr (KIND < K_CK_6) {
              stc<T>(A.kout, s, kk);
            } else {
              // K_CK_6: new state and the error norm (solve.f90:365-375, 184)
              const cx<T> sc = COMP ? F[0][p][NF - 1] : oc[p][0];
              const cx<T> err = F[0][p][COMP ? 2 : 1] + rmul((T)ck_dc(6), kk);
              cx<double> np;
              if constexpr (IS_DBL) {
                np = F[0][p][0] + rmul((T)ck_c(6), kk);
                stc<T>(A.psi_out, s, np);
                const double e = cabs_(cdiv(err, sc));
                acc_emax = (e > acc_emax || e != e) ? e : acc_emax;
              } else {
                if constexpr (COMP) {
                  const cx<double> dl = todbl(F[0][p][1]) + rmul(ck_c(6), todbl(kk));
                  np = (todbl(F[0][p][0]) + todbl(oc[p][0])) + dl;
                  const cx<T> hi = mk((T)np.re, (T)np.im);
                  stc<T>(A.psi_out, s, hi);
                  stc<T>(A.lo_out, s, mk((T)(np.re - (double)hi.re), (T)(np.im - (double)hi.im)));
                  // |err/scale|^2, maximised; its square root after the march
                  const double e2 = err_ratio2(err, sc);
                  acc_emax = (e2 > acc_emax || e2 != e2) ? e2 : acc_emax;
                } else {
                  const cx<T> nf = F[0][p][0] + rmul((T)ck_c(6), kk);
                  stc<T>(A.psi_out, s, nf);
                  np = todbl(nf);
                  // abs(err/scale) (solve.f90:184) in single precision
                  const float e = cabs_(cdiv(err, sc));
                  acc_emax = ((double)e > acc_emax || e != e) ? (double)e : acc_emax;
                }
              }
              const double wn = wx * wy[p] * nweight(k, nz, A.w1, A.w2, A.w3);
              if constexpr (IS_DBL) {
                const double mm = cabs_(np);
                acc_norm += wn * (mm * mm);
              } else {
                acc_norm += wn * (np.re * np.re + np.im * np.im);
              }
            }
          } else if constexpr (KI::final_) {
            // K_EULER, K_RK2_2, K_RK4_4
            cx<T> ps, incr;  // psi' = ps + dt*incr
            if constexpr (KIND == K_EULER) {
              ps = c0;
              incr = f;
            } else if constexpr (KIND == K_RK2_2) {
              ps = F[0][p][0];
              incr = f;
            } else {
              ps = F[0][p][0];
              incr = F[0][p][1] + rdiv(f, A.d6, 1.0 / 6.0);
            }
            cx<double> np;
            if constexpr (IS_DBL) {
              np = ps + cmul(mk((T)A.dtr, (T)A.dti), incr);
              stc<T>(A.psi_out, s, np);
            } else if constexpr (COMP) {
              np = (todbl(ps) + todbl(oc[p][0])) + cmul(mk(A.dtr, A.dti), todbl(incr));
              const cx<T> hi = mk((T)np.re, (T)np.im);
              stc<T>(A.psi_out, s, hi);
              stc<T>(A.lo_out, s, mk((T)(np.re - (double)hi.re), (T)(np.im - (double)hi.im)));
            } else {
              const cx<T> nf = ps + cmul(mk((T)A.dtr, (T)A.dti), incr);
              stc<T>(A.psi_out, s, nf);
              np = todbl(nf);
            }
            const double wn = wx * wy[p] * nweight(k, nz, A.w1, A.w2, A.w3);
            if constexpr (IS_DBL) {
              const double mm = cabs_(np);
              acc_norm += wn * (mm * mm);
            } else {
              acc_norm += wn * (np.re * np.re + np.im * np.im);
            }
          }
        }
      }
    }

    // shift queue and FIFOs
#pragma unroll
    for (int m = 0; m < 2 * R; ++m)
#pragma unroll
      for (int p = 0; p < P; ++p) q[m][p] = q[m + 1][p];
#pragma unroll
    for (int m = 0; m < R; ++m)
#pragma unroll
      for (int i = 0; i < HPT; ++i) G[m][i] = G[m + 1][i];
    if constexpr (NF > 0) {
#pragma unroll
      for (int m = 0; m < R; ++m)
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
          for (int f = 0; f < NF; ++f) F[m][p][f] = F[m + 1][p][f];
    }
  };
  // unrolled by D so that every prefetch buffer has a fixed register home
  for (int k = z0; k < z1; k += D) {
    step(k, rn[0], hn[0]);
    if constexpr (D > 1) {
      if (k + 1 < z1) step(k + 1, rn[1], hn[1]);
    }
  }

  // The compensated form maximised |err/scale|^2: the maximum of |err/scale|
  // is its square root (monotone; NaN stays NaN).
  if constexpr (KIND == K_CK_6 && COMP) acc_emax = sqrt(acc_emax);
  // Deterministic block reductions: xor-shuffle within the wave, then the
  // waves in order.
  if constexpr (KI::final_ || KIND == K_ENERGY) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      acc_norm += __shfl_xor(acc_norm, o);
      acc_en += __shfl_xor(acc_en, o);
      double om = __shfl_xor(acc_emax, o);
      acc_emax = (om > acc_emax || om != om) ? om : acc_emax;
    }
    if (lane == 0) {
      red[0][wave] = acc_norm;
      red[1][wave] = acc_emax;
      red[2][wave] = acc_en;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double sn = 0.0, se = 0.0, mx = 0.0;
      for (int w = 0; w < NW; ++w) {
        sn += red[0][w];
        se += red[2][w];
        mx = (red[1][w] > mx || red[1][w] != red[1][w]) ? red[1][w] : mx;
      }
      if (pn) pn[slot] = sn;
      if (pe) pe[slot] = mx;
      if (pen) pen[slot] = se;
    }
  }
}

// Single simulation: one workgroup per tile.
template <typename T, int R, int NW, int P, int KIND, bool COMP>
__global__ void __launch_bounds__(64 * NW) stage_kernel(const StageArgs A) {
  // XCD-aware tile order: the hardware deals consecutive workgroups round
  // robin over the 8 XCDs; give each XCD a contiguous range of tiles so
  // neighbouring tiles (which share ghost rows) run on one L2.
  const int nb = A.nblocks;
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int q8 = nb >> 3, r8 = nb & 7;
  const int L = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
  const int tx = L % A.ntx;
  const int rest = L / A.ntx;
  const int ty = rest % A.nty;
  const int tz = rest / A.nty;
  const int z0 = tz * A.zc;
  stage_tile<T, R, NW, P, KIND, COMP>(A, tx, ty, z0, min(z0 + A.zc, A.nz), b, A.part_norm, A.part_emax,
                                      A.part_energy);
}

// Many simulations: a persistent grid walks a table of (instance, tile)
// work items.  Each instance's parameters (grid, coefficients, dt, field
// pointers) live in device memory, where the per-instance step control
// updates dt and swaps the state buffers between launches; instances that
// have finished are skipped.  Partials go to the item's slot of the batch
// arrays, so each instance's reduction order does not depend on the others.
template <typename T, int R, int NW, int P, int KIND, bool COMP>
__global__ void __launch_bounds__(64 * NW)
    stage_kernel_batched(const StageArgs *__restrict__ insts, const WorkItem *__restrict__ work, int nwork,
                         const int *__restrict__ active, double *pn, double *pe) {
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem it = work[w];
    if (active[it.inst])
      stage_tile<T, R, NW, P, KIND, COMP>(insts[it.inst], it.tx, it.ty, it.z0, it.z1, it.slot, pn, pe, nullptr);
    __syncthreads();  // the next item reuses the LDS tile
  }
}

// Sum / max of the per-block partials in a fixed order (one workgroup).
// out[0] = sum(part_norm) * scale_norm, out[1] = max(part_emax) / eps
// (maxval(...)/eps, solve.f90:184), out[2] = sum(part_energy) * scale_energy;
// hist (if set) gets out[0].
static __global__ void __launch_bounds__(256) reduce_partials(const double *pn, const double *pe, const double *pen, int n,
                                                       double scale_norm, double eps, double scale_energy,
                                                       double *out, double *hist) {
  __shared__ double s0[256], s1[256], s2[256];
  double a = 0.0, m = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    if (pn) a += pn[i];
    if (pe) m = (pe[i] > m || pe[i] != pe[i]) ? pe[i] : m;
    if (pen) c += pen[i];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = m;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      double o = s1[threadIdx.x + w];
      s1[threadIdx.x] = (o > s1[threadIdx.x] || o != o) ? o : s1[threadIdx.x];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0] * scale_norm;
    out[1] = s1[0] / eps;
    out[2] = s2[0] * scale_energy;
    if (hist) *hist = out[0];
  }
}

}  // namespace gpe
