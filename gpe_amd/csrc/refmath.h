// refmath.h -- the host libm operations the reference's FP64 path calls,
// restated for the device.
//
// The reference, compiled by flang, evaluates abs(z) of a complex as the
// host libm's hypot and x**y as its pow (solve.f90:184,197-229; abs also in
// get_rhs, solve.f90:430-436, and the rk45 error scale, 176-178).  The
// device library's own hypot and pow are accurate to an ulp but round
// differently from the host's on some arguments, and one ulp of the rk45
// error estimate moves dt at the 1e-11 level (the estimate is a difference
// of stage values ~1e6 times larger than itself), which then moves the
// whole trajectory.  These restatements remove that source of difference.
//
//   hypot_ref: the algorithm of the host libm's hypot (glibc 2.35, x86-64,
//     built without FMA): operands ordered and scaled into range, h =
//     sqrt(ax^2 + ay^2), then the correction of C. F. Borges, "An Improved
//     Algorithm for hypot(a,b)" (2019), h -= (t1 + t2)/(2h).  The CPU test
//     of this header compares a host build with the host libm's hypot on
//     tens of millions of arguments over the whole double range: equal in
//     every bit.
//   pow_cr: x**y rounded correctly, from log and exp in double-double
//     arithmetic (about 2^-95 relative) and one final rounding.  The host
//     libm's pow is accurate to about 0.52 ulp; it equals the correctly
//     rounded value on all but ~0.08% of the step control's arguments (the
//     CPU test counts them), where the single-simulation path, which calls
//     the host pow itself, and a batched instance can differ by one ulp
//     in dt.
//
// Both are built from IEEE double +, -, *, /, sqrt, fma, ldexp, frexp and
// rint, which round (or are exact) identically on the host and the device,
// so a host build of this header gives the device's bits.  Callers compile
// it with -ffp-contract=off (no contraction the host build would not do).
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define GPE_HD __host__ __device__ __forceinline__
#else
#define GPE_HD inline
#endif

namespace gpe {
namespace refm {

// --- hypot ------------------------------------------------------------------
// ax >= ay > 0, both in range: h = sqrt(ax^2 + ay^2), then the correction.
// Both correction formulas are evaluated and one is selected (no divergent
// branch on the device); each is the same expression either way.
GPE_HD double hypot_kernel(double ax, double ay) {
  const double h = std::sqrt(ax * ax + ay * ay);
  const bool near = h <= 2.0 * ay;
  const double d1 = h - ay, d2 = h - ax;
  const double t1a = ax * (2.0 * d1 - ax);
  const double t2a = (d1 - 2.0 * (ax - ay)) * d1;
  const double t1b = 2.0 * d2 * (ax - 2.0 * ay);
  const double t2b = (4.0 * d2 - ay) * ay + d2 * d2;
  const double t1 = near ? t1a : t1b, t2 = near ? t2a : t2b;
  return h - (t1 + t2) / (2.0 * h);
}

// Operands out of the kernel's range: non-finite, very large or very small,
// or so different in size that the smaller one does not count.
GPE_HD double hypot_edge(double x, double y) {
  constexpr double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
  if (!std::isfinite(x) || !std::isfinite(y)) {
    if (std::isinf(x) || std::isinf(y)) return INFINITY;
    return x + y;  // NaN
  }
  x = std::fabs(x);
  y = std::fabs(y);
  const double ax = x < y ? y : x, ay = x < y ? x : y;
  if (ax > kLarge) {
    if (ay <= ax * kEps) return ax + ay;
    return hypot_kernel(ax * kScale, ay * kScale) / kScale;
  }
  if (ay < kTiny) {
    if (ax >= ay / kEps) return ax + ay;
    return hypot_kernel(ax / kScale, ay / kScale) * kScale;
  }
  return ax + ay;  // ay <= ax * kEps
}

GPE_HD double hypot_ref(double x, double y) {
  constexpr double kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
  const double fx = std::fabs(x), fy = std::fabs(y);
  const double ax = fx < fy ? fy : fx, ay = fx < fy ? fx : fy;
  // false for NaN and infinite operands, which take the edge path
  if (ax <= kLarge && ay >= kTiny && ay > ax * kEps) return hypot_kernel(ax, ay);
  return hypot_edge(x, y);
}

// --- double-double arithmetic (Dekker 1971; Knuth TAOCP vol. 2, 4.2.2) -----
struct dd {
  double hi, lo;
};

GPE_HD dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
GPE_HD dd fast_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return {s, b - (s - a)};
}
GPE_HD dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, std::fma(a, b, -p)};
}
GPE_HD dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
GPE_HD dd neg(dd a) { return {-a.hi, -a.lo}; }
GPE_HD dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}
GPE_HD dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}
GPE_HD dd div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  dd r = add(a, neg(mul_d(b, q1)));
  const double q2 = r.hi / b.hi;
  r = add(r, neg(mul_d(b, q2)));
  const double q3 = r.hi / b.hi;
  return add(fast_two_sum(q1, q2), dd{q3, 0.0});
}
GPE_HD dd recip(double n) {  // 1/n (n a small integer)
  const double h = 1.0 / n;
  return {h, std::fma(-h, n, 1.0) / n};
}

// ln 2 in three parts (about 2^-160)
constexpr double kLn2a = 0x1.62e42fefa39efp-1;
constexpr double kLn2b = 0x1.abc9e3b39803fp-56;
constexpr double kLn2c = 0x1.7b57a079a1934p-111;

// log x (x > 0 finite): x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.1716,
// atanh(s)/s = sum_{k=0..22} s^(2k)/(2k+1) (remainder below 2^-110)
GPE_HD dd log_dd(double x) {
  int e = 0;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const dd s = div(dd{m - 1.0, 0.0}, two_sum(m, 1.0));  // m - 1 is exact
  const dd s2 = mul(s, s);
  constexpr int K = 22;
  dd p = recip(2.0 * K + 1.0);
  for (int k = K - 1; k >= 0; --k) p = add(mul(p, s2), recip(2.0 * k + 1.0));
  const dd lm = mul(s, p);
  const double de = (double)e;  // at most 11 bits: e*kLn2a is exact as a double-double
  dd l2 = add(two_prod(de, kLn2a), two_prod(de, kLn2b));
  l2 = add(l2, dd{de * kLn2c, 0.0});
  return add(l2, dd{2.0 * lm.hi, 2.0 * lm.lo});
}

// x**y rounded correctly; non-finite or non-positive x, non-finite y and
// results near the ends of the range take the library pow
GPE_HD double pow_cr(double x, double y) {
  if (!(x > 0.0) || !std::isfinite(x) || !std::isfinite(y)) return std::pow(x, y);
  if (x == 1.0 || y == 0.0) return 1.0;
  const dd L = log_dd(x);
  const dd t = add(two_prod(y, L.hi), dd{y * L.lo, 0.0});
  if (!(std::fabs(t.hi) < 700.0)) return std::pow(x, y);
  // exp t = 2^k exp(r), r = t - k ln2 (k has at most 11 bits)
  const double k = std::rint(t.hi * 1.4426950408889634);
  dd r = add(t, neg(two_prod(k, kLn2a)));
  r = add(r, neg(two_prod(k, kLn2b)));
  r = add(r, dd{-(k * kLn2c), 0.0});
  // expm1(a), a = r/2^10 (|a| < 3.4e-4), by its Taylor series to a^10/10!
  const dd a{std::ldexp(r.hi, -10), std::ldexp(r.lo, -10)};
  dd q = recip(3628800.0);
  for (int n = 9; n >= 1; --n) {
    double f = 1.0;  // n!, exact
    for (int j = 2; j <= n; ++j) f *= (double)j;
    q = add(mul(q, a), recip(f));
  }
  dd E = mul(q, a);
  // expm1(2u) = expm1(u) (expm1(u) + 2), ten times: E = expm1(r)
  for (int i = 0; i < 10; ++i) E = mul(E, add(E, dd{2.0, 0.0}));
  const dd v = add(dd{1.0, 0.0}, E);  // normalised: v.hi is v rounded to nearest
  return std::ldexp(v.hi, (int)k);
}

}  // namespace refm
}  // namespace gpe
