/*
 * gpe_oracle.c -- CPU restatement of the reference's GP time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests compare the HIP library with it; it is
 * never part of the product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Compiled twice (see oracle/Makefile): -DREAL=double -DPFX=orc64 and
 * -DREAL=float -DPFX=orc32, both with -ffp-contract=off so every operation
 * is one IEEE rounding in the order the reference's Fortran expressions
 * evaluate (left to right, no fused multiply-add).
 *
 * What follows which reference lines:
 *   index maps        derivs.f90:276-398 (second_index, fourth_index); y/z
 *                     periodic ghosts come from the one-rank halo exchange
 *                     variables.f90:203-375, i.e. a plain wrap
 *   deriv_xx/yy/zz    derivs.f90:146-272;  laplacian variables.f90:55-71
 *   deriv_x           derivs.f90:25-64
 *   grid / Vtrap      ic.f90:32-69, ic.f90:713-731;  sphere ic.f90:640-662
 *   get_rhs           solve.f90:382-440
 *   euler/rk2/rk4     solve.f90:57-154
 *   rkqs / rkck       solve.f90:158-378
 *   get_norm          variables.f90:554-576, integrate_x/y/z 730-840
 *   energy / mass     variables.f90:580-628
 *   renormalise       variables.f90:715-726
 *   linelength        variables.f90:896-1162
 *   zeros             io.f90:886-1393 (get_zeros, get_extra_zeros,
 *                     get_re_im_zeros)
 *   imprint_vortex_line  variables.f90:1166-1180 with vortex_line, amp,
 *                     get_theta, get_r (ic.f90:314-353, 735-889)
 *
 * Complex arithmetic follows what this image's flang (the compiler the
 * reference is built with here) emits for the reference's expressions,
 * checked value by value against flang-built code:
 *   - abs() of a complex is the C library hypot()/hypotf();
 *   - every complex division, including complex / real (Fortran promotes the
 *     real divisor to complex), is the logb-scaled textbook quotient of
 *     compiler-rt's __divdc3/__divsc3: scale c, d by 2^-logb(max|c|,|d|),
 *     then ((a*c + b*d)/(c*c + d*d), (b*c - a*d)/(c*c + d*d)) scaled back;
 *   - complex * real and complex + real equal the componentwise forms;
 *   - z**2 equals z*z = (x*x - y*y, x*y + y*x).
 * Other deliberate choices:
 *   - energy() sums row -> plane -> volume instead of one flat sum(); the
 *     Fortran SUM intrinsic leaves the order to the processor, and the
 *     hierarchical order keeps the oracle accurate to ~1e-15 at 512^3.
 *   - the rk45 attempt loop fails on a NaN anywhere in the error field, and
 *     after 1000 rejected attempts, instead of spinning (the reference has no
 *     guard, solve.f90:180-212).
 *
 * How each part is pinned (tests/test_oracle*.py, tests/test_refmath.py,
 * tests/test_linelength_oracle.py, tests/test_zeros_oracle.py,
 * tests/test_imprint_oracle.py): the stencils and index maps bit for bit
 * against the reference's derivs.f90 compiled unchanged (golden vectors in
 * tests/golden/ref_derivs.npz); the right-hand sides, integrators, step
 * control, norm, energy and mass by plane-wave and cosine-mode closed forms,
 * the uniform stationary state and scalar restatements of the integrators;
 * linelength, the zeros passes and imprint_vortex_line by closed forms and
 * line-by-line transcriptions.  Parity unpinned, at the last-bit level only:
 * the order of energy()'s flat sum (left to the processor by the Fortran
 * standard) and glibc tanh against another libm's (eqns 1 and 3 only).
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#include "gpe_oracle.h"

#ifndef REAL
#define REAL double
#define PFX orc64
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(PFX, name)

#define IS_FLOAT (sizeof(REAL) == 4)
static inline REAL r_sqrt(REAL v) { return IS_FLOAT ? (REAL)sqrtf((float)v) : (REAL)sqrt((double)v); }
static inline REAL r_tanh(REAL v) { return IS_FLOAT ? (REAL)tanhf((float)v) : (REAL)tanh((double)v); }
static inline REAL r_pow(REAL a, REAL b) { return IS_FLOAT ? (REAL)powf((float)a, (float)b) : (REAL)pow((double)a, (double)b); }
static inline REAL r_abs(REAL v) { return v < 0 ? -v : v; }

/* Step-control pow.  The reference calls the C library's pow (accurate to
 * about 0.52 ulp, not correctly rounded) for errmax**dt_decrease and
 * errmax**dt_increase (solve.f90:197-220); that is the default here.  After
 * PFX_set_control_pow(1) the double build rounds those two powers correctly
 * instead: libquadmath's powq (113-bit significand) rounded to double, with
 * a counter of results lying so close to a double rounding boundary that
 * the conversion might not be the correctly rounded value (expected never).
 * A step control on the device, which cannot call the host library (the
 * batched instances), rounds pow correctly; this mode pins it bit for bit.
 * errcon (solve.f90:173) keeps the C library's pow in both modes. */
static int g_ctl_pow_cr = 0;
static long g_ctl_pow_ambiguous = 0;
static double pow_cr_quad(double x, double y) {
  __float128 r = powq((__float128)x, (__float128)y);
  double d = (double)r;
  if (isfinite(d) && d != 0.0) {
    __float128 up = ((__float128)d + (__float128)nextafter(d, INFINITY)) / 2;
    __float128 dn = ((__float128)d + (__float128)nextafter(d, -INFINITY)) / 2;
    __float128 dist = fminq(fabsq(r - up), fabsq(r - dn));
    if (dist <= fabsq(r) * 0x1p-100Q) g_ctl_pow_ambiguous++;
  }
  return d;
}
static inline REAL ctl_pow(REAL a, REAL b) {
  if (!IS_FLOAT && g_ctl_pow_cr) return (REAL)pow_cr_quad((double)a, (double)b);
  return r_pow(a, b);
}
void FN(set_control_pow)(int correctly_rounded) { g_ctl_pow_cr = correctly_rounded; }
double FN(control_pow)(double x, double y) { return (double)ctl_pow((REAL)x, (REAL)y); }
long FN(control_pow_ambiguous)(void) { return g_ctl_pow_ambiguous; }

typedef struct { REAL re, im; } cx;

static inline cx cx_make(REAL r, REAL i) { cx c; c.re = r; c.im = i; return c; }
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
static inline cx cx_mul(cx a, cx b) {
  REAL r1 = a.re * b.re, r2 = a.im * b.im, i1 = a.re * b.im, i2 = a.im * b.re;
  return cx_make(r1 - r2, i1 + i2);
}
static inline cx cx_scale(REAL s, cx a) { return cx_make(s * a.re, s * a.im); }
static inline REAL r_hypot(REAL a, REAL b) {
  return IS_FLOAT ? (REAL)hypotf((float)a, (float)b) : (REAL)hypot((double)a, (double)b);
}
static inline REAL r_ldexp(REAL v, int e) { return IS_FLOAT ? (REAL)ldexpf((float)v, e) : (REAL)ldexp((double)v, e); }
static inline REAL cx_absv(cx a) { return r_hypot(a.re, a.im); }
/* (a + ib) / (c + id), logb-scaled as compiler-rt's __divdc3/__divsc3. */
static inline cx cx_div(cx a, cx b) {
  REAL c = b.re, d = b.im;
  REAL m = r_abs(c) > r_abs(d) ? r_abs(c) : r_abs(d);
  int e = 0;
  if (m != (REAL)0.0 && isfinite((double)m)) {
    e = ilogb((double)m);
    c = r_ldexp(c, -e);
    d = r_ldexp(d, -e);
  }
  REAL cc = c * c, dd = d * d;
  REAL den = cc + dd;
  REAL ac = a.re * c, bd = a.im * d, bc = a.im * c, ad = a.re * d;
  REAL x = r_ldexp((ac + bd) / den, -e);
  REAL y = r_ldexp((bc - ad) / den, -e);
  return cx_make(x, y);
}
/* complex / real: Fortran promotes the divisor to (s, 0). */
static inline cx cx_rdiv(cx a, REAL s) { return cx_div(a, cx_make(s, (REAL)0.0)); }

static inline cx ld(const REAL *f, long s) { return cx_make(f[2 * s], f[2 * s + 1]); }
static inline void st(REAL *f, long s, cx v) { f[2 * s] = v.re; f[2 * s + 1] = v.im; }

/* ------------------------------------------------------------------------ */
/* Context: index maps, coordinates and constants in precision REAL.        */

typedef struct {
  const orc_params *p;
  int nx, ny, nz;
  long n;
  int *mx[4], *my[4], *mz[4]; /* minus1, minus2, plus1, plus2 per axis */
  REAL *x, *y, *z;
  REAL dx, dy, dz, dx2, dy2, dz2, den_x, den_y, den_z;
  REAL g, mu, ox2, oy2, oz2, urhs, diss_amp;
  REAL a_diss, b_diss, c_diss;
  int nt;
} ctx;

/* second_index / fourth_index (derivs.f90:276-398) for one index.  For the
 * y and z axes in periodic mode the reference returns indx+-1, indx+-2 and
 * reads the ghost planes that the one-rank halo exchange filled with the
 * wrapped planes (variables.f90:215-248); that is the wrap below. */
static void bc_index(int bcs, int order, int indx, int n1, int xaxis, int out[4]) {
  int minus1 = indx - 1, minus2 = indx - 2, plus1 = indx + 1, plus2 = indx + 2;
  int n = n1; /* the reference passes nx1 = nx-1 as n */
  if (order == 2) {
    if (bcs == 1) {
      if (indx == 0) { minus1 = n; plus1 = 1; }
      if (indx == n) { minus1 = n - 1; plus1 = 0; }
      (void)xaxis;
    } else {
      if (indx == 0) { minus1 = 1; plus1 = 1; }
      if (indx == n) { minus1 = n - 1; plus1 = n - 1; }
    }
    minus2 = minus1; plus2 = plus1; /* unused for order 2 */
  } else {
    if (bcs == 1) {
      if (indx == 0) { minus1 = n; minus2 = n - 1; plus1 = 1; plus2 = 2; }
      if (indx == 1) { minus1 = 0; minus2 = n; plus1 = 2; plus2 = 3; }
      if (indx == n - 1) { minus1 = n - 2; minus2 = n - 3; plus1 = n; plus2 = 0; }
      if (indx == n) { minus1 = n - 1; minus2 = n - 2; plus1 = 0; plus2 = 1; }
    } else {
      if (indx == 0) { minus1 = 1; minus2 = 2; plus1 = 1; plus2 = 2; }
      if (indx == 1) { minus1 = 0; minus2 = 1; plus1 = 2; plus2 = 3; }
      if (indx == n - 1) { minus1 = n - 2; minus2 = n - 3; plus1 = n; plus2 = n - 1; }
      if (indx == n) { minus1 = n - 1; minus2 = n - 2; plus1 = n - 1; plus2 = n - 2; }
    }
  }
  out[0] = minus1; out[1] = minus2; out[2] = plus1; out[3] = plus2;
}

static int *axis_maps(const orc_params *p, int n, int xaxis, int *m[4]) {
  int *buf = (int *)malloc(sizeof(int) * 4 * (size_t)n);
  for (int q = 0; q < 4; q++) m[q] = buf + (size_t)q * n;
  for (int i = 0; i < n; i++) {
    int o[4];
    bc_index(p->bcs, p->order, i, n - 1, xaxis, o);
    for (int q = 0; q < 4; q++) m[q][i] = o[q];
  }
  return buf;
}

static void ctx_init(ctx *c, const orc_params *p) {
  c->p = p;
  c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
  c->n = (long)p->nx * p->ny * p->nz;
  axis_maps(p, p->nx, 1, c->mx);
  axis_maps(p, p->ny, 0, c->my);
  axis_maps(p, p->nz, 0, c->mz);
  c->dx = (REAL)p->dx; c->dy = (REAL)p->dy; c->dz = (REAL)p->dz;
  c->dx2 = c->dx * c->dx; c->dy2 = c->dy * c->dy; c->dz2 = c->dz * c->dz; /* ic.f90:49-51 */
  c->den_x = (REAL)12.0 * c->dx2; c->den_y = (REAL)12.0 * c->dy2; c->den_z = (REAL)12.0 * c->dz2;
  c->x = (REAL *)malloc(sizeof(REAL) * (size_t)(p->nx + p->ny + p->nz));
  c->y = c->x + p->nx; c->z = c->y + p->ny;
  REAL xl = (REAL)p->xl, yl = (REAL)p->yl, zl = (REAL)p->zl;
  for (int i = 0; i < p->nx; i++) c->x[i] = xl + (REAL)i * c->dx; /* ic.f90:53-56 */
  for (int j = 0; j < p->ny; j++) c->y[j] = yl + (REAL)j * c->dy;
  for (int k = 0; k < p->nz; k++) c->z[k] = zl + (REAL)k * c->dz;
  c->g = (REAL)p->g; c->mu = (REAL)p->mu;
  REAL ox = (REAL)p->omx, oy = (REAL)p->omy, oz = (REAL)p->omz;
  c->ox2 = ox * ox; c->oy2 = oy * oy; c->oz2 = oz * oz;
  c->urhs = (REAL)p->urhs; c->diss_amp = (REAL)p->diss_amp;
  c->a_diss = c->x[p->nx - 3]; c->b_diss = c->y[p->ny - 3]; c->c_diss = c->z[p->nz - 3]; /* solve.f90:397-399 */
  c->nt = p->nthreads > 0 ? p->nthreads : 1;
}

static void ctx_free(ctx *c) {
  free(c->mx[0]); free(c->my[0]); free(c->mz[0]); free(c->x);
}

static inline long IDX(const ctx *c, int i, int j, int k) {
  return (long)i + (long)c->nx * ((long)j + (long)c->ny * (long)k);
}

/* deriv_xx / deriv_yy / deriv_zz at one site, summed as dxx + dyy + dzz. */
static inline cx lap_site(const ctx *c, const REAL *f, int i, int j, int k) {
  long s = IDX(c, i, j, k);
  cx fc = ld(f, s);
  cx dxx, dyy, dzz;
  if (c->p->order == 2) {
    cx xp = ld(f, IDX(c, c->mx[2][i], j, k)), xm = ld(f, IDX(c, c->mx[0][i], j, k));
    cx yp = ld(f, IDX(c, i, c->my[2][j], k)), ym = ld(f, IDX(c, i, c->my[0][j], k));
    cx zp = ld(f, IDX(c, i, j, c->mz[2][k])), zm = ld(f, IDX(c, i, j, c->mz[0][k]));
    cx two = cx_scale((REAL)2.0, fc);
    dxx = cx_rdiv(cx_add(cx_sub(xp, two), xm), c->dx2);
    dyy = cx_rdiv(cx_add(cx_sub(yp, two), ym), c->dy2);
    dzz = cx_rdiv(cx_add(cx_sub(zp, two), zm), c->dz2);
  } else {
    cx thirty = cx_scale((REAL)30.0, fc);
#define D4(P2, P1, M1, M2, DEN)                                                \
  cx_rdiv(cx_sub(cx_add(cx_sub(cx_add(cx_make(-(P2).re, -(P2).im),             \
                                      cx_scale((REAL)16.0, P1)),               \
                               thirty),                                        \
                        cx_scale((REAL)16.0, M1)),                             \
                 M2),                                                          \
          DEN)
    cx xp1 = ld(f, IDX(c, c->mx[2][i], j, k)), xp2 = ld(f, IDX(c, c->mx[3][i], j, k));
    cx xm1 = ld(f, IDX(c, c->mx[0][i], j, k)), xm2 = ld(f, IDX(c, c->mx[1][i], j, k));
    cx yp1 = ld(f, IDX(c, i, c->my[2][j], k)), yp2 = ld(f, IDX(c, i, c->my[3][j], k));
    cx ym1 = ld(f, IDX(c, i, c->my[0][j], k)), ym2 = ld(f, IDX(c, i, c->my[1][j], k));
    cx zp1 = ld(f, IDX(c, i, j, c->mz[2][k])), zp2 = ld(f, IDX(c, i, j, c->mz[3][k]));
    cx zm1 = ld(f, IDX(c, i, j, c->mz[0][k])), zm2 = ld(f, IDX(c, i, j, c->mz[1][k]));
    dxx = D4(xp2, xp1, xm1, xm2, c->den_x);
    dyy = D4(yp2, yp1, ym1, ym2, c->den_y);
    dzz = D4(zp2, zp1, zm1, zm2, c->den_z);
#undef D4
  }
  return cx_add(cx_add(dxx, dyy), dzz); /* variables.f90:68 */
}

/* deriv_x (derivs.f90:25-64) at one site. */
static inline cx dx_site(const ctx *c, const REAL *f, int i, int j, int k) {
  if (c->p->order == 2) {
    cx xp = ld(f, IDX(c, c->mx[2][i], j, k)), xm = ld(f, IDX(c, c->mx[0][i], j, k));
    return cx_rdiv(cx_sub(xp, xm), (REAL)2.0 * c->dx);
  } else {
    cx xp1 = ld(f, IDX(c, c->mx[2][i], j, k)), xp2 = ld(f, IDX(c, c->mx[3][i], j, k));
    cx xm1 = ld(f, IDX(c, c->mx[0][i], j, k)), xm2 = ld(f, IDX(c, c->mx[1][i], j, k));
    cx a = cx_add(cx_make(-xp2.re, -xp2.im), cx_scale((REAL)8.0, xp1));
    a = cx_add(cx_sub(a, cx_scale((REAL)8.0, xm1)), xm2);
    return cx_rdiv(a, (REAL)12.0 * c->dx);
  }
}

/* get_rhs (solve.f90:382-440) at one site. */
static inline cx rhs_site(const ctx *c, const REAL *f, int i, int j, int k) {
  cx u = ld(f, IDX(c, i, j, k));
  cx lap = lap_site(c, f, i, j, k);
  int eqn = c->p->eqn;
  if (eqn == 4) {
    REAL x = c->x[i], y = c->y[j], z = c->z[k];
    REAL V = (REAL)0.5 * ((c->ox2 * (x * x) + c->oy2 * (y * y)) + c->oz2 * (z * z));
    cx u2 = cx_mul(u, u);              /* in_var**2 */
    REAL s = cx_absv(u2);              /* abs(in_var**2) */
    cx t1 = cx_scale((REAL)-0.5, lap); /* -0.5*laplacian */
    cx t2 = cx_scale(V, u);            /* Vtrap*psi */
    cx t3 = cx_scale(c->g, cx_scale(s, u)); /* g*(abs(psi**2)*psi) */
    cx t4 = cx_scale(c->mu, u);        /* mu*psi */
    cx inner = cx_sub(cx_add(cx_add(t1, t2), t3), t4);
    return cx_make(inner.im, -inner.re); /* -eye*(...) */
  }
  REAL m = cx_absv(u);
  REAL m2 = m * m; /* abs(psi)**2 */
  cx r;
  if (eqn == 1 || eqn == 3) {
    REAL diss = (REAL)0.0;
    if (c->diss_amp != (REAL)0.0) {
      REAL x = c->x[i], y = c->y[j], z = c->z[k];
      REAL tx = (REAL)1.0 + r_tanh(x - c->a_diss) * r_tanh(x + c->a_diss);
      REAL ty = (REAL)1.0 + r_tanh(y - c->b_diss) * r_tanh(y + c->b_diss);
      REAL tz = (REAL)1.0 + r_tanh(z - c->c_diss) * r_tanh(z + c->c_diss);
      diss = c->diss_amp * ((tx + ty) + tz); /* solve.f90:404-406 */
    }
    REAL om = (REAL)1.0 - m2;
    cx X = cx_add(lap, cx_scale(om, u));
    cx fac = cx_make((REAL)0.5 * diss, (REAL)0.5); /* 0.5*(eye+diss) */
    r = cx_mul(fac, X);
  } else { /* eqn 2 */
    cx X = cx_sub(lap, cx_scale(m2, u));
    r = cx_make(-X.im, X.re); /* eye*(...) */
  }
  if (c->urhs != (REAL)0.0) r = cx_add(r, cx_scale(c->urhs, dx_site(c, f, i, j, k)));
  if (eqn == 3) {
    REAL x = c->x[i], y = c->y[j], z = c->z[k];
    REAL rr = r_sqrt((x * x + y * y) + z * z);
    REAL sph = (REAL)0.5 * ((REAL)1.0 + r_tanh((rr - (REAL)10.0) - (REAL)2.0));
    r = cx_make(r.re * sph, r.im * sph);
  }
  return r;
}

/* ------------------------------------------------------------------------ */
/* Whole-field sweeps.  Every site is independent, so the OpenMP split over */
/* z-planes does not change any value.                                      */

typedef enum { OP_RHS, OP_LAP, OP_DX } field_op;

static void sweep(const ctx *c, field_op op, const REAL *f, REAL *out) {
  int nz = c->nz, ny = c->ny, nx = c->nx;
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        cx v;
        if (op == OP_RHS) v = rhs_site(c, f, i, j, k);
        else if (op == OP_LAP) v = lap_site(c, f, i, j, k);
        else v = dx_site(c, f, i, j, k);
        st(out, IDX(c, i, j, k), v);
      }
}

/* tmp = old + h*k (complex h), the stage input of rk2/rk4 (solve.f90:92,125). */
static void axpy_c(const ctx *c, cx h, const REAL *old, const REAL *k, REAL *tmp) {
  long n = c->n;
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long s = 0; s < n; s++) st(tmp, s, cx_add(ld(old, s), cx_mul(h, ld(k, s))));
}

/* ------------------------------------------------------------------------ */
/* Integrators.                                                             */

typedef struct {
  REAL *k[7]; /* k[1..6]; k[0] unused */
  REAL *tmp, *err, *scal;
} work;

static int work_alloc(work *w, long n) {
  memset(w, 0, sizeof(*w));
  size_t bytes = sizeof(REAL) * 2 * (size_t)n;
  for (int q = 1; q <= 6; q++) if (!(w->k[q] = (REAL *)malloc(bytes))) return -1;
  if (!(w->tmp = (REAL *)malloc(bytes))) return -1;
  if (!(w->err = (REAL *)malloc(bytes))) return -1;
  if (!(w->scal = (REAL *)malloc(bytes))) return -1;
  return 0;
}
static void work_free(work *w) {
  for (int q = 1; q <= 6; q++) free(w->k[q]);
  free(w->tmp); free(w->err); free(w->scal);
}

typedef struct { REAL dtr, dti, t, im_t; } tstate;

static void euler(const ctx *c, tstate *ts, const REAL *old, REAL *new_, work *w) {
  sweep(c, OP_RHS, old, w->k[1]);
  cx dt = cx_make(ts->dtr, ts->dti);
  long n = c->n;
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long s = 0; s < n; s++) st(new_, s, cx_add(ld(old, s), cx_mul(dt, ld(w->k[1], s))));
  ts->t = ts->t + ts->dtr;
  ts->im_t = ts->im_t + r_abs(ts->dti);
}

static void rk2(const ctx *c, tstate *ts, const REAL *old, REAL *new_, work *w) {
  cx dt = cx_make(ts->dtr, ts->dti);
  cx h = cx_make((REAL)0.5 * ts->dtr, (REAL)0.5 * ts->dti); /* 0.5_pr*dt */
  sweep(c, OP_RHS, old, w->k[1]);
  axpy_c(c, h, old, w->k[1], w->tmp);
  sweep(c, OP_RHS, w->tmp, w->k[2]);
  long n = c->n;
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long s = 0; s < n; s++) st(new_, s, cx_add(ld(old, s), cx_mul(dt, ld(w->k[2], s))));
  ts->t = ts->t + ts->dtr;
  ts->im_t = ts->im_t + r_abs(ts->dti);
}

static void rk4(const ctx *c, tstate *ts, const REAL *old, REAL *new_, work *w) {
  cx dt = cx_make(ts->dtr, ts->dti);
  cx h = cx_make((REAL)0.5 * ts->dtr, (REAL)0.5 * ts->dti);
  sweep(c, OP_RHS, old, w->k[1]);
  axpy_c(c, h, old, w->k[1], w->tmp);
  sweep(c, OP_RHS, w->tmp, w->k[2]);
  axpy_c(c, h, old, w->k[2], w->tmp);
  sweep(c, OP_RHS, w->tmp, w->k[3]);
  axpy_c(c, dt, old, w->k[3], w->tmp);
  sweep(c, OP_RHS, w->tmp, w->k[4]);
  long n = c->n;
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long s = 0; s < n; s++) { /* solve.f90:144-148 */
    cx a = cx_add(cx_rdiv(ld(w->k[1], s), (REAL)6.0), cx_rdiv(ld(w->k[2], s), (REAL)3.0));
    a = cx_add(a, cx_rdiv(ld(w->k[3], s), (REAL)3.0));
    a = cx_add(a, cx_rdiv(ld(w->k[4], s), (REAL)6.0));
    st(new_, s, cx_add(ld(old, s), cx_mul(dt, a)));
  }
  ts->t = ts->t + ts->dtr;
  ts->im_t = ts->im_t + r_abs(ts->dti);
}

/* Cash-Karp tableau, solve.f90:255-281, evaluated in precision REAL. */
typedef struct {
  REAL b[7][6];
  REAL c[7], dc[7];
} ck_tab;

static void ck_init(ck_tab *T) {
  memset(T, 0, sizeof(*T));
  T->b[2][1] = (REAL)0.2;
  T->b[3][1] = (REAL)0.075; T->b[3][2] = (REAL)0.225;
  T->b[4][1] = (REAL)0.3; T->b[4][2] = (REAL)-0.9; T->b[4][3] = (REAL)1.2;
  T->b[5][1] = (REAL)-11.0 / (REAL)54.0; T->b[5][2] = (REAL)2.5;
  T->b[5][3] = (REAL)-70.0 / (REAL)27.0; T->b[5][4] = (REAL)35.0 / (REAL)27.0;
  T->b[6][1] = (REAL)1631.0 / (REAL)55296.0; T->b[6][2] = (REAL)175.0 / (REAL)512.0;
  T->b[6][3] = (REAL)575.0 / (REAL)13824.0; T->b[6][4] = (REAL)44275.0 / (REAL)110592.0;
  T->b[6][5] = (REAL)253.0 / (REAL)4096.0;
  T->c[1] = (REAL)37.0 / (REAL)378.0; T->c[2] = (REAL)0.0; T->c[3] = (REAL)250.0 / (REAL)621.0;
  T->c[4] = (REAL)125.0 / (REAL)594.0; T->c[5] = (REAL)0.0; T->c[6] = (REAL)512.0 / (REAL)1771.0;
  T->dc[1] = T->c[1] - (REAL)2825.0 / (REAL)27648.0;
  T->dc[2] = T->c[2] - (REAL)0.0;
  T->dc[3] = T->c[3] - (REAL)18575.0 / (REAL)48384.0;
  T->dc[4] = T->c[4] - (REAL)13525.0 / (REAL)55296.0;
  T->dc[5] = T->c[5] - (REAL)277.0 / (REAL)14336.0;
  T->dc[6] = T->c[6] - (REAL)0.25;
}

/* rkck (solve.f90:248-378): new and err from old with the current dt. */
static void rkck(const ctx *c, const ck_tab *T, cx dt, const REAL *old, REAL *new_, REAL *err,
                 work *w) {
  long n = c->n;
  for (int s = 1; s <= 6; s++) {
    const REAL *in = (s == 1) ? old : w->tmp;
    sweep(c, OP_RHS, in, w->k[s]);
    REAL *ks = w->k[s];
#pragma omp parallel for schedule(static) num_threads(c->nt)
    for (long q = 0; q < n; q++) {
      st(ks, q, cx_mul(dt, ld(ks, q))); /* kk_s = dt*kk_s */
      if (s < 6) {
        cx a = ld(old, q);
        for (int m = 1; m <= s; m++) a = cx_add(a, cx_scale(T->b[s + 1][m], ld(w->k[m], q)));
        st(w->tmp, q, a);
      }
    }
  }
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long q = 0; q < n; q++) {
    cx a = ld(old, q);
    for (int m = 1; m <= 6; m++) a = cx_add(a, cx_scale(T->c[m], ld(w->k[m], q)));
    st(new_, q, a);
    cx e = cx_scale(T->dc[1], ld(w->k[1], q));
    for (int m = 2; m <= 6; m++) e = cx_add(e, cx_scale(T->dc[m], ld(w->k[m], q)));
    st(err, q, e);
  }
}

static REAL max_err(const ctx *c, const REAL *err, const REAL *scal) {
  int nz = c->nz;
  long plane = (long)c->nx * c->ny;
  REAL *pm = (REAL *)malloc(sizeof(REAL) * (size_t)nz);
  int bad = 0;
#pragma omp parallel for schedule(static) num_threads(c->nt) reduction(| : bad)
  for (int k = 0; k < nz; k++) {
    REAL m = (REAL)0.0;
    for (long q = (long)k * plane; q < (long)(k + 1) * plane; q++) {
      REAL v = cx_absv(cx_div(ld(err, q), ld(scal, q)));
      if (!(v == v)) bad = 1;
      if (v > m) m = v;
    }
    pm[k] = m;
  }
  REAL m = (REAL)0.0;
  for (int k = 0; k < nz; k++) if (pm[k] > m) m = pm[k];
  free(pm);
  if (bad) return (REAL)NAN;
  return m;
}

/* rkqs (solve.f90:158-244).  Returns 0, or -1 on a non-finite error norm
 * or more than 1000 rejected attempts. */
static int rkqs(const ctx *c, tstate *ts, const REAL *old, REAL *new_, work *w, double *errmax_out,
                int *nrej) {
  const orc_params *p = c->p;
  ck_tab T;
  ck_init(&T);
  REAL safety = (REAL)p->safety, eps = (REAL)p->eps;
  REAL dt_inc = (REAL)p->dt_increase, dt_dec = (REAL)p->dt_decrease;
  REAL errcon = r_pow((REAL)5.0 / safety, (REAL)1.0 / dt_inc); /* solve.f90:173 */
  long n = c->n;
  cx dt = cx_make(ts->dtr, ts->dti);
  /* psi_scal = abs(in) + dt*abs(f(in)) + 1e-30 (solve.f90:176-178) */
  sweep(c, OP_RHS, old, w->scal);
#pragma omp parallel for schedule(static) num_threads(c->nt)
  for (long q = 0; q < n; q++) {
    REAL af = cx_absv(ld(w->scal, q));
    REAL au = cx_absv(ld(old, q));
    cx d = cx_make(dt.re * af, dt.im * af);
    st(w->scal, q, cx_make((au + d.re) + (REAL)1e-30, d.im));
  }
  REAL errmax;
  int rejects = 0;
  for (;;) {
    rkck(c, &T, dt, old, new_, w->err, w);
    errmax = max_err(c, w->err, w->scal) / eps; /* solve.f90:184 */
    if (!(errmax == errmax)) return -1;
    if (errmax <= (REAL)1.0) break;
    if (++rejects > 1000) return -1;
    REAL pw = ctl_pow(errmax, dt_dec);
    cx dtt = cx_make(safety * dt.re, safety * dt.im);
    dtt = cx_make(dtt.re * pw, dtt.im * pw); /* safety*dt*(errmax**dt_decrease) */
    REAL adt = cx_absv(dt);
    REAL a = cx_absv(dtt), b = (REAL)0.1 * adt;
    REAL mx = a > b ? a : b;            /* sign(max(.,.), abs(dt)): abs(dt) >= 0 */
    if (!p->real_time) dt = cx_make((REAL)0.0, -mx);
    else dt = cx_make(mx, (REAL)0.0);
  }
  cx dt_next;
  if (errmax > errcon) {
    REAL pw = ctl_pow(errmax, dt_inc);
    cx a = cx_make(safety * dt.re, safety * dt.im);
    dt_next = cx_make(a.re * pw, a.im * pw);
  } else {
    dt_next = cx_make((REAL)5.0 * dt.re, (REAL)5.0 * dt.im);
  }
  ts->t = ts->t + dt.re;              /* solve.f90:225 */
  ts->dtr = dt_next.re; ts->dti = dt_next.im;
  ts->im_t = ts->im_t + r_abs(dt_next.im); /* uses the new dt (solve.f90:229) */
  *errmax_out = (double)errmax;
  *nrej = rejects;
  return 0;
}

static int solver_ws(const ctx *c, orc_state *s, const REAL *in, REAL *out, work *w) {
  tstate ts;
  ts.dtr = (REAL)s->dt_re; ts.dti = (REAL)s->dt_im; ts.t = (REAL)s->t; ts.im_t = (REAL)s->im_t;
  int rc = 0;
  s->n_rejects = 0;
  switch (c->p->scheme) {
    case 1: euler(c, &ts, in, out, w); break;
    case 2: rk2(c, &ts, in, out, w); break;
    case 4: rk4(c, &ts, in, out, w); break;
    case 45: rc = rkqs(c, &ts, in, out, w, &s->errmax, &s->n_rejects); break;
    default: return -2;
  }
  s->dt_re = ts.dtr; s->dt_im = ts.dti; s->t = ts.t; s->im_t = ts.im_t;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Public entry points.                                                     */

void FN(grid)(const orc_params *p, REAL *x, REAL *y, REAL *z) {
  ctx c;
  ctx_init(&c, p);
  memcpy(x, c.x, sizeof(REAL) * (size_t)p->nx);
  memcpy(y, c.y, sizeof(REAL) * (size_t)p->ny);
  memcpy(z, c.z, sizeof(REAL) * (size_t)p->nz);
  ctx_free(&c);
}

void FN(laplacian)(const orc_params *p, const REAL *f, REAL *lap) {
  ctx c; ctx_init(&c, p); sweep(&c, OP_LAP, f, lap); ctx_free(&c);
}

void FN(deriv_x)(const orc_params *p, const REAL *f, REAL *fx) {
  ctx c; ctx_init(&c, p); sweep(&c, OP_DX, f, fx); ctx_free(&c);
}

void FN(rhs)(const orc_params *p, const REAL *f, REAL *rhs) {
  ctx c; ctx_init(&c, p); sweep(&c, OP_RHS, f, rhs); ctx_free(&c);
}

int FN(solver)(const orc_params *p, orc_state *s, const REAL *in, REAL *out) {
  ctx c; ctx_init(&c, p);
  work w;
  int rc = work_alloc(&w, c.n);
  if (rc == 0) rc = solver_ws(&c, s, in, out, &w);
  work_free(&w);
  ctx_free(&c);
  return rc;
}

/* nsteps calls of solver() as the time loop of gpe.f90:136-280 makes them
 * (psi%old = psi%new; solver(psi%old, psi%new); p = p+1), with get_norm of
 * every new state into norms[] when norms is not NULL (gpe.f90:243-245). */
int FN(run)(const orc_params *p, orc_state *s, REAL *psi, int nsteps, double *norms) {
  ctx c; ctx_init(&c, p);
  work w;
  int rc = work_alloc(&w, c.n);
  REAL *out = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)c.n);
  if (!out) rc = -1;
  for (int q = 0; q < nsteps && rc == 0; q++) {
    rc = solver_ws(&c, s, psi, out, &w);
    if (rc) break;
    memcpy(psi, out, sizeof(REAL) * 2 * (size_t)c.n);
    s->p += 1;
    if (norms) norms[q] = FN(norm)(p, psi);
  }
  free(out);
  work_free(&w);
  ctx_free(&c);
  return rc;
}

/* get_norm (variables.f90:554-576) with integrate_x/y/z (730-840). */
double FN(norm)(const orc_params *p, const REAL *f) {
  int nx = p->nx, ny = p->ny, nz = p->nz;
  const REAL c1 = (REAL)3.0 / (REAL)8.0, c2 = (REAL)7.0 / (REAL)6.0, c3 = (REAL)23.0 / (REAL)24.0;
  REAL dx = (REAL)p->dx, dy = (REAL)p->dy, dz = (REAL)p->dz;
  REAL *xint = (REAL *)malloc(sizeof(REAL) * (size_t)ny * nz);
  int nt = p->nthreads > 0 ? p->nthreads : 1;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      const REAL *row = f + 2 * ((long)nx * ((long)j + (long)ny * k));
#define DEN(i) (cx_absv(cx_make(row[2 * (i)], row[2 * (i) + 1])) * cx_absv(cx_make(row[2 * (i)], row[2 * (i) + 1])))
      REAL mid = (REAL)0.0;
      for (int i = 3; i <= nx - 4; i++) mid = mid + DEN(i);
      REAL v = c1 * DEN(0) + c2 * DEN(1);
      v = v + c3 * DEN(2);
      v = v + mid;
      v = v + c3 * DEN(nx - 3);
      v = v + c2 * DEN(nx - 2);
      v = v + c1 * DEN(nx - 1);
#undef DEN
      xint[j + (long)ny * k] = v * dx;
    }
  REAL zsum = (REAL)0.0;
  for (int k = 0; k < nz; k++) {
    REAL t = (REAL)0.0;
    for (int j = 0; j < ny; j++) {
      REAL v = xint[j + (long)ny * k];
      if (j == 0) v = c1 * v;
      if (j == 1) v = c2 * v;
      if (j == 2) v = c3 * v;
      if (j == ny - 3) v = c3 * v;
      if (j == ny - 2) v = c2 * v;
      if (j == ny - 1) v = c1 * v;
      t = t + v;
    }
    REAL yint = t * dy;
    if (k == 0) yint = c1 * yint;
    if (k == 1) yint = c2 * yint;
    if (k == 2) yint = c3 * yint;
    if (k == nz - 3) yint = c3 * yint;
    if (k == nz - 2) yint = c2 * yint;
    if (k == nz - 1) yint = c1 * yint;
    zsum = zsum + yint;
  }
  free(xint);
  return (double)(zsum * dz);
}

/* energy (variables.f90:580-604): sum of Re(-lap*conj(psi)) + 0.5|psi|^4,
 * times dx*dy*dz. */
double FN(energy)(const orc_params *p, const REAL *f) {
  ctx c; ctx_init(&c, p);
  int nx = c.nx, ny = c.ny, nz = c.nz;
  REAL *plane = (REAL *)malloc(sizeof(REAL) * (size_t)nz);
#pragma omp parallel for schedule(static) num_threads(c.nt)
  for (int k = 0; k < nz; k++) {
    REAL ps = (REAL)0.0;
    for (int j = 0; j < ny; j++) {
      REAL rs = (REAL)0.0;
      for (int i = 0; i < nx; i++) {
        cx u = ld(f, IDX(&c, i, j, k));
        cx l = lap_site(&c, f, i, j, k);
        cx prod = cx_mul(cx_make(-l.re, -l.im), cx_make(u.re, -u.im));
        REAL m = cx_absv(u);
        REAL m2 = m * m;
        REAL v = prod.re + (REAL)0.5 * (m2 * m2);
        rs = rs + v;
      }
      ps = ps + rs;
    }
    plane[k] = ps;
  }
  REAL tot = (REAL)0.0;
  for (int k = 0; k < nz; k++) tot = tot + plane[k];
  free(plane);
  REAL E = ((tot * c.dx) * c.dy) * c.dz;
  ctx_free(&c);
  return (double)E;
}

/* mass (variables.f90:608-628). */
double FN(mass)(const orc_params *p, const REAL *f) {
  long n = (long)p->nx * p->ny * p->nz;
  long plane = (long)p->nx * p->ny;
  int nz = p->nz;
  REAL *pl = (REAL *)malloc(sizeof(REAL) * (size_t)nz);
  int nt = p->nthreads > 0 ? p->nthreads : 1;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int k = 0; k < nz; k++) {
    REAL s = (REAL)0.0;
    for (long q = (long)k * plane; q < (long)(k + 1) * plane; q++) {
      REAL m = cx_absv(ld(f, q));
      s = s + m * m;
    }
    pl[k] = s;
  }
  REAL tot = (REAL)0.0;
  for (int k = 0; k < nz; k++) tot = tot + pl[k];
  free(pl);
  (void)n;
  return (double)(((tot * (REAL)p->dx) * (REAL)p->dy) * (REAL)p->dz);
}

/* renormalise (variables.f90:715-726): var = sqrt(nn)*var/sqrt(norm). */
void FN(renormalise)(const orc_params *p, REAL *f, double nn, double norm) {
  long n = (long)p->nx * p->ny * p->nz;
  REAL a = r_sqrt((REAL)nn), b = r_sqrt((REAL)norm);
  for (long q = 0; q < n; q++) st(f, q, cx_rdiv(cx_scale(a, ld(f, q)), b));
}

/* ---- linelength (variables.f90:896-1162) ---------------------------------
 * Total vortex-line length: every interior cell (corners i..i+1, j..j+1,
 * k..k+1 with 1 <= i,j,k <= n-2) is tested face by face for a point where
 * the zero lines of Re(psi) and Im(psi) -- each traced as the segment
 * between its crossings of the face's edges -- intersect inside the face.
 * Two such points give a segment whose length is added (halved when it
 * lies in a face shared with a neighbouring cell, quartered on an edge); a
 * face that is entirely a zero of one component with a sign change of the
 * other adds 1/2 and ends the cell, and an edge that is a zero of both adds
 * 1/4 and ends the cell.  The result is scaled by dx.
 *
 * The point arrays of the reference keep their values from cell to cell, so
 * a cell in which only ONE point is found measures from that point to the
 * second point of the most recent earlier cell (in x-fastest order) that
 * stored one; this restatement keeps that behaviour, with (0, 0, 0) before
 * any cell has stored a second point. */
typedef struct {
  REAL x[7], y[7], z[7];
} ll_pts;

/* Corner offsets (di, dj, dk) of u(1..4) for faces 1..6: front (k), back
 * (k+1), right (i+1), left (i), top (j+1), bottom (j). */
static const int ll_corner[6][4][3] = {
    {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}, {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},
    {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}}, {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},
    {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}}, {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}}};

/* Edge crossings of one component w(1..5) (w5 = w1) of a face, in face
 * coordinates; returns the count index the reference calls lu / lv. */
static int ll_crossings(const REAL w[6], REAL *xs, REAL *ys) {
  int l = 1;
  for (int q = 1; q <= 4; q++) {
    if (w[q] == (REAL)0.0) {
      /* corner q; the index is not advanced (variables.f90:1025-1037) */
      static const REAL cxq[5] = {0, 0, 1, 1, 0}, cyq[5] = {0, 0, 0, 1, 1};
      xs[l] = cxq[q];
      ys[l] = cyq[q];
    } else if (w[q] * w[q + 1] < (REAL)0.0) {
      switch (q) {
        case 1: xs[l] = r_abs(w[1] / (w[1] - w[2])); ys[l] = (REAL)0.0; break;
        case 2: xs[l] = (REAL)1.0; ys[l] = r_abs(w[2] / (w[2] - w[3])); break;
        case 3: xs[l] = r_abs(w[4] / (w[3] - w[4])); ys[l] = (REAL)1.0; break;
        default: xs[l] = (REAL)0.0; ys[l] = r_abs(w[5] / (w[4] - w[5])); break;
      }
      if (l == 1 || xs[l] != xs[l - 1] || ys[l] != ys[l - 1]) l++;
    }
  }
  return l;
}

static int ll_in01(REAL v) { return v == (REAL)0.0 || v == (REAL)1.0; }

double FN(linelength)(const orc_params *p, const REAL *f) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  ll_pts P;
  memset(&P, 0, sizeof P);
  REAL L = (REAL)0.0;
  for (int k = 1; k <= nz - 2; k++)
    for (int j = 1; j <= ny - 2; j++)
      for (int i = 1; i <= nx - 2; i++) {
        int m = 1, done = 0;
        for (int n = 0; n < 6 && !done; n++) {
          REAL u[6], v[6];
          for (int q = 0; q < 4; q++) {
            const int *o = ll_corner[n][q];
            const long s = ((long)(k + o[2]) * ny + (j + o[1])) * nx + (i + o[0]);
            u[q + 1] = f[2 * s];
            v[q + 1] = f[2 * s + 1];
          }
          u[5] = u[1];
          v[5] = v[1];
          const int uz = u[1] == 0 && u[2] == 0 && u[3] == 0 && u[4] == 0;
          const int vz = v[1] == 0 && v[2] == 0 && v[3] == 0 && v[4] == 0;
          const int vs = v[1] * v[2] < 0 || v[2] * v[3] < 0 || v[3] * v[4] < 0 || v[4] * v[1] < 0;
          const int us = u[1] * u[2] < 0 || u[2] * u[3] < 0 || u[3] * u[4] < 0 || u[4] * u[1] < 0;
          if ((uz && vs) || (vz && us)) {  /* a face of zeros (variables.f90:981-993) */
            L = L + (REAL)0.5;
            done = 1;
            break;
          }
          for (int q = 1; q <= 4 && !done; q++)
            if (u[q] == 0 && u[q + 1] == 0 && v[q] == 0 && v[q + 1] == 0) {  /* an edge (995-1002) */
              L = L + (REAL)0.25;
              done = 1;
            }
          if (done) break;
          REAL xu[7], yu[7], xv[7], yv[7];
          const int lu = ll_crossings(u, xu, yu);
          const int lv = ll_crossings(v, xv, yv);
          if (lu > 2 && lv > 2) {
            REAL den = xv[2] * (yu[1] - yu[2]) + xv[1] * (yu[2] - yu[1]) + (xu[1] - xu[2]) * (yv[1] - yv[2]);
            if (den == (REAL)0.0) den = den + (REAL)0.0000001;
            const REAL xx = (xu[1] * (xv[2] * (yv[1] - yu[2]) + xv[1] * (yu[2] - yv[2])) +
                             xu[2] * (xv[2] * (yu[1] - yv[1]) + xv[1] * (yv[2] - yu[1]))) / den;
            const REAL yy = (xv[2] * (yu[1] - yu[2]) * yv[1] + xu[1] * yu[2] * yv[1] - xv[1] * yu[1] * yv[2] -
                             xu[1] * yu[2] * yv[2] + xv[1] * yu[2] * yv[2] + xu[2] * yu[1] * (yv[2] - yv[1])) / den;
            if (xx >= 0 && xx <= 1 && yy >= 0 && yy <= 1 && m <= 6) {
              REAL px, py, pz;
              switch (n) {
                case 0: px = xx; py = yy; pz = 0; break;
                case 1: px = xx; py = yy; pz = 1; break;
                case 2: px = 1; py = yy; pz = xx; break;
                case 3: px = 0; py = yy; pz = xx; break;
                case 4: px = xx; py = 1; pz = yy; break;
                default: px = xx; py = 0; pz = yy; break;
              }
              P.x[m] = px;
              P.y[m] = py;
              P.z[m] = pz;
              m++;
            }
          }
        }
        if (!done && m > 1) {
          REAL a = (REAL)1.0;
          const int ex = P.x[1] == P.x[2] && ll_in01(P.x[1]);
          const int ey = P.y[1] == P.y[2] && ll_in01(P.y[1]);
          const int ez = P.z[1] == P.z[2] && ll_in01(P.z[1]);
          if (ex || ey) {  /* (the reference tests y twice, z never, here) */
            a = a * (REAL)0.5;
            if ((ex && ey) || (ey && ez)) a = a * (REAL)0.5;
          }
          const REAL d = r_sqrt((P.x[1] - P.x[2]) * (P.x[1] - P.x[2]) + (P.y[1] - P.y[2]) * (P.y[1] - P.y[2]) +
                                (P.z[1] - P.z[2]) * (P.z[1] - P.z[2]));
          L = L + a * d;
        }
      }
  return (double)(L * (REAL)p->dx);
}

/* ---- vortex cores: get_zeros, get_extra_zeros, get_re_im_zeros -----------
 * (io.f90:886-1236 and 1237-1393.)  Every interior lower corner (i, j, k)
 * with 1 <= i <= nx-2, 1 <= j <= ny-2, 1 <= k <= nz-2 is tested; output is
 * in the reference's file order: section by section, k slowest, i fastest.
 *
 * which = 0, the zeros<p>.dat file, six sections:
 *   0..2  edges (i,j,k)->(i+1,j,k), ->(i,j+1,k), ->(i,j,k+1) on which Re and
 *         Im both change sign (or vanish at the first corner, io.f90:924-931);
 *         the point is where the linear interpolant of Re vanishes; an edge
 *         with equal Re at both ends is skipped;
 *   3..5  faces xy, xz, yz (corners in the order the section header lists)
 *         around which Re and Im change sign on alternate edges, one of two
 *         patterns (io.f90:1039-1098 and the xz/yz copies); the point is the
 *         intersection of the line through the two Re crossings with the
 *         line through the two Im crossings, formed with the reference's
 *         slopes m(1), m(2) (no guard against a zero denominator).
 * which = 1 / 2, the re_zeros / im_zeros files, three sections: edges in x,
 * y, z on which that component vanishes at the first corner or changes
 * sign.  In z the reference joins the two tests with .and. (io.f90:1304 and
 * 1375), which no finite value satisfies, so that section stays empty; this
 * restatement keeps the test as written.
 *
 * Each point carries its section and the linear index i + nx*(j + ny*k) of
 * the lower corner.  Returns the number of points; the first `cap` are
 * written. */
static int zr_edge(REAL r0, REAL i0, REAL r1, REAL i1) {
  return (r0 == (REAL)0.0 && i0 == (REAL)0.0) || (r0 == (REAL)0.0 && i0 * i1 < (REAL)0.0) ||
         (i0 == (REAL)0.0 && r0 * r1 < (REAL)0.0) || (r0 * r1 < (REAL)0.0 && i0 * i1 < (REAL)0.0);
}

/* zero = -w1*a0/d + w0*a1/d with d = w0 - w1 (io.f90:933-935) */
static REAL zr_interp(REAL w0, REAL w1, REAL a0, REAL a1, REAL d) { return -(w1 * a0 / d) + w0 * a1 / d; }

/* Face with corners c1 = (a0,b0), c2 = (a1,b0), c3 = (a1,b1), c4 = (a0,b1):
 * u changes sign on edges 1-2 and 3-4 (v does not), v on edges 2-3 and
 * 4-1 (u does not).  io.f90:1039-1065 with u = Re, v = Im; 1066-1092 with
 * u = Im, v = Re. */
static int zr_face_pattern(const REAL *u, const REAL *v) {
  return (u[1] * u[2] < (REAL)0.0 && v[1] * v[2] >= (REAL)0.0) && (v[2] * v[3] < (REAL)0.0 && u[2] * u[3] >= (REAL)0.0) &&
         (u[3] * u[4] < (REAL)0.0 && v[3] * v[4] >= (REAL)0.0) && (v[4] * v[1] < (REAL)0.0 && u[4] * u[1] >= (REAL)0.0);
}

static void zr_face_point(const REAL *u, const REAL *v, REAL a0, REAL a1, REAL b0, REAL b1, REAL *ap, REAL *bp) {
  const REAL z1 = zr_interp(u[1], u[2], a0, a1, u[1] - u[2]);
  const REAL z2 = zr_interp(v[2], v[3], b0, b1, v[2] - v[3]);
  const REAL z3 = zr_interp(u[3], u[4], a1, a0, u[3] - u[4]);
  const REAL z4 = zr_interp(v[4], v[1], b1, b0, v[4] - v[1]);
  const REAL m1 = (b0 - b1) / (z1 - z3);
  const REAL m2 = (z2 - z4) / (a1 - a0);
  const REAL a = (z4 - a0 * m2 - b0 + z1 * m1) / (m1 - m2);
  *ap = a;
  *bp = a * m1 + b0 - z1 * m1;
}

long FN(zeros)(const orc_params *p, const REAL *f, int which, orc_zero *out, long cap) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  ctx c;
  ctx_init(&c, p);
  const int nsec = which == 0 ? 6 : 3;
  long n = 0;
#define ZRE(ii, jj, kk) f[2 * (((long)(kk) * ny + (jj)) * nx + (ii))]
#define ZIM(ii, jj, kk) f[2 * (((long)(kk) * ny + (jj)) * nx + (ii)) + 1]
  for (int sec = 0; sec < nsec; sec++)
    for (int k = 1; k <= nz - 2; k++)
      for (int j = 1; j <= ny - 2; j++)
        for (int i = 1; i <= nx - 2; i++) {
          REAL px = c.x[i], py = c.y[j], pz = c.z[k];
          int hit = 0;
          if (which == 0 && sec < 3) {
            const int di = sec == 0, dj = sec == 1, dk = sec == 2;
            const REAL r0 = ZRE(i, j, k), i0 = ZIM(i, j, k);
            const REAL r1 = ZRE(i + di, j + dj, k + dk), i1 = ZIM(i + di, j + dj, k + dk);
            if (zr_edge(r0, i0, r1, i1)) {
              const REAL d = r0 - r1;
              if (d != (REAL)0.0) {
                hit = 1;
                if (sec == 0) px = zr_interp(r0, r1, c.x[i], c.x[i + 1], d);
                if (sec == 1) py = zr_interp(r0, r1, c.y[j], c.y[j + 1], d);
                if (sec == 2) pz = zr_interp(r0, r1, c.z[k], c.z[k + 1], d);
              }
            }
          } else if (which == 0) {
            /* corner offsets of c1..c4 and the face axes (a, b) */
            int o[5][3];
            REAL a0, a1, b0, b1;
            if (sec == 3) {        /* i,j,k --> i+1,j,k --> i+1,j+1,k --> i,j+1,k */
              int t[5][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
              memcpy(o, t, sizeof t);
              a0 = c.x[i]; a1 = c.x[i + 1]; b0 = c.y[j]; b1 = c.y[j + 1];
            } else if (sec == 4) { /* i,j,k --> i+1,j,k --> i+1,j,k+1 --> i,j,k+1 */
              int t[5][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}};
              memcpy(o, t, sizeof t);
              a0 = c.x[i]; a1 = c.x[i + 1]; b0 = c.z[k]; b1 = c.z[k + 1];
            } else {               /* i,j,k --> i,j,k+1 --> i,j+1,k+1 --> i,j+1,k */
              int t[5][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}};
              memcpy(o, t, sizeof t);
              a0 = c.z[k]; a1 = c.z[k + 1]; b0 = c.y[j]; b1 = c.y[j + 1];
            }
            REAL re[5], im[5];
            for (int q = 1; q <= 4; q++) {
              re[q] = ZRE(i + o[q][0], j + o[q][1], k + o[q][2]);
              im[q] = ZIM(i + o[q][0], j + o[q][1], k + o[q][2]);
            }
            REAL ap = 0, bp = 0;
            if (zr_face_pattern(re, im)) {
              zr_face_point(re, im, a0, a1, b0, b1, &ap, &bp);
              hit = 1;
            } else if (zr_face_pattern(im, re)) {
              zr_face_point(im, re, a0, a1, b0, b1, &ap, &bp);
              hit = 1;
            }
            if (hit) {
              if (sec == 3) { px = ap; py = bp; }
              else if (sec == 4) { px = ap; pz = bp; }
              else { pz = ap; py = bp; }
            }
          } else {
            const int di = sec == 0, dj = sec == 1, dk = sec == 2;
            const REAL w0 = which == 1 ? ZRE(i, j, k) : ZIM(i, j, k);
            const REAL w1 = which == 1 ? ZRE(i + di, j + dj, k + dk) : ZIM(i + di, j + dj, k + dk);
            const int cond = sec < 2 ? (w0 == (REAL)0.0 || w0 * w1 < (REAL)0.0)
                                     : (w0 == (REAL)0.0 && w0 * w1 < (REAL)0.0);
            if (cond) {
              const REAL d = w0 - w1;
              if (d != (REAL)0.0) {
                hit = 1;
                if (sec == 0) px = zr_interp(w0, w1, c.x[i], c.x[i + 1], d);
                if (sec == 1) py = zr_interp(w0, w1, c.y[j], c.y[j + 1], d);
                if (sec == 2) pz = zr_interp(w0, w1, c.z[k], c.z[k + 1], d);
              }
            }
          }
          if (hit) {
            if (n < cap) {
              out[n].section = sec;
              out[n].reserved = 0;
              out[n].cell = ((long long)k * ny + j) * nx + i;
              out[n].x = (double)px;
              out[n].y = (double)py;
              out[n].z = (double)pz;
            }
            n++;
          }
        }
#undef ZRE
#undef ZIM
  ctx_free(&c);
  return n;
}

/* ---- imprint_vortex_line (variables.f90:1166-1180) ---------------------
 * psi = abs(psi) * vortex_line(vl): vortex_line (ic.f90:332-353) is
 * exp(i theta), times amp(r) = 1 - exp(-0.7 r**1.15) (ic.f90:314-328) unless
 * vl%imprint_phase, with theta from get_theta (ic.f90:824-889) and r from
 * get_r (ic.f90:735-800) for a line along `dir` displaced by
 * (a1 sin, a2 cos)(2 pi (c - c0)/ll).  With ll ~ 0 and both amplitudes ~ 0
 * the wavelength becomes 1 (safe1, safe2); ll ~ 0 with an amplitude is the
 * reference's emergency_stop (returns -1 here). */
static REAL r_sin(REAL v) { return IS_FLOAT ? (REAL)sinf((float)v) : (REAL)sin((double)v); }
static REAL r_cos(REAL v) { return IS_FLOAT ? (REAL)cosf((float)v) : (REAL)cos((double)v); }
static REAL r_exp(REAL v) { return IS_FLOAT ? (REAL)expf((float)v) : (REAL)exp((double)v); }
static REAL r_atan2(REAL a, REAL b) { return IS_FLOAT ? (REAL)atan2f((float)a, (float)b) : (REAL)atan2((double)a, (double)b); }

int FN(imprint_vortex_line)(const orc_params *p, REAL *f, const orc_line *vl, double scal_d) {
  const REAL eps = IS_FLOAT ? (REAL)1.1920929e-07f : (REAL)2.220446049250313e-16;
  const REAL x0 = (REAL)vl->x0, y0 = (REAL)vl->y0, z0 = (REAL)vl->z0, a1 = (REAL)vl->amp1, a2 = (REAL)vl->amp2;
  const REAL ll = (REAL)vl->ll, sgn = (REAL)vl->sgn, scal = (REAL)scal_d;
  REAL safe1 = (REAL)0.0, safe2 = (REAL)0.0;
  if (r_abs(ll) < eps) {
    if (r_abs(a1) < eps) safe1 = (REAL)1.0; else return -1;
    if (r_abs(a2) < eps) safe2 = (REAL)1.0; else return -1;
  }
  if (vl->dir < 0 || vl->dir > 2) return -1;
  const REAL twopi = (REAL)2.0 * (REAL)3.14159265358979323846;  /* 2.0_pr*pi (parameters.f90:124) */
  const REAL c1 = (REAL)-0.7, c2 = (REAL)1.15;
  ctx c;
  ctx_init(&c, p);
  for (int k = 0; k < p->nz; k++)
    for (int j = 0; j < p->ny; j++)
      for (int i = 0; i < p->nx; i++) {
        const REAL x = c.x[i], y = c.y[j], z = c.z[k];
        /* (A, B, the coordinate along the line and its origin) per dir */
        REAL A, B;
        if (vl->dir == 0) {        /* 'x': atan2(z.., y..) */
          const REAL t = x - x0;
          A = y - y0 - a1 * r_sin(twopi * t / (ll + safe1));
          B = z - z0 - a2 * r_cos(twopi * t / (ll + safe2));
        } else if (vl->dir == 1) { /* 'y': atan2(x.., z..) */
          const REAL t = y - y0;
          A = z - z0 - a1 * r_sin(twopi * t / (ll + safe1));
          B = x - x0 - a2 * r_cos(twopi * t / (ll + safe2));
        } else {                   /* 'z': atan2(y.., x..) */
          const REAL t = z - z0;
          A = x - x0 - a1 * r_sin(twopi * t / (ll + safe1));
          B = y - y0 - a2 * r_cos(twopi * t / (ll + safe2));
        }
        const REAL theta = sgn * r_atan2(scal * B, scal * A);
        REAL vr = r_cos(theta), vi = r_sin(theta);
        if (!vl->imprint_phase) {
          const REAL r = r_sqrt(scal * (A * A) + scal * (B * B));
          const REAL am = (REAL)1.0 - r_exp(c1 * r_pow(r, c2));
          vr = am * vr;
          vi = am * vi;
        }
        const long s = ((long)k * p->ny + j) * p->nx + i;
        const REAL m = cx_absv(ld(f, s));
        st(f, s, cx_make(m * vr, m * vi));
      }
  ctx_free(&c);
  return 0;
}
