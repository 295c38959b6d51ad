/*
 * gpe_oracle.h -- CPU restatement of the reference's GP time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP library in gpe_amd/
 * with it: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product links it or falls back to it.
 *
 * Two builds of the same source: REAL=double (symbols orc64_*) restates the
 * reference built with pr=selected_real_kind(15,307); REAL=float (orc32_*)
 * restates the pr=selected_real_kind(6,37) build (parameters.in:22-23).
 *
 * Layout: a field is nx*ny*nz interleaved complex values (re, im), x fastest
 * (Fortran order of psi(0:nx1, js:je, ks:ke) on one rank, no halo).  Halo
 * semantics of the reference (variables.f90:203-375 with one rank) are
 * replaced by the equivalent index maps.
 */
#ifndef GPE_ORACLE_H
#define GPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int nx, ny, nz;
  int scheme;      /* 1 euler, 2 rk2, 4 rk4, 45 rk45 (solve.f90:34-50) */
  int eqn;         /* eqn_to_solve 1..4 (solve.f90:414-437) */
  int bcs;         /* 1 periodic, 2 reflective (derivs.f90:285-314) */
  int order;       /* 2 or 4 */
  int real_time;   /* logical real_time (solve.f90:197) */
  double xl, yl, zl;  /* -xr, -yr, -zr (ic.f90:43-45) */
  double dx, dy, dz;  /* (xr-xl)/nx1 ... (ic.f90:46-48) */
  double g, mu, omx, omy, omz, urhs, diss_amp;
  double eps, safety, dt_increase, dt_decrease;  /* parameters.f90:90-93 */
  int nthreads;    /* OpenMP threads (results do not depend on it) */
} orc_params;

typedef struct orc_state {
  double dt_re, dt_im;  /* complex dt (parameters.f90:117) */
  double t, im_t;       /* parameters.f90:118-119 */
  double errmax;        /* rk45: errmax of the accepted attempt */
  int n_rejects;        /* rk45: rejected attempts in the last call */
  int p;                /* step counter (gpe.f90:279) */
} orc_state;

/* One vortex-core point (get_zeros / get_extra_zeros / get_re_im_zeros). */
typedef struct orc_zero {
  int section;        /* section of the reference's output file, from 0 */
  int reserved;
  long long cell;     /* i + nx*(j + ny*k) of the lower corner */
  double x, y, z;
} orc_zero;

/* A vortex line (type line_param, parameters.in:33-43); dir 0/1/2 = x/y/z. */
typedef struct orc_line {
  double x0, y0, z0, amp1, amp2, ll, sgn;
  int dir;
  int imprint_phase;
} orc_line;

#define ORC_DECL(P, REAL)                                                     \
  void P##_grid(const orc_params *p, REAL *x, REAL *y, REAL *z);             \
  void P##_laplacian(const orc_params *p, const REAL *f, REAL *lap);        \
  void P##_deriv_x(const orc_params *p, const REAL *f, REAL *fx);           \
  void P##_rhs(const orc_params *p, const REAL *f, REAL *rhs);              \
  int P##_solver(const orc_params *p, orc_state *st, const REAL *in,        \
                 REAL *out);                                                 \
  int P##_run(const orc_params *p, orc_state *st, REAL *psi, int nsteps,     \
              double *norms);                                                \
  double P##_norm(const orc_params *p, const REAL *f);                      \
  double P##_energy(const orc_params *p, const REAL *f);                    \
  double P##_mass(const orc_params *p, const REAL *f);                      \
  void P##_renormalise(const orc_params *p, REAL *f, double nn, double norm); \
  double P##_linelength(const orc_params *p, const REAL *f);                \
  long P##_zeros(const orc_params *p, const REAL *f, int which,             \
                 orc_zero *out, long cap);                                   \
  int P##_imprint_vortex_line(const orc_params *p, REAL *f,                 \
                              const orc_line *vl, double scal);             \
  void P##_set_control_pow(int correctly_rounded);                          \
  double P##_control_pow(double x, double y);                               \
  long P##_control_pow_ambiguous(void);

ORC_DECL(orc64, double)
ORC_DECL(orc32, float)

#ifdef __cplusplus
}
#endif
#endif
