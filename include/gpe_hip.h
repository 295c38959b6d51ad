/*
 * gpe_hip.h -- C ABI of the MI355X (gfx950) GP time-stepping library.
 *
 * The library replaces the reference's time-stepping hot path: the GP
 * right-hand side (solve.f90:382-440) on the 2nd/4th-order centred-FD
 * Laplacian (derivs.f90:146-272, variables.f90:55-71), fused into the
 * Euler / RK2 / RK4 / Cash-Karp RKF45 stage updates (solve.f90:57-378),
 * with the norm (variables.f90:554-576, 730-840), energy
 * (variables.f90:580-604) and adaptive-error (solve.f90:176-190)
 * reductions on the device.
 *
 * Plain C: pointers, sizes and POD structs only.  No call aborts; every
 * entry point returns GPE_OK (0) or a negative GPE_E* code, and
 * gpe_last_error() gives the message of the last failure on the calling
 * thread.  One host thread per context.
 *
 * Field layout (host and device): nx*ny*nz complex values, real and
 * imaginary parts interleaved, x fastest -- the Fortran order of
 * psi(0:nx1, js:je, ks:ke) on one rank.
 */
#ifndef GPE_HIP_H
#define GPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define GPE_OK 0
#define GPE_EINVAL (-1)   /* bad argument or parameter */
#define GPE_ENOMEM (-2)   /* device or host allocation failed */
#define GPE_EHIP (-3)     /* a HIP runtime call failed */
#define GPE_ENONFINITE (-4) /* NaN in the adaptive error norm */
#define GPE_ESTEP (-5)    /* adaptive step control gave up (too many rejections) */
#define GPE_ESTATE (-6)   /* call out of order (e.g. no field uploaded) */
#define GPE_EFULL (-7)    /* batch: the instance does not fit the HBM budget */
#define GPE_EIO (-8)      /* a file could not be read or parsed */

/* scheme (run_params%scheme, solve.f90:34-50) */
#define GPE_EULER 1
#define GPE_RK2 2
#define GPE_RK4 4
#define GPE_RK45 45

/* precision of the device path */
#define GPE_FP64 0  /* complex128 storage and arithmetic: the parity path */
#define GPE_FP32 1  /* complex64 storage and arithmetic */
#define GPE_FP32C 2 /* complex64 storage + hi/lo compensated psi update,
                       FP64 reductions and step control */

/* host buffer element type for uploads and downloads */
#define GPE_HOST_C128 0
#define GPE_HOST_C64 1

/* Run parameters: run_params / eqn_params (parameters.f90:24-56), the
 * grid from get_grid (ic.f90:32-69) and the adaptive-step constants
 * (parameters.f90:90-93). */
typedef struct gpe_params {
  int32_t nx, ny, nz;      /* grid (parameters.in:27-29); each >= 5 */
  int32_t scheme;          /* GPE_EULER | GPE_RK2 | GPE_RK4 | GPE_RK45 */
  int32_t eqn;             /* eqn_to_solve 1..4 */
  int32_t bcs;             /* 1 periodic, 2 reflective */
  int32_t order;           /* 2 or 4 */
  int32_t precision;       /* GPE_FP64 | GPE_FP32 | GPE_FP32C */
  int32_t real_time;       /* 1 real time, 0 imaginary time */
  int32_t reserved;
  double xl, yl, zl;       /* -xr, -yr, -zr */
  double dx, dy, dz;       /* (xr-xl)/(nx-1), ... */
  double g, mu;            /* eqn 4 coefficients */
  double omx, omy, omz;    /* trap frequencies (Vtrap, ic.f90:713-731) */
  double urhs, diss_amp;   /* eqns 1-3 */
  double eps, safety, dt_increase, dt_decrease; /* rk45 control */
} gpe_params;

/* Time-stepping state carried across calls (parameters.f90:117-119,129). */
typedef struct gpe_step_state {
  double dt_re, dt_im;     /* complex dt; rk45 updates it */
  double t, im_t;
  double errmax;           /* out: errmax of the last accepted rk45 attempt */
  int32_t n_rejects;       /* out: rejected rk45 attempts in the last call */
  int32_t p;               /* step counter, advanced by gpe_sim_advance */
} gpe_step_state;

typedef struct gpe_ctx gpe_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char *gpe_last_error(void);

/* Create a context for one simulation: validates params, allocates the
 * device fields the scheme needs and a stream.  Replaces the automatic
 * arrays of rk4/rkck (solve.f90:117-121, 284-290). */
int gpe_ctx_create(const gpe_params *params, gpe_ctx **out);
void gpe_ctx_destroy(gpe_ctx *ctx);

/* Drop-in for solver(var_in, var_out) (solve.f90:25-53), as called by
 * gpe.f90:230.  var_in has `halo` ghost layers in y and z (2 for the
 * Fortran caller: psi%old(0:nx1, js-2:je+2, ks-2:ke+2)); only its interior
 * is read -- boundaries are applied in the kernels.  var_out is
 * (nx, ny, nz).  Element type: complex128 for GPE_FP64, complex64 for
 * GPE_FP32.  st is updated as solver() updates t, im_t and dt.  Synchronous
 * with respect to both host buffers. */
int gpe_solver_step(gpe_ctx *ctx, const void *var_in, int32_t halo, void *var_out,
                    gpe_step_state *st);

/* The same step with the device field authoritative across calls: the
 * drop-in for unchanged gpe.f90 when the host does not read psi between
 * most steps (gpe.f90:167-176 copies psi%new into psi%old and calls
 * solver(); only its periodic I/O, 178-226, reads psi).
 *   GPE_STEP_IN_RESIDENT   var_in's interior equals the field the context
 *                          holds (the result of its previous step): skip
 *                          the upload.  GPE_ESTATE if the context holds no
 *                          field yet.
 *   GPE_STEP_OUT_DEFERRED  do not write var_out: the context keeps the
 *                          result and var_out keeps its old contents until
 *                          gpe_sim_download (or a step without this flag)
 *                          writes it.
 * flags = 0 is gpe_solver_step. */
#define GPE_STEP_IN_RESIDENT 1
#define GPE_STEP_OUT_DEFERRED 2
int gpe_solver_step_ex(gpe_ctx *ctx, const void *var_in, int32_t halo, void *var_out,
                       gpe_step_state *st, int32_t flags);
/* Bytes of field data this context has moved host -> device and device ->
 * host (uploads, downloads, drop-in steps).  No reference counterpart:
 * it shows what the device-authoritative drop-in saves. */
int gpe_transfer_bytes(gpe_ctx *ctx, int64_t *to_device, int64_t *to_host);

/* Device-resident path (no host traffic per step). */
int gpe_sim_upload(gpe_ctx *ctx, const void *psi, int32_t host_type);
int gpe_sim_download(gpe_ctx *ctx, void *psi, int32_t host_type);
int gpe_sim_set_state(gpe_ctx *ctx, const gpe_step_state *st);
int gpe_sim_get_state(gpe_ctx *ctx, gpe_step_state *st);
/* nsteps solver() calls on the resident field (the time loop of
 * gpe.f90:136-280 without its host I/O).  If record_norm is nonzero,
 * get_norm of each new state is kept (gpe.f90:243-245); read it with
 * gpe_sim_norm_history. */
int gpe_sim_advance(gpe_ctx *ctx, int64_t nsteps, int32_t record_norm);
int gpe_sim_norm_history(gpe_ctx *ctx, double *out, int64_t n);
/* Reductions of the resident field, in the reference's functional forms. */
int gpe_sim_norm(gpe_ctx *ctx, double *norm);     /* get_norm */
int gpe_sim_energy(gpe_ctx *ctx, double *energy); /* energy */
int gpe_sim_mass(gpe_ctx *ctx, double *mass);     /* mass */
/* The eqn-4 Hamiltonian that the dynamics conserve,
 * H = K + sum (V |psi|^2 + g/2 |psi|^4 - mu |psi|^2) dV with the kinetic
 * energy K = (1/2) sum Re(-lap psi psi*) dV of the finite-difference
 * Laplacian, formed from energy() (variables.f90:580-604, which weighs
 * 2K + (1/2)|psi|^4) and one pass for the other sums.  No reference
 * counterpart: the reference reports energy() only.  eqn_to_solve must be
 * 4 (GPE_EINVAL otherwise). */
int gpe_sim_hamiltonian(gpe_ctx *ctx, double *hamiltonian);
/* linelength (variables.f90:896-1162): total vortex-line length, the
 * reference's cell-by-cell zero tracing (including its reuse of the previous
 * cell's second point in cells that find a single point). */
int gpe_sim_linelength(gpe_ctx *ctx, double *length);
/* Vortex-core points of the resident field, as the reference's zeros files
 * hold them: which = 0 is zeros<p>.dat (get_zeros + get_extra_zeros,
 * io.f90:886-1236; sections 0..2 = edges in x, y, z on which Re and Im both
 * change sign, 3..5 = xy, xz, yz faces around which their zero lines
 * cross), 1 is re_zeros<p>.dat and 2 is im_zeros<p>.dat (get_re_im_zeros,
 * io.f90:1237-1393; sections 0..2 = edges in x, y, z).  Points come in the
 * file's order (section, then k, j, i of the cell's lower corner), in the
 * precision of the storage form (FP32: single-precision arithmetic).
 * Writes min(count, cap) points to out and the number found to *count;
 * call with cap = 0 to size the buffer. */
typedef struct gpe_zero_point {
  int32_t section;   /* section of the file, from 0 */
  int32_t reserved;
  int64_t cell;      /* i + nx*(j + ny*k) of the lower corner */
  double x, y, z;    /* the point */
} gpe_zero_point;
int gpe_sim_zeros(gpe_ctx *ctx, int32_t which, gpe_zero_point *out, int64_t cap, int64_t *count);
/* A vortex line (type line_param, parameters.in:33-43). */
typedef struct gpe_line_param {
  double x0, y0, z0;      /* position; the coordinate along dir is the origin of the disturbance */
  double amp1, amp2;      /* disturbance amplitudes (amp1*sin, amp2*cos) */
  double ll;              /* disturbance wavelength */
  double sgn;             /* sign of the winding */
  int32_t dir;            /* direction of the line: 0 'x', 1 'y', 2 'z' */
  int32_t imprint_phase;  /* nonzero: phase only, no core amplitude */
} gpe_line_param;
/* imprint_vortex_line (variables.f90:1166-1180), which gpe.f90:272 applies
 * after each imaginary-time step: psi = abs(psi) * vortex_line(vl), with
 * the eqn_params scaling `scal`.  A zero wavelength with a non-zero
 * amplitude is the reference's emergency stop (GPE_EINVAL). */
int gpe_sim_imprint_vortex_line(gpe_ctx *ctx, const gpe_line_param *vl, double scal);
/* Random-phase initial state built on the device, replacing the resident
 * field.  The reference's random_phase (ic.f90:195-274) gives the modes
 * inside a sphere of wavenumbers a fixed amplitude and a random phase and
 * transforms back with 2DECOMP&FFT (ic.f90:910-939), phases drawn from the
 * Fortran generator seeded by the clock.  Here: every mode
 * k = 2 pi m/(n d) (m in fftfreq order) with |k| <= kmax gets amplitude 1
 * and phase 2 pi u, u from SplitMix64(seed) in the order of the modes' flat
 * index; inverse FFT (hipFFT); rescaled so <|psi|^2> = 1; converted to the
 * context's storage form. */
int gpe_sim_random_phase(gpe_ctx *ctx, uint64_t seed, double kmax);
/* Create, ahead of use and from any thread, the cached inverse-FFT plan
 * that gpe_sim_random_phase and gpe_batch_add_random_phase use for this
 * grid (plans are kept for the life of the process; creating one costs far
 * more than running it). */
int gpe_random_phase_prepare(int32_t nx, int32_t ny, int32_t nz);
/* The mode list of that state, computed on the host (no device work):
 * writes min(count, cap) flat indices (i + nx*(j + ny*k)) and unit phasors
 * (re, im pairs) and sets *count. */
int gpe_random_phase_modes(int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz, double kmax,
                           uint64_t seed, int64_t *flat_index, double *phasor, int64_t cap, int64_t *count);
/* FFT-side diagnostics of the resident field.  Each starts from the
 * unitary forward DFT a of psi, as the reference's fft(.., 'forward')
 * leaves it (ic.f90:910-939: 2DECOMP&FFT's forward transform divided by
 * sqrt(nx*ny*nz)), computed here with hipFFT in FP64 whatever the storage
 * form; a(i,j,k) takes wavenumber index i = 0..nx-1 (likewise j, k), with
 * k^2 formed by the reference's wrap (i <= (nx-1)/2+1 ? i : nx-i)^2 + ...
 * They need one complex128 work field beside the context's (allocated on
 * first use).
 *
 * condensed_particles (io.f90:304-395): the record it appends to mass.dat,
 * from M = mass(psi) and n0 = |a(0,0,0)|^2, with the grid extent
 * xr = -xl (likewise y, z) and time = the context's t. */
typedef struct gpe_condensed {
  double time, im_t;
  double density;          /* M / (8 xr yr zr) */
  double n0_per_site;      /* n0 / V, V = nx*ny*nz */
  double mass, n0;         /* M, n0 */
  double temp, temp2;      /* (M-n0)/tot, (M/(8 xr yr zr) - n0/V)/tot */
  double energy_per_site;  /* H / V */
} gpe_condensed;
int gpe_sim_condensed_particles(gpe_ctx *ctx, gpe_condensed *out);
/* spectrum (io.f90:623-716): for the 7 octave shells
 * m-1 <= |log2(|k|/(2 pi))| < m, the number of modes nharm[m-1] and the
 * mean occupation eta[m-1] (0 for an empty shell), and the integral
 * distribution idf[kk-1] = sum of |a|^2 over modes with 0 < |k| <= kk,
 * kk = 1..nx/2 (nidf must be nx/2).  Written by the reference to
 * spectrum<p>.dat, idf<p>.dat and eta_time.dat. */
int gpe_sim_spectrum(gpe_ctx *ctx, int32_t *nharm, double *eta, double *idf, int32_t nidf);
/* filtered_surface (io.f90:524-619): the field whose spectrum is a times
 * max(1 - m k^2/filter_kc^2, 0) (m = fscale*ind, or 1), transformed back
 * with the unitary inverse DFT, into the host array out (nx*ny*nz of
 * host_type); with length != NULL also its vortex-line length (the
 * reference's save_linelength(filtered, 1)).  The resident field is
 * unchanged. */
int gpe_sim_filtered(gpe_ctx *ctx, double filter_kc, double m, void *out, int32_t host_type, double *length);
/* renormalise (variables.f90:715-726): psi = sqrt(nn)*psi/sqrt(norm). */
int gpe_sim_renormalise(gpe_ctx *ctx, double nn, double norm);
/* f(psi) of the resident field into a host buffer (get_rhs, solve.f90:382). */
int gpe_sim_rhs(gpe_ctx *ctx, void *out, int32_t host_type);

/* energy() (variables.f90:580-604, as gpe_sim_energy) of a field the caller
 * holds in device memory, allocating nothing.  No reference counterpart:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).
 *   params  grid, spacing, order, bcs and precision of the field (the
 *           scheme and step-control fields are not read)
 *   psi     device field, nx*ny*nz complex128 (GPE_FP64) or complex64
 *           (GPE_FP32, and the high word of GPE_FP32C)
 *   energy  device pointer to one double, written on `stream`
 *   stream  a hipStream_t, or NULL for the null stream
 * Asynchronous: the result is in *energy when the stream reaches it.  The
 * per-workgroup partial sums live in a static device array of 65,536
 * doubles per GPU (512 KiB, held until the process ends); grids that need
 * more workgroups than that (with the default (8,2) tile and automatic
 * z-chunks: far beyond one GPU's memory) return GPE_EINVAL.  Calls that
 * overlap on one GPU (different streams or threads) share that array and
 * race; calls on one stream are ordered.  The result equals gpe_sim_energy's
 * on a context with the default launch configuration, bit for bit. */
int gpe_energy_noalloc(const gpe_params *params, const void *psi, double *energy, void *stream);

/* Launch configuration of the stage kernels.  Each stage kind (0 .. 14)
 * and each fused stage pair (16 .. 20; see gpe_prof_query for names) runs
 * with one of several compiled tile variants (waves x rows per wave) and a
 * z-chunk length (0 = automatic); a pair with variant -1 runs as its two
 * stages.  No reference counterpart: the reference has no device kernels to
 * configure. */
int gpe_tune_set(gpe_ctx *ctx, int32_t kind, int32_t variant, int32_t zchunk);
/* Fuse consecutive RK4 / Cash-Karp stages pairwise into one sweep each
 * (RK4: 7 field passes per step instead of 13; Cash-Karp: 14 per attempt
 * instead of 27).  Results are bitwise those of the stage-by-stage sweeps.
 * Default: allowed for GPE_FP32 / GPE_FP32C, with every pair unfused until
 * gpe_autotune or gpe_tune_set (variant >= 0) selects its fused sweep.
 * GPE_FP64 has no fused kernels: enable = 1 on a GPE_FP64 context returns
 * GPE_EINVAL. */
int gpe_set_fusion(gpe_ctx *ctx, int32_t enable);
int gpe_tune_get(gpe_ctx *ctx, int32_t kind, int32_t *variant, int32_t *zchunk, int32_t *waves,
                 int32_t *rows_per_wave);
/* Stored launch configurations.  The autotuner's picks are kept in a JSON
 * table keyed by problem, "NXxNYxNZ/oORDER/PREC/SCHEME/bBCS" with "/fused"
 * appended where the context allows fused pairs (PREC fp64 | fp32 | fp32c,
 * SCHEME euler | rk2 | rk4 | rk45), each entry mapping stage-kind names
 * (gpe_prof_query) to [variant, zchunk].  gpe_tune_key writes the context's
 * key into buf (cap bytes, NUL-terminated).  gpe_tune_load applies the entry
 * for the context's problem from the table at path; path NULL or "" means
 * $GPE_TUNING_FILE, else tuning.json beside this library.  A missing table
 * or key leaves the configuration unchanged and sets *applied = 0 (1 when
 * an entry was applied).  Kind names this library does not know are
 * skipped.  A malformed table, or an entry gpe_tune_set would reject,
 * returns GPE_EIO with the configuration unchanged.  Speed only: results
 * do not depend on the launch configuration.  No reference counterpart. */
int gpe_tune_key(gpe_ctx *ctx, char *buf, int64_t cap);
int gpe_tune_load(gpe_ctx *ctx, const char *path, int32_t *applied);
/* Time every variant and several z-chunks of each stage kind (and, with
 * fusion on, of each stage pair) of the context's scheme on the resident
 * field (reps launches each) and keep the fastest; a pair runs fused only
 * where that beats its two stages.  The state is unchanged. */
int gpe_autotune(gpe_ctx *ctx, int32_t reps);

/* Kernel timing with HIP events recorded on the context's stream around
 * every stage launch.  enable=1 resets the counters. */
int gpe_prof_enable(gpe_ctx *ctx, int32_t enable);
/* For kernel kind `kind` (0..GPE_PROF_KINDS-1): launches, total ms and
 * algorithmic HBM bytes per launch; the name is written to name[64]. */
#define GPE_PROF_KINDS 24
int gpe_prof_query(gpe_ctx *ctx, int32_t kind, int64_t *launches, double *total_ms,
                   double *bytes_per_launch, char *name);

/* Measured HBM ceiling of this GPU: read, write and copy bandwidth (GB/s,
 * counting bytes read + written) of plain float4 kernels over arrays of
 * `bytes`.  No reference counterpart: the reference point the roofline
 * fractions are read against, beside the 8 TB/s datasheet peak. */
int gpe_probe_bandwidth(int64_t bytes, int32_t reps, double *read_gbs, double *write_gbs,
                        double *copy_gbs);

/* ---- Batches: many independent simulations on one GPU -----------------
 *
 * No reference counterpart: the reference runs one compile-time grid per
 * job (parameters.in:23-29).  Each instance of a batch is a solver() loop
 * of its own (solve.f90:25-53 as called by gpe.f90:230), with its own grid,
 * boundary conditions, order, trap, dt and step count, in FP64 or FP32C.
 * Per stage kind, all instances of one (precision, order, scheme) run in
 * one persistent launch over a table of (instance, tile) work items; after
 * the last stage a control kernel advances every instance (t, dt, the rk45
 * accept/reject of rkqs, solve.f90:180-229) on the device.  One iteration
 * (one rk4 step, or one rk45 attempt) is captured in a hipGraph and
 * replayed.  An instance's results do not depend on what else is in the
 * batch.  Instance device memory: psi, psi' and the scheme's work fields
 * (RK4: 3, RKF45: 6) plus psi's low word (x2) in FP32C; admission is
 * against an HBM budget.  Calls that change the set of instances
 * synchronise the batch's stream. */
typedef struct gpe_batch gpe_batch;

/* hbm_budget: device bytes the instances may use (0 = 90% of free memory). */
int gpe_batch_create(int64_t hbm_budget, gpe_batch **out);
void gpe_batch_destroy(gpe_batch *b);
/* Device bytes one instance with these parameters occupies. */
int gpe_batch_footprint(const gpe_params *params, int64_t *bytes);
/* Add an instance that will take nsteps solver() steps from state st and
 * field psi (nx*ny*nz, complex128 or complex64 per host_type).  Returns
 * GPE_EFULL, adding nothing, when it does not fit the budget. */
int gpe_batch_add(gpe_batch *b, const gpe_params *params, const void *psi, int32_t host_type,
                  const gpe_step_state *st, int64_t nsteps, int32_t *id);
/* gpe_batch_add with the random-phase state of gpe_sim_random_phase built
 * on the device in the instance's memory (no host field, no upload). */
int gpe_batch_add_random_phase(gpe_batch *b, const gpe_params *params, uint64_t seed, double kmax,
                               const gpe_step_state *st, int64_t nsteps, int32_t *id);
int gpe_batch_remove(gpe_batch *b, int32_t id);
/* Enqueue `iterations` iterations (asynchronous).  0 only prepares: it
 * rebuilds the work tables and captures the iteration graph if the set of
 * instances changed. */
int gpe_batch_run(gpe_batch *b, int32_t iterations);
int gpe_batch_sync(gpe_batch *b);
/* Instances that stopped since the last poll: ids and status (1 finished,
 * -1 NaN in the rk45 error norm, -2 more than 1000 rejections); n <= cap. */
int gpe_batch_poll(gpe_batch *b, int32_t *ids, int32_t *status, int32_t cap, int32_t *n);
/* State of one instance: st as solver() leaves it (p = steps done),
 * status (0 running), get_norm of its current field, and the rk45
 * attempts rejected in the step in progress (0 at a step boundary, where
 * st, the field and the step count are a complete restart point). */
int gpe_batch_query(gpe_batch *b, int32_t id, gpe_step_state *st, int64_t *steps_done, int32_t *status,
                    double *norm, int32_t *pending_rejects);
int gpe_batch_download(gpe_batch *b, int32_t id, void *psi, int32_t host_type);
int gpe_batch_stats(gpe_batch *b, int32_t *instances, int32_t *running, int64_t *bytes_in_use,
                    int64_t *budget);
/* 1: replay the captured iteration graph (default); 0: launch kernels. */
int gpe_batch_use_graph(gpe_batch *b, int32_t enable);
/* Tile variant and z-chunk (0 = whole column) of one stage kind (0 .. 14)
 * or fused stage pair (16 .. 20, variant -1 = run as its two stages) for
 * the instances of (precision, order, scheme). */
int gpe_batch_tune_set(gpe_batch *b, int32_t precision, int32_t order, int32_t scheme, int32_t kind,
                       int32_t variant, int32_t zchunk);
int gpe_batch_tune_get(gpe_batch *b, int32_t precision, int32_t order, int32_t scheme, int32_t kind,
                       int32_t *variant, int32_t *zchunk);
/* Time every variant and z-chunk of every stage kind and stage pair over
 * the current instances and keep the fastest (a pair runs fused only where
 * that beats its two stages).  Commits no step. */
int gpe_batch_autotune(gpe_batch *b, int32_t reps);

/* ---- Serving: continuous batching behind a request queue ---------------
 *
 * No reference counterpart (the reference runs one simulation per job).
 * Requests -- a solver() loop of nsteps steps from a state and a field --
 * are queued in submission order.  Each gpe_serve_step admits queued
 * requests first come first served while their device footprint fits the
 * HBM budget and the resident grid points stay within max_sites (the first
 * request that does not fit waits, and the ones behind it), runs
 * `iterations` batch iterations (every resident request advances one step,
 * rk45 one attempt; requests admitted now join the running iteration), and
 * retires finished requests: field and state are kept in host memory until
 * gpe_serve_result, and their device memory returns to the pool.  A
 * request's result equals the same request run alone in a batch. */
typedef struct gpe_serve gpe_serve;
/* hbm_budget: device bytes for resident requests (0 = 90% of free memory);
 * max_sites: cap on resident grid points (0 = none). */
int gpe_serve_create(int64_t hbm_budget, int64_t max_sites, gpe_serve **out);
void gpe_serve_destroy(gpe_serve *s);
/* Queue a request.  psi: the initial field (nx*ny*nz of host_type), copied;
 * or NULL for the random-phase state of gpe_sim_random_phase(seed, kmax),
 * built on the device at admission.  GPE_EFULL if the request can never fit
 * the budget.  *ticket names the request. */
int gpe_serve_submit(gpe_serve *s, const gpe_params *params, const void *psi, int32_t host_type, uint64_t seed,
                     double kmax, const gpe_step_state *st, int64_t nsteps, int64_t *ticket);
/* One scheduling round (admit, run `iterations`, retire); the counts of
 * resident and queued requests after it.  Synchronous. */
int gpe_serve_step(gpe_serve *s, int32_t iterations, int32_t *running, int32_t *queued);
/* Requests finished since the last poll: tickets and status (1 done, -1 NaN
 * in the rk45 error norm, -2 more than 1000 rejections, -3 could not be
 * admitted, or its result could not be read back: the gpe_serve_step that
 * met it returned the error); n <= cap. */
int gpe_serve_poll(gpe_serve *s, int64_t *tickets, int32_t *status, int32_t cap, int32_t *n);
/* Field (host_type) and state of a finished request; releases it
 * (GPE_ESTATE if it has not finished, was collected or has status -3). */
int gpe_serve_result(gpe_serve *s, int64_t ticket, void *psi, int32_t host_type, gpe_step_state *st);

#ifdef __cplusplus
}
#endif
#endif /* GPE_HIP_H */
