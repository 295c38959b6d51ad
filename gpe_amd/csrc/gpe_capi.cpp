// gpe_capi.cpp -- C ABI of the GP time-stepping library (include/gpe_hip.h).
//
// Host side of the path: owns the device fields, sequences the fused stage
// kernels for each scheme (solve.f90:57-378), runs the rk45 step-size
// control of rkqs (solve.f90:158-244) and replaces the reference's
// solver(var_in, var_out) call (solve.f90:25-53).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/gpe_hip.h"
#include "host_common.h"
#include "gpe_ic.h"
#include "gpe_spectral.h"
#include "launch.h"

using namespace gpe;



namespace {

#define HIPCHK(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail(GPE_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

}  // namespace


namespace {

// Profiling / tuning ids: stage kinds 0..K_COUNT-1, fused pairs KP0 + pair.
constexpr int KP0 = 16;

const char *kind_name(int k) {
  static const char *names[] = {"rhs",   "euler", "rk2_1", "rk2_2", "rk4_1", "rk4_2", "rk4_3", "rk4_4",
                                "ck_1",  "ck_2",  "ck_3",  "ck_4",  "ck_5",  "ck_6",  "energy"};
  static const char *pairs[] = {"rk4_12", "rk4_34", "ck_12", "ck_34", "ck_56"};
  if (k >= 0 && k < K_COUNT) return names[k];
  if (k >= KP0 && k < KP0 + N_PAIR_KINDS) return pairs[k - KP0];
  return "?";
}

// Fields read + written per site by each kind in the fused-sweep traffic
// model (units of one stored complex): RK4 13 passes per step, Cash-Karp 27
// per attempt (the error scale is counted as register-resident, so the
// scale field FP64 and plain FP32 keep is traffic above the model).  The compensated
// low word (read + write in the final stage) is added by the caller.
int kind_passes(int k) {
  switch (k) {
    case K_RHS: return 2;
    case K_EULER: return 2;
    case K_RK2_1: return 2;
    case K_RK2_2: return 3;
    case K_RK4_1: return 2;
    case K_RK4_2: return 4;
    case K_RK4_3: return 3;
    case K_RK4_4: return 4;
    case K_CK_1: return 2;
    case K_CK_2: return 3;
    case K_CK_3: return 4;
    case K_CK_4: return 5;
    case K_CK_5: return 6;
    case K_CK_6: return 7;
    case K_ENERGY: return 1;
    // fused pairs: the two stages' shares of the model
    case KP0 + P_RK4_12: return 2 + 4;
    case KP0 + P_RK4_34: return 3 + 4;
    case KP0 + P_CK_12: return 2 + 3;
    case KP0 + P_CK_34: return 4 + 5;
    case KP0 + P_CK_56: return 6 + 7;
    default: return 0;
  }
}

bool kind_final(int k) {
  return k == K_EULER || k == K_RK2_2 || k == K_RK4_4 || k == K_CK_6 || k == KP0 + P_RK4_34 || k == KP0 + P_CK_56;
}

}  // namespace


struct gpe_ctx {
  gpe_params P{};
  int R = 1;
  int prec = GPE_FP64;
  size_t n = 0;
  size_t esz = 16;  // bytes per stored complex
  bool comp = false;
  hipStream_t stream = nullptr;
  // fields
  void *psi = nullptr, *psi2 = nullptr, *lo = nullptr, *lo2 = nullptr;
  void *k[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void *acc = nullptr, *scale = nullptr;
  // reductions
  double *part = nullptr;  // 3 * part_cap
  int part_cap = 0;
  double *dscal = nullptr;   // device scalars: norm, emax, energy
  double *hscal = nullptr;   // pinned mirror
  double *hist = nullptr;
  int64_t hist_cap = 0, hist_len = 0;
  bool uploaded = false;
  int64_t bytes_h2d = 0, bytes_d2h = 0;  // host field traffic
  gpe_step_state st{};
  StageArgs base{};
  // launch configuration per stage kind: tile variant (index into kTiles)
  // and z-chunk length (0 = automatic)
  int vid[K_COUNT];
  int zc[K_COUNT];
  // fused stage pairs: on/off, tile variant and z-chunk per pair kind
  bool fuse = false;
  bool pfuse[N_PAIR_KINDS];  // per pair: fused sweep (autotuner's pick) or the two stages
  int pvid[N_PAIR_KINDS];
  int pzc[N_PAIR_KINDS];
  // profiling
  bool prof = false;
  struct Ev {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Ev> pending;
  std::vector<hipEvent_t> pool;
  int64_t prof_n[GPE_PROF_KINDS] = {};
  // line-length point lists (grown on demand)
  void *ll_buf = nullptr;
  int ll_cap = 0;
  void *zr_buf = nullptr;  // zeros pass: cap points + a counter
  int zr_cap = 0;
  // FFT-side diagnostics: a complex128 field and small scratch (on demand)
  void *spec = nullptr, *spec_aux = nullptr;
  double prof_ms[GPE_PROF_KINDS] = {};
};

namespace {

int grid_for_v(gpe_ctx *c, int vid, int zc, StageArgs &A) { return stage_grid(c->P.nx, c->P.ny, c->P.nz, vid, zc, A); }

int grid_for(gpe_ctx *c, int kind, StageArgs &A) { return grid_for_v(c, c->vid[kind], c->zc[kind], A); }

// Grid of a fused pair: tiles step by the (64-2R) x (TY-2R) interiors; the
// z-chunks are long (each chunk primes 2R extra planes of stage A).
int grid_for_pair_v(gpe_ctx *c, int vid, int zc, StageArgs &B) {
  const PairTile t = kPairTiles[vid];
  const int TXO = 64 - 2 * c->R, TYO = t.nw * t.p - 2 * c->R;
  B.ntx = (c->P.nx + TXO - 1) / TXO;
  B.nty = (c->P.ny + TYO - 1) / TYO;
  const int tiles2d = B.ntx * B.nty;
  if (zc <= 0) {
    int nzc = (1024 + tiles2d - 1) / tiles2d;
    zc = (c->P.nz + nzc - 1) / nzc;
    if (zc < 64 && tiles2d * ((c->P.nz + 63) / 64) >= 512) zc = 64;
  }
  if (zc < 8) zc = 8;
  if (zc > c->P.nz) zc = c->P.nz;
  B.zc = zc;
  B.nzc = (c->P.nz + zc - 1) / zc;
  B.nblocks = tiles2d * B.nzc;
  return B.nblocks;
}

hipEvent_t take_event(gpe_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void prof_drain(gpe_ctx *c) {
  for (auto &ev : c->pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(ev.b);
    if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess && ev.kind >= 0 && ev.kind < GPE_PROF_KINDS) {
      c->prof_n[ev.kind] += 1;
      c->prof_ms[ev.kind] += ms;
    }
    c->pool.push_back(ev.a);
    c->pool.push_back(ev.b);
  }
  c->pending.clear();
}

int run_kind(gpe_ctx *c, int kind, StageArgs &A) {
  Range range(kind_name(kind));
  grid_for(c, kind, A);
  if (A.nblocks > c->part_cap) return fail(GPE_EINVAL, "internal: %d workgroups exceed partial buffer", A.nblocks);
  hipEvent_t ea = nullptr, eb = nullptr;
  if (c->prof) {
    ea = take_event(c);
    eb = take_event(c);
    if (ea) (void)hipEventRecord(ea, c->stream);
  }
  hipError_t e = launch_stage((int)c->esz, c->vid[kind], kind, c->R, c->comp, A, c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "stage %s launch: %s", kind_name(kind), hipGetErrorString(e));
  if (c->prof && ea && eb) {
    (void)hipEventRecord(eb, c->stream);
    c->pending.push_back({kind, ea, eb});
    if (c->pending.size() > 4096) prof_drain(c);
  }
  return GPE_OK;
}

int run_pair(gpe_ctx *c, int pair, const StageArgs &A, StageArgs &B) {
  Range range(kind_name(KP0 + pair));
  grid_for_pair_v(c, c->pvid[pair], c->pzc[pair], B);
  if (B.nblocks > c->part_cap) return fail(GPE_EINVAL, "internal: %d workgroups exceed partial buffer", B.nblocks);
  hipEvent_t ea = nullptr, eb = nullptr;
  if (c->prof) {
    ea = take_event(c);
    eb = take_event(c);
    if (ea) (void)hipEventRecord(ea, c->stream);
  }
  hipError_t e = launch_pair((int)c->esz, c->pvid[pair], pair, c->R, c->comp, A, B, c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "pair %s launch: %s", kind_name(KP0 + pair), hipGetErrorString(e));
  if (c->prof && ea && eb) {
    (void)hipEventRecord(eb, c->stream);
    c->pending.push_back({KP0 + pair, ea, eb});
    if (c->pending.size() > 4096) prof_drain(c);
  }
  return GPE_OK;
}

// Stages kA, kB = kA+1 of pair `pair`: one fused sweep, or one sweep each.
int run_two(gpe_ctx *c, int pair, StageArgs &A, StageArgs &B) {
  static const int first[N_PAIR_KINDS] = {K_RK4_1, K_RK4_3, K_CK_1, K_CK_3, K_CK_5};
  if (c->fuse && c->pfuse[pair]) return run_pair(c, pair, A, B);
  int rc = run_kind(c, first[pair], A);
  return rc ? rc : run_kind(c, first[pair] + 1, B);
}

StageArgs args_for(gpe_ctx *c) {
  StageArgs A = c->base;
  A.dtr = c->st.dt_re;
  A.dti = c->st.dt_im;
  A.hr = 0.5 * c->st.dt_re;
  A.hi = 0.5 * c->st.dt_im;
  A.inv_absdt = 1.0 / std::hypot(c->st.dt_re, c->st.dt_im);
  A.part_norm = c->part;
  A.part_emax = c->part + c->part_cap;
  A.part_energy = c->part + 2 * c->part_cap;
  return A;
}

int reduce_after(gpe_ctx *c, const StageArgs &A, bool emax, bool energy, double *hist_slot) {
  const double dv = c->P.dx * c->P.dy * c->P.dz;
  hipError_t e = launch_reduce(A.part_norm, emax ? A.part_emax : nullptr, energy ? A.part_energy : nullptr,
                               A.nblocks, dv, c->P.eps, dv, c->dscal, hist_slot, c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "reduce launch: %s", hipGetErrorString(e));
  return GPE_OK;
}

void swap_state(gpe_ctx *c) {
  std::swap(c->psi, c->psi2);
  if (c->comp) std::swap(c->lo, c->lo2);
}

// One solver() call on the resident field.
int step_once(gpe_ctx *c, double *hist_slot) {
  int rc;
  const int sch = c->P.scheme;
  if (sch == GPE_EULER) {
    StageArgs A = args_for(c);
    A.psi = c->psi;
    A.lo = c->lo;
    A.psi_out = c->psi2;
    A.lo_out = c->lo2;
    if ((rc = run_kind(c, K_EULER, A))) return rc;
    if ((rc = reduce_after(c, A, false, false, hist_slot))) return rc;
    swap_state(c);
  } else if (sch == GPE_RK2) {
    StageArgs A = args_for(c);
    A.psi = c->psi;
    A.kout = c->k[0];
    if ((rc = run_kind(c, K_RK2_1, A))) return rc;
    A = args_for(c);
    A.psi = c->psi;
    A.lo = c->lo;
    A.kin[0] = c->k[0];
    A.psi_out = c->psi2;
    A.lo_out = c->lo2;
    if ((rc = run_kind(c, K_RK2_2, A))) return rc;
    if ((rc = reduce_after(c, A, false, false, hist_slot))) return rc;
    swap_state(c);
  } else if (sch == GPE_RK4 && c->fuse) {
    // stages (1,2) and (3,4) pairwise: fused, k1 and k3 never leave the sweep
    StageArgs A1 = args_for(c), A2 = args_for(c), A3 = args_for(c), A4 = args_for(c);
    A1.psi = c->psi;
    A1.kout = c->k[0];
    A2.psi = c->psi;
    A2.kin[0] = c->k[0];
    A2.kout = c->k[1];
    A2.acc_out = c->acc;
    if ((rc = run_two(c, P_RK4_12, A1, A2))) return rc;
    A3.psi = c->psi;
    A3.kin[0] = c->k[1];
    A3.kout = c->k[0];
    A4.psi = c->psi;
    A4.lo = c->lo;
    A4.kin[0] = c->k[0];
    A4.acc_in = c->acc;
    A4.psi_out = c->psi2;
    A4.lo_out = c->lo2;
    if ((rc = run_two(c, P_RK4_34, A3, A4))) return rc;
    if ((rc = reduce_after(c, A4, false, false, hist_slot))) return rc;
    swap_state(c);
  } else if (sch == GPE_RK4) {
    // k[0]: k1 then k3 (k1 is folded into acc by stage 2), k[1]: k2
    StageArgs A = args_for(c);
    A.psi = c->psi;
    A.kout = c->k[0];
    if ((rc = run_kind(c, K_RK4_1, A))) return rc;
    A = args_for(c);
    A.psi = c->psi;
    A.kin[0] = c->k[0];
    A.kout = c->k[1];
    A.acc_out = c->acc;
    if ((rc = run_kind(c, K_RK4_2, A))) return rc;
    A = args_for(c);
    A.psi = c->psi;
    A.kin[0] = c->k[1];
    A.kout = c->k[0];
    if ((rc = run_kind(c, K_RK4_3, A))) return rc;
    A = args_for(c);
    A.psi = c->psi;
    A.lo = c->lo;
    A.kin[0] = c->k[0];
    A.acc_in = c->acc;
    A.psi_out = c->psi2;
    A.lo_out = c->lo2;
    if ((rc = run_kind(c, K_RK4_4, A))) return rc;
    if ((rc = reduce_after(c, A, false, false, hist_slot))) return rc;
    swap_state(c);
  } else {
    return GPE_EINVAL;  // rk45 is driven by step_rk45
  }
  const double dtr = c->st.dt_re, dti = c->st.dt_im;
  if (c->prec == GPE_FP32) {
    c->st.t = (double)((float)c->st.t + (float)dtr);
    c->st.im_t = (double)((float)c->st.im_t + std::fabs((float)dti));
  } else {
    c->st.t = c->st.t + dtr;
    c->st.im_t = c->st.im_t + std::fabs(dti);
  }
  return GPE_OK;
}

// rkqs (solve.f90:158-244) with the arithmetic of the storage precision for
// the plain FP32 path (the reference built with pr = single) and FP64
// otherwise.
template <typename RT>
struct Ctl {
  static RT pw(RT a, RT b) { return (RT)std::pow(a, b); }
};

template <typename RT>
int rkqs(gpe_ctx *c, double *hist_slot) {
  // eps divides the error on the device (the final stage reports maxval/eps)
  const RT safety = (RT)c->P.safety;
  const RT dt_inc = (RT)c->P.dt_increase, dt_dec = (RT)c->P.dt_decrease;
  const RT errcon = Ctl<RT>::pw((RT)5.0 / safety, (RT)1.0 / dt_inc);
  RT dtr = (RT)c->st.dt_re, dti = (RT)c->st.dt_im;
  const double sdtr = c->st.dt_re, sdti = c->st.dt_im;
  RT errmax = 0;
  int rejects = 0;
  int rc;
  for (int attempt = 0;; ++attempt) {
    Range range("rk45 attempt");
    c->st.dt_re = (double)dtr;
    c->st.dt_im = (double)dti;
    // descriptor of Cash-Karp stage s (kk_s = dt f(psi + sum_m b_sm kk_m))
    auto ck_args = [&](int s) {
      StageArgs A = args_for(c);
      A.sdtr = sdtr;
      A.sdti = sdti;
      A.psi = c->psi;
      A.lo = c->lo;
      for (int m = 0; m < s - 1; ++m) A.kin[m] = c->k[m];
      if (s < 6) {
        A.kout = c->k[s - 1];
        if (s == 1) {
          A.write_scale = attempt == 0;
          A.scale_out = c->scale;
        }
      } else {
        A.scale_in = c->scale;
        A.psi_out = c->psi2;
        A.lo_out = c->lo2;
      }
      return A;
    };
    if (c->fuse) {
      for (int s = 1; s <= 5; s += 2) {
        StageArgs A = ck_args(s), B = ck_args(s + 1);
        if ((rc = run_two(c, P_CK_12 + s / 2, A, B))) return rc;
        if (s == 5 && (rc = reduce_after(c, B, true, false, hist_slot))) return rc;
      }
    } else {
      for (int s = 1; s <= 6; ++s) {
        StageArgs A = ck_args(s);
        if ((rc = run_kind(c, K_CK_1 + s - 1, A))) return rc;
        if (s == 6 && (rc = reduce_after(c, A, true, false, hist_slot))) return rc;
      }
    }
    HIPCHK(hipMemcpyAsync(c->hscal, c->dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double em = c->hscal[1];  // maxval(|err/psi_scal|)/eps
    errmax = (RT)em;
    if (!(errmax == errmax)) return fail(GPE_ENONFINITE, "rk45: NaN in the error norm (t=%g)", c->st.t);
    if (errmax <= (RT)1.0) break;
    if (++rejects > 1000) return fail(GPE_ESTEP, "rk45: more than 1000 rejected attempts (t=%g)", c->st.t);
    const RT pwv = Ctl<RT>::pw(errmax, dt_dec);
    RT ar = safety * dtr, ai = safety * dti;
    ar = ar * pwv;
    ai = ai * pwv;
    const RT a = (RT)std::hypot(ar, ai), b = (RT)0.1 * (RT)std::hypot(dtr, dti);
    const RT mx = a > b ? a : b;
    if (!c->P.real_time) {
      dtr = (RT)0.0;
      dti = -mx;
    } else {
      dtr = mx;
      dti = (RT)0.0;
    }
  }
  RT nr, ni;
  if (errmax > errcon) {
    const RT pwv = Ctl<RT>::pw(errmax, dt_inc);
    RT ar = safety * dtr, ai = safety * dti;
    nr = ar * pwv;
    ni = ai * pwv;
  } else {
    nr = (RT)5.0 * dtr;
    ni = (RT)5.0 * dti;
  }
  c->st.t = (double)((RT)c->st.t + dtr);
  c->st.dt_re = (double)nr;
  c->st.dt_im = (double)ni;
  c->st.im_t = (double)((RT)c->st.im_t + (RT)std::fabs(ni));
  c->st.errmax = (double)errmax;
  c->st.n_rejects = rejects;
  swap_state(c);
  return GPE_OK;
}

int solver_call(gpe_ctx *c, double *hist_slot) {
  Range range("solver");
  if (c->P.scheme == GPE_RK45) return c->prec == GPE_FP32 ? rkqs<float>(c, hist_slot) : rkqs<double>(c, hist_slot);
  return step_once(c, hist_slot);
}

void free_all(gpe_ctx *c) {
  void *ptrs[] = {c->psi, c->psi2, c->lo, c->lo2, c->k[0], c->k[1], c->k[2], c->k[3], c->k[4], c->k[5],
                  c->acc, c->scale, c->part, c->dscal, c->hist, c->ll_buf, c->zr_buf, c->spec, c->spec_aux};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  for (auto &ev : c->pending) {
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int check_params(const gpe_params *p) {
  if (!p) return fail(GPE_EINVAL, "params is NULL");
  if (p->nx < 5 || p->ny < 5 || p->nz < 5) return fail(GPE_EINVAL, "grid must be at least 5 points per axis");
  if (p->scheme != GPE_EULER && p->scheme != GPE_RK2 && p->scheme != GPE_RK4 && p->scheme != GPE_RK45)
    return fail(GPE_EINVAL, "ERROR: Unrecognised time stepping scheme.");
  if (p->eqn < 1 || p->eqn > 4) return fail(GPE_EINVAL, "eqn_to_solve must be 1..4");
  if (p->bcs != 1 && p->bcs != 2) return fail(GPE_EINVAL, "bcs must be 1 (periodic) or 2 (reflective)");
  if (p->order != 2 && p->order != 4) return fail(GPE_EINVAL, "order must be 2 or 4");
  if (p->precision < GPE_FP64 || p->precision > GPE_FP32C) return fail(GPE_EINVAL, "unknown precision");
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0)) return fail(GPE_EINVAL, "grid spacing must be positive");
  return GPE_OK;
}

int alloc_field(gpe_ctx *c, void **p) {
  hipError_t e = hipMalloc(p, c->n * c->esz);
  if (e != hipSuccess) return fail(GPE_ENOMEM, "hipMalloc of %zu bytes: %s", c->n * c->esz, hipGetErrorString(e));
  return GPE_OK;
}

int to_device(gpe_ctx *c, const void *src, int host_type, size_t src_pitch_planes, size_t row_off) {
  const int rc = field_upload(c->prec, (size_t)c->P.nx * c->P.ny, c->P.nz, src, host_type, src_pitch_planes,
                              row_off, c->psi, c->lo);
  if (rc == GPE_OK) c->bytes_h2d += (int64_t)(c->n * (host_type == GPE_HOST_C128 ? 16 : 8));
  return rc;
}

int from_device(gpe_ctx *c, void *dst, int host_type) {
  GPE_HIPCHK(hipStreamSynchronize(c->stream));
  const int rc = field_download(c->prec, (size_t)c->P.nx * c->P.ny, c->P.nz, c->psi, c->lo, dst, host_type);
  if (rc == GPE_OK) c->bytes_d2h += (int64_t)(c->n * (host_type == GPE_HOST_C128 ? 16 : 8));
  return rc;
}

}  // namespace

extern "C" {

const char *gpe_last_error(void) { return g_err.c_str(); }

int gpe_ctx_create(const gpe_params *params, gpe_ctx **out) {
  if (!out) return fail(GPE_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  gpe_ctx *c = new (std::nothrow) gpe_ctx();
  if (!c) return fail(GPE_ENOMEM, "host allocation failed");
  c->P = *params;
  c->R = params->order / 2;
  c->prec = params->precision;
  c->comp = params->precision == GPE_FP32C;
  c->esz = params->precision == GPE_FP64 ? 16 : 8;
  c->n = (size_t)params->nx * params->ny * params->nz;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    rc = fail(GPE_EHIP, "hipStreamCreate failed");
    delete c;
    return rc;
  }
  int nk = 0;
  switch (params->scheme) {
    case GPE_EULER: nk = 0; break;
    case GPE_RK2: nk = 1; break;
    case GPE_RK4: nk = 2; break;
    case GPE_RK45: nk = 5; break;
  }
  if ((rc = alloc_field(c, &c->psi)) || (rc = alloc_field(c, &c->psi2))) goto bad;
  if (c->comp && ((rc = alloc_field(c, &c->lo)) || (rc = alloc_field(c, &c->lo2)))) goto bad;
  for (int m = 0; m < nk; ++m)
    if ((rc = alloc_field(c, &c->k[m]))) goto bad;
  if (params->scheme == GPE_RK4 && (rc = alloc_field(c, &c->acc))) goto bad;
  // the compensated form forms the rk45 error scale in the last stage
  if (params->scheme == GPE_RK45 && !c->comp && (rc = alloc_field(c, &c->scale))) goto bad;
  {
    // default launch configuration; partial buffers sized for the largest
    // grid any variant and z-chunk can produce
    for (int kk = 0; kk < K_COUNT; ++kk) {
      c->vid[kk] = variant_allowed(c->prec == GPE_FP64, kk, kDefaultVid) ? kDefaultVid : kVidF64Ck;
      c->zc[kk] = 0;
    }
    // every pair runs as its two stages until the autotuner (or
    // gpe_tune_set with a variant >= 0) finds the fused sweep faster: the
    // one-stage kernels hold up at every grid size, a fused pair's register
    // pressure does not (RKF45 at 1024^3: 148 ms per step fused against
    // 53 ms unfused).  The FP32 forms allow fusion; FP64 (the parity path)
    // never fuses.
    for (int q = 0; q < N_PAIR_KINDS; ++q) {
      c->pvid[q] = 0;
      c->pzc[q] = 0;
      c->pfuse[q] = false;
    }
    c->fuse = c->prec != GPE_FP64;
    int cap = 0;
    for (int v = 0; v < N_TILE_VARIANTS; ++v) {
      StageArgs tmp{};
      int nb = grid_for_v(c, v, 8, tmp);
      if (nb > cap) cap = nb;
    }
    for (int v = 0; v < N_PAIR_VARIANTS; ++v) {
      StageArgs tmp{};
      int nb = grid_for_pair_v(c, v, 8, tmp);
      if (nb > cap) cap = nb;
    }
    if (cap < PW_BLOCKS) cap = PW_BLOCKS;
    c->part_cap = cap;
    if (hipMalloc((void **)&c->part, 3 * sizeof(double) * (size_t)cap) != hipSuccess) {
      rc = fail(GPE_ENOMEM, "partials");
      goto bad;
    }
    if (hipMalloc((void **)&c->dscal, 4 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **)&c->hscal, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      rc = fail(GPE_ENOMEM, "scalars");
      goto bad;
    }
  }
  fill_base(params, c->prec, c->base);
  *out = c;
  return GPE_OK;
bad:
  free_all(c);
  delete c;
  return rc;
}

void gpe_ctx_destroy(gpe_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  delete ctx;
}

int gpe_sim_upload(gpe_ctx *c, const void *psi, int32_t host_type) {
  if (!c || !psi) return fail(GPE_EINVAL, "NULL argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t plane = (size_t)c->P.nx * c->P.ny;
  int rc = to_device(c, psi, host_type, plane, 0);
  if (rc) return rc;
  if (c->comp && host_type != GPE_HOST_C128 && host_type != GPE_HOST_C64) return fail(GPE_EINVAL, "host type");
  c->uploaded = true;
  c->hist_len = 0;
  return GPE_OK;
}

int gpe_sim_download(gpe_ctx *c, void *psi, int32_t host_type) {
  if (!c || !psi) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  return from_device(c, psi, host_type);
}

int gpe_sim_set_state(gpe_ctx *c, const gpe_step_state *st) {
  if (!c || !st) return fail(GPE_EINVAL, "NULL argument");
  c->st = *st;
  return GPE_OK;
}

int gpe_sim_get_state(gpe_ctx *c, gpe_step_state *st) {
  if (!c || !st) return fail(GPE_EINVAL, "NULL argument");
  *st = c->st;
  return GPE_OK;
}

int gpe_sim_advance(gpe_ctx *c, int64_t nsteps, int32_t record_norm) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (nsteps < 0) return fail(GPE_EINVAL, "nsteps < 0");
  if (record_norm) {
    if (c->hist_cap < nsteps) {
      if (c->hist) (void)hipFree(c->hist);
      c->hist = nullptr;
      HIPCHK(hipMalloc((void **)&c->hist, sizeof(double) * (size_t)(nsteps > 0 ? nsteps : 1)));
      c->hist_cap = nsteps;
    }
    c->hist_len = 0;
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    double *slot = record_norm ? c->hist + s : nullptr;
    int rc = solver_call(c, slot);
    if (rc) return rc;
    c->st.p += 1;
    if (record_norm) c->hist_len = s + 1;
  }
  if (c->prof) prof_drain(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return GPE_OK;
}

int gpe_sim_norm_history(gpe_ctx *c, double *out, int64_t n) {
  if (!c || !out) return fail(GPE_EINVAL, "NULL argument");
  if (n > c->hist_len) return fail(GPE_EINVAL, "only %lld norms recorded", (long long)c->hist_len);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (n > 0) HIPCHK(hipMemcpy(out, c->hist, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  return GPE_OK;
}

static int norm_like(gpe_ctx *c, int weighted, double *out) {
  if (!c || !out) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  hipError_t e;
  const StageArgs &A = c->base;
  if (c->prec == GPE_FP64)
    e = launch_norm_f64(c->psi, A.nx, A.ny, A.nz, weighted, A.w1, A.w2, A.w3, c->part, c->stream);
  else
    e = launch_norm_f32(c->psi, c->comp ? c->lo : nullptr, A.nx, A.ny, A.nz, weighted, A.w1, A.w2, A.w3, c->part,
                        c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "norm launch: %s", hipGetErrorString(e));
  const double dv = c->P.dx * c->P.dy * c->P.dz;
  e = launch_reduce(c->part, nullptr, nullptr, PW_BLOCKS, dv, c->P.eps, dv, c->dscal, nullptr, c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "reduce launch: %s", hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(c->hscal, c->dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = c->hscal[0];
  return GPE_OK;
}

int gpe_sim_norm(gpe_ctx *c, double *norm) { return norm_like(c, 1, norm); }
int gpe_sim_mass(gpe_ctx *c, double *mass) { return norm_like(c, 0, mass); }

int gpe_sim_energy(gpe_ctx *c, double *energy) {
  if (!c || !energy) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  StageArgs A = args_for(c);
  A.psi = c->psi;
  int rc = run_kind(c, K_ENERGY, A);
  if (rc) return rc;
  if ((rc = reduce_after(c, A, false, true, nullptr))) return rc;
  HIPCHK(hipMemcpyAsync(c->hscal, c->dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *energy = c->hscal[2];
  return GPE_OK;
}

}  // extern "C"

namespace {
// Segment of a one-point cell to the latest earlier second point, with the
// reference's face/edge weights (variables.f90:1140-1153), in T.
template <typename T>
T ll_segment_host(T x1, T y1, T z1, T x2, T y2, T z2) {
  auto in01 = [](T v) { return v == (T)0 || v == (T)1; };
  T a = (T)1;
  const bool ex = x1 == x2 && in01(x1), ey = y1 == y2 && in01(y1), ez = z1 == z2 && in01(z1);
  if (ex || ey) {
    a = a * (T)0.5;
    if ((ex && ey) || (ey && ez)) a = a * (T)0.5;
  }
  return a * std::sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
}

template <typename T>
double ll_singles(std::vector<LinePoint> &w, std::vector<LinePoint> &s) {
  auto by_cell = [](const LinePoint &a, const LinePoint &b) { return a.cell < b.cell; };
  std::sort(w.begin(), w.end(), by_cell);
  std::sort(s.begin(), s.end(), by_cell);
  double sum = 0.0;
  T x2 = 0, y2 = 0, z2 = 0;  // before any cell stored a second point
  size_t j = 0;
  for (const LinePoint &p : s) {
    while (j < w.size() && w[j].cell < p.cell) {
      x2 = (T)w[j].x;
      y2 = (T)w[j].y;
      z2 = (T)w[j].z;
      ++j;
    }
    sum += (double)ll_segment_host<T>((T)p.x, (T)p.y, (T)p.z, x2, y2, z2);
  }
  return sum;
}
}  // namespace

extern "C" {

int gpe_sim_hamiltonian(gpe_ctx *c, double *h) {
  if (!c || !h) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (c->P.eqn != 4) return fail(GPE_EINVAL, "the Hamiltonian is that of eqn_to_solve = 4");
  if (c->part_cap < PW_BLOCKS) return fail(GPE_EINVAL, "internal: partial buffer below %d", PW_BLOCKS);
  double E = 0.0;
  if (int rc = gpe_sim_energy(c, &E)) return rc;
  double *p2 = c->part, *p4 = c->part + c->part_cap, *pv = c->part + 2 * (size_t)c->part_cap;
  // sums of |psi|^2, |psi|^4 and V |psi|^2 over the field (hi + lo in the
  // compensated form)
  auto moments = [&](const void *lo, double &s2, double &s4, double &sv) -> int {
    hipError_t e = launch_moments(c->prec == GPE_FP64 ? 0 : 1, c->psi, lo, c->base, p2, p4, pv, c->stream);
    if (e != hipSuccess) return fail(GPE_EHIP, "moments launch: %s", hipGetErrorString(e));
    std::vector<double> hp(3 * (size_t)c->part_cap);
    HIPCHK(hipMemcpyAsync(hp.data(), c->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    s2 = s4 = sv = 0.0;
    for (int b = 0; b < PW_BLOCKS; ++b) {
      s2 += hp[b];
      s4 += hp[c->part_cap + b];
      sv += hp[2 * (size_t)c->part_cap + b];
    }
    return GPE_OK;
  };
  double s2 = 0.0, s4 = 0.0, sv = 0.0;
  if (int rc = moments(c->comp ? c->lo : nullptr, s2, s4, sv)) return rc;
  // energy() = (sum Re(-lap psi psi*) + 1/2 |psi|^4) dV: twice the kinetic
  // energy plus half the quartic sum.  The compensated form's energy() reads
  // the high word only, so the quartic sum taken out of it is the high
  // word's too (the potential terms below keep hi + lo).
  double s4e = s4;
  if (c->comp) {
    double a = 0.0, b = 0.0;
    if (int rc = moments(nullptr, a, s4e, b)) return rc;
  }
  const double dv = c->P.dx * c->P.dy * c->P.dz;
  const double kin = 0.5 * (E - 0.5 * s4e * dv);
  *h = kin + (0.5 * c->P.g * s4 + sv - c->P.mu * s2) * dv;
  return GPE_OK;
}

}  // extern "C"

namespace {
int linelength_of(gpe_ctx *c, int prec, const void *psi, const void *lo, double *length);
}  // namespace

extern "C" {

int gpe_sim_linelength(gpe_ctx *c, double *length) {
  if (!c || !length) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  return linelength_of(c, c->prec, c->psi, c->comp ? c->lo : nullptr, length);
}

}  // extern "C"

namespace {
// linelength (variables.f90:896-1162) of a device field of the context's
// grid in storage form prec (lo: the compensated low word or NULL).
int linelength_of(gpe_ctx *c, int prec, const void *psi, const void *lo, double *length) {
  const long long ncell = (long long)(c->P.nx - 2) * (c->P.ny - 2) * (c->P.nz - 2);
  const int nb = (int)std::min<long long>((ncell + 255) / 256, std::min(c->part_cap, 2048));
  int cnt[2] = {0, 0};
  for (;;) {
    if (!c->ll_buf || c->ll_cap < 1) {
      const int cap = c->ll_cap > 0 ? c->ll_cap : (1 << 16);
      if (c->ll_buf) (void)hipFree(c->ll_buf);
      c->ll_buf = nullptr;
      if (hipMalloc(&c->ll_buf, 2 * sizeof(LinePoint) * (size_t)cap + 256) != hipSuccess)
        return fail(GPE_ENOMEM, "line-length lists of %d points", cap);
      c->ll_cap = cap;
    }
    LinePoint *w = (LinePoint *)c->ll_buf, *sg = w + c->ll_cap;
    int *counts = (int *)(sg + c->ll_cap);
    HIPCHK(hipMemsetAsync(counts, 0, 2 * sizeof(int), c->stream));
    hipError_t e = launch_linelength(prec, psi, lo, c->P.nx, c->P.ny, c->P.nz, c->part, nb, w, sg, counts, c->ll_cap,
                                     c->stream);
    if (e != hipSuccess) return fail(GPE_EHIP, "linelength launch: %s", hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(cnt, counts, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int need = std::max(cnt[0], cnt[1]);
    if (need <= c->ll_cap) break;
    // lists overflowed: grow them and run the pass again
    (void)hipFree(c->ll_buf);
    c->ll_buf = nullptr;
    c->ll_cap = need + need / 4;
  }
  std::vector<double> part(nb);
  std::vector<LinePoint> w(cnt[0]), sg(cnt[1]);
  LinePoint *dw = (LinePoint *)c->ll_buf;
  HIPCHK(hipMemcpy(part.data(), c->part, sizeof(double) * nb, hipMemcpyDeviceToHost));
  if (cnt[0]) HIPCHK(hipMemcpy(w.data(), dw, sizeof(LinePoint) * cnt[0], hipMemcpyDeviceToHost));
  if (cnt[1]) HIPCHK(hipMemcpy(sg.data(), dw + c->ll_cap, sizeof(LinePoint) * cnt[1], hipMemcpyDeviceToHost));
  double sum = 0.0;
  for (double v : part) sum += v;
  sum += prec == GPE_FP32 ? ll_singles<float>(w, sg) : ll_singles<double>(w, sg);
  *length = prec == GPE_FP32 ? (double)((float)sum * (float)c->P.dx) : sum * c->P.dx;
  return GPE_OK;
}

// The unitary forward DFT of the resident field into c->spec.
int spectral_of_field(gpe_ctx *c) {
  if (!c->spec && hipMalloc(&c->spec, c->n * 16) != hipSuccess) {
    c->spec = nullptr;
    return fail(GPE_ENOMEM, "spectral work field of %zu bytes", c->n * 16);
  }
  if (!c->spec_aux &&
      hipMalloc(&c->spec_aux, sizeof(double) * spectral_scratch_doubles(c->P.nx, c->P.ny, c->P.nz)) != hipSuccess) {
    c->spec_aux = nullptr;
    return fail(GPE_ENOMEM, "spectral scratch");
  }
  return spectral_forward(c->P.nx, c->P.ny, c->P.nz, c->prec, c->psi, c->comp ? c->lo : nullptr,
                          static_cast<double2 *>(c->spec), c->stream);
}

}  // namespace

extern "C" {

int gpe_sim_condensed_particles(gpe_ctx *c, gpe_condensed *out) {
  if (!c || !out) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (c->P.nx < 2) return fail(GPE_EINVAL, "condensed_particles needs nx >= 2");
  double M = 0.0;
  if (int rc = gpe_sim_mass(c, &M)) return rc;
  if (int rc = spectral_of_field(c)) return rc;
  double2 a0;
  HIPCHK(hipMemcpyAsync(&a0, c->spec, sizeof a0, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double m0 = std::hypot(a0.x, a0.y);
  const double n0 = m0 * m0;
  const double V = (double)c->n;
  const double rho0 = n0 / V;
  double tot = 0.0;
  if (int rc = spectral_condensed_sum(c->P.nx, c->P.ny, c->P.nz, c->P.dx, c->P.dy, c->P.dz, rho0,
                                      static_cast<double *>(c->spec_aux), &tot, c->stream))
    return rc;
  // io.f90:371-381, with xr = -xl (the grid runs from -xr to xr)
  const double box = 8.0 * (-c->P.xl) * (-c->P.yl) * (-c->P.zl);
  const double temp = (M - n0) / tot;
  const double E0 = (1.0 / (2.0 * V)) * (M * M + (M - n0) * (M - n0));
  const double H = E0 + temp * (V - 1.0);
  out->time = c->st.t;
  out->im_t = c->st.im_t;
  out->density = M / box;
  out->n0_per_site = n0 / V;
  out->mass = M;
  out->n0 = n0;
  out->temp = temp;
  out->temp2 = ((M / box) - (n0 / V)) / tot;
  out->energy_per_site = H / V;
  return GPE_OK;
}

int gpe_sim_spectrum(gpe_ctx *c, int32_t *nharm, double *eta, double *idf, int32_t nidf) {
  if (!c || !nharm || !eta || (nidf > 0 && !idf)) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  const int nb = c->P.nx / 2;
  if (nidf != nb) return fail(GPE_EINVAL, "idf holds nx/2 = %d values, not %d", nb, nidf);
  if (int rc = spectral_of_field(c)) return rc;
  long long nh[7];
  double es[7];
  std::vector<double> bins(nb);
  double *aux = static_cast<double *>(c->spec_aux);
  if (int rc = spectral_spectrum(c->P.nx, c->P.ny, c->P.nz, static_cast<const double2 *>(c->spec), nb, aux,
                                 reinterpret_cast<long long *>(aux + (size_t)c->P.nz * (7 + nb)), nh, es, bins.data(),
                                 c->stream))
    return rc;
  // io.f90:685-689: the shell mean, left at 0 where the shell holds nothing
  for (int m = 0; m < 7; ++m) {
    nharm[m] = (int32_t)nh[m];
    eta[m] = es[m] == 0.0 ? 0.0 : es[m] / (double)nh[m];
  }
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    acc += bins[b];
    idf[b] = acc;
  }
  return GPE_OK;
}

int gpe_sim_filtered(gpe_ctx *c, double filter_kc, double m, void *out, int32_t host_type, double *length) {
  if (!c || !out) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (host_type != GPE_HOST_C128 && host_type != GPE_HOST_C64) return fail(GPE_EINVAL, "host type");
  if (!(filter_kc != 0.0)) return fail(GPE_EINVAL, "filter_kc must be non-zero");
  if (int rc = spectral_of_field(c)) return rc;
  double2 *a = static_cast<double2 *>(c->spec);
  if (int rc = spectral_filter(c->P.nx, c->P.ny, c->P.nz, a, m, filter_kc, c->stream)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (host_type == GPE_HOST_C128) {
    if (int rc = field_download(GPE_FP64, (size_t)c->P.nx * c->P.ny, c->P.nz, a, nullptr, out, host_type)) return rc;
  } else {
    // the filtered field is FP64: round it to complex64 on the host
    std::vector<double> tmp(2 * c->n);
    if (int rc = field_download(GPE_FP64, (size_t)c->P.nx * c->P.ny, c->P.nz, a, nullptr, tmp.data(), GPE_HOST_C128))
      return rc;
    float *o = static_cast<float *>(out);
    for (size_t q = 0; q < 2 * c->n; ++q) o[q] = (float)tmp[q];
  }
  if (length) return linelength_of(c, GPE_FP64, a, nullptr, length);
  return GPE_OK;
}

int gpe_sim_imprint_vortex_line(gpe_ctx *c, const gpe_line_param *vl, double scal) {
  if (!c || !vl) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (vl->dir < 0 || vl->dir > 2) return fail(GPE_EINVAL, "line direction must be 0 (x), 1 (y) or 2 (z)");
  LineArgs L{};
  // epsilon(ll) of the reference's real kind (ic.f90:840-853)
  const double eps = c->prec == GPE_FP32 ? (double)FLT_EPSILON : DBL_EPSILON;
  if (std::fabs(vl->ll) < eps) {
    if (std::fabs(vl->amp1) >= eps || std::fabs(vl->amp2) >= eps)
      return fail(GPE_EINVAL, "vortex line disturbance wavelength is zero, but amplitude is non-zero.");
    L.safe1 = L.safe2 = 1.0;
  }
  L.xl = c->P.xl;
  L.yl = c->P.yl;
  L.zl = c->P.zl;
  L.dx = c->P.dx;
  L.dy = c->P.dy;
  L.dz = c->P.dz;
  L.x0 = vl->x0;
  L.y0 = vl->y0;
  L.z0 = vl->z0;
  L.a1 = vl->amp1;
  L.a2 = vl->amp2;
  L.ll = vl->ll;
  L.sgn = vl->sgn;
  L.scal = scal;
  L.nx = c->P.nx;
  L.ny = c->P.ny;
  L.nz = c->P.nz;
  L.dir = vl->dir;
  L.phase_only = vl->imprint_phase != 0;
  hipError_t e = launch_imprint(c->prec, c->psi, c->comp ? c->lo : nullptr, L, c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "imprint launch: %s", hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GPE_OK;
}

int gpe_sim_random_phase(gpe_ctx *c, uint64_t seed, double kmax) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = random_phase_device(c->P.nx, c->P.ny, c->P.nz, c->P.dx, c->P.dy, c->P.dz, kmax, seed, c->prec, c->psi,
                               c->comp ? c->lo : nullptr, c->stream);
  if (rc) return rc;
  c->uploaded = true;
  c->hist_len = 0;
  return GPE_OK;
}

int gpe_sim_zeros(gpe_ctx *c, int32_t which, gpe_zero_point *out, int64_t cap, int64_t *count) {
  if (!c || !count) return fail(GPE_EINVAL, "NULL argument");
  if (which < 0 || which > 2) return fail(GPE_EINVAL, "which must be 0 (zeros), 1 (re_zeros) or 2 (im_zeros)");
  if (cap < 0 || (cap > 0 && !out)) return fail(GPE_EINVAL, "bad output buffer");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  static_assert(sizeof(gpe_zero_point) == sizeof(ZeroPoint), "gpe_zero_point layout");
  const long long ncell = (long long)(c->P.nx - 2) * (c->P.ny - 2) * (c->P.nz - 2);
  const int nb = (int)std::max<long long>(1, std::min<long long>((ncell + 255) / 256, 4096));
  const double xyzl[3] = {c->P.xl, c->P.yl, c->P.zl}, dxyz[3] = {c->P.dx, c->P.dy, c->P.dz};
  int cnt = 0;
  for (;;) {
    if (!c->zr_buf) {
      const int zc = c->zr_cap > 0 ? c->zr_cap : (1 << 16);
      if (hipMalloc(&c->zr_buf, sizeof(ZeroPoint) * (size_t)zc + 256) != hipSuccess)
        return fail(GPE_ENOMEM, "zero list of %d points", zc);
      c->zr_cap = zc;
    }
    ZeroPoint *pts = (ZeroPoint *)c->zr_buf;
    int *counter = (int *)(pts + c->zr_cap);
    HIPCHK(hipMemsetAsync(counter, 0, sizeof(int), c->stream));
    hipError_t e = launch_zeros(c->prec, c->psi, c->comp ? c->lo : nullptr, c->P.nx, c->P.ny, c->P.nz, xyzl, dxyz,
                                which, pts, counter, c->zr_cap, nb, c->stream);
    if (e != hipSuccess) return fail(GPE_EHIP, "zeros launch: %s", hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(&cnt, counter, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (cnt <= c->zr_cap) break;
    // the list overflowed: grow it and run the pass again
    if (cnt > (1 << 28)) return fail(GPE_ENOMEM, "%d zero points", cnt);
    (void)hipFree(c->zr_buf);
    c->zr_buf = nullptr;
    c->zr_cap = cnt + cnt / 4;
  }
  std::vector<ZeroPoint> h(cnt);
  if (cnt) HIPCHK(hipMemcpy(h.data(), c->zr_buf, sizeof(ZeroPoint) * (size_t)cnt, hipMemcpyDeviceToHost));
  // the reference's file order: section, then k, j, i
  std::sort(h.begin(), h.end(), [](const ZeroPoint &a, const ZeroPoint &b) {
    return a.section != b.section ? a.section < b.section : a.cell < b.cell;
  });
  const int64_t nout = std::min<int64_t>(cap, cnt);
  if (nout) std::memcpy(out, h.data(), sizeof(ZeroPoint) * (size_t)nout);
  *count = cnt;
  return GPE_OK;
}

int gpe_sim_renormalise(gpe_ctx *c, double nn, double norm) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  hipError_t e;
  if (c->prec == GPE_FP64)
    e = launch_renorm_f64(c->psi, c->n, std::sqrt(nn), make_divc(std::sqrt(norm)), c->stream);
  else
    e = launch_renorm_f32(c->psi, c->comp ? c->lo : nullptr, c->n, std::sqrt(nn), std::sqrt(norm), c->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "renormalise launch: %s", hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GPE_OK;
}

int gpe_sim_rhs(gpe_ctx *c, void *out, int32_t host_type) {
  if (!c || !out) return fail(GPE_EINVAL, "NULL argument");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  void *tmp = c->psi2;  // scratch: the next-state buffer is free between steps
  StageArgs A = args_for(c);
  A.psi = c->psi;
  A.kout = tmp;
  int rc = run_kind(c, K_RHS, A);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->prec == GPE_FP64) {
    if (host_type != GPE_HOST_C128) return fail(GPE_EINVAL, "FP64 context returns complex128 host data");
    HIPCHK(hipMemcpy(out, tmp, c->n * 16, hipMemcpyDeviceToHost));
  } else {
    if (host_type == GPE_HOST_C64) {
      HIPCHK(hipMemcpy(out, tmp, c->n * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(2 * c->n);
      HIPCHK(hipMemcpy(f.data(), tmp, c->n * 8, hipMemcpyDeviceToHost));
      double *d = (double *)out;
      for (size_t q = 0; q < 2 * c->n; ++q) d[q] = f[q];
    }
  }
  return GPE_OK;
}

int gpe_solver_step(gpe_ctx *c, const void *var_in, int32_t halo, void *var_out, gpe_step_state *st) {
  return gpe_solver_step_ex(c, var_in, halo, var_out, st, 0);
}

int gpe_transfer_bytes(gpe_ctx *c, int64_t *to_device, int64_t *to_host) {
  if (!c || !to_device || !to_host) return fail(GPE_EINVAL, "NULL argument");
  *to_device = c->bytes_h2d;
  *to_host = c->bytes_d2h;
  return GPE_OK;
}

int gpe_solver_step_ex(gpe_ctx *c, const void *var_in, int32_t halo, void *var_out, gpe_step_state *st,
                       int32_t flags) {
  if (!c || !var_in || !var_out || !st) return fail(GPE_EINVAL, "NULL argument");
  if (halo < 0) return fail(GPE_EINVAL, "halo < 0");
  if (flags & ~(GPE_STEP_IN_RESIDENT | GPE_STEP_OUT_DEFERRED)) return fail(GPE_EINVAL, "unknown step flags");
  if (c->prec == GPE_FP32C) return fail(GPE_EINVAL, "the compensated form is device-resident only");
  const int ht = c->prec == GPE_FP64 ? GPE_HOST_C128 : GPE_HOST_C64;
  int rc;
  if (flags & GPE_STEP_IN_RESIDENT) {
    if (!c->uploaded) return fail(GPE_ESTATE, "GPE_STEP_IN_RESIDENT before any field was uploaded");
  } else {
    const size_t nx = c->P.nx, nyh = (size_t)c->P.ny + 2 * halo;
    const size_t pitch = nx * nyh;                                // elements per host z-plane
    const size_t off = (size_t)halo * nx + (size_t)halo * pitch;  // (0, js, ks)
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = to_device(c, var_in, ht, pitch, off);
    if (rc) return rc;
    c->uploaded = true;
  }
  c->st = *st;
  rc = solver_call(c, nullptr);
  if (rc) return rc;
  if (flags & GPE_STEP_OUT_DEFERRED)
    HIPCHK(hipStreamSynchronize(c->stream));
  else if ((rc = from_device(c, var_out, ht)))
    return rc;
  gpe_step_state o = c->st;
  o.p = st->p;
  *st = o;
  if (c->prof) prof_drain(c);
  return GPE_OK;
}


int gpe_set_fusion(gpe_ctx *c, int32_t enable) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (enable && c->prec == GPE_FP64) return fail(GPE_EINVAL, "FP64 runs its stages one by one (no fused pairs)");
  c->fuse = enable != 0;
  return GPE_OK;
}

int gpe_tune_set(gpe_ctx *c, int32_t kind, int32_t variant, int32_t zchunk) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (kind >= KP0 && kind < KP0 + N_PAIR_KINDS) {
    const int nv = c->prec == GPE_FP64 ? N_PAIR_VARIANTS_F64 : N_PAIR_VARIANTS;
    if (variant < -1 || variant >= nv) return fail(GPE_EINVAL, "pair variant out of range");
    if (zchunk < 0) return fail(GPE_EINVAL, "zchunk < 0");
    // variant -1: run this pair as its two stages
    c->pfuse[kind - KP0] = variant >= 0;
    if (variant >= 0) c->pvid[kind - KP0] = variant;
    c->pzc[kind - KP0] = zchunk;
    return GPE_OK;
  }
  if (kind < 0 || kind >= K_COUNT) return fail(GPE_EINVAL, "kind out of range");
  if (variant < 0 || variant >= N_TILE_VARIANTS) return fail(GPE_EINVAL, "variant out of range");
  if (!variant_allowed(c->prec == GPE_FP64, kind, variant))
    return fail(GPE_EINVAL, "FP64 Cash-Karp stages do not run tile variant %d", variant);
  if (zchunk < 0) return fail(GPE_EINVAL, "zchunk < 0");
  c->vid[kind] = variant;
  c->zc[kind] = zchunk;
  return GPE_OK;
}

int gpe_tune_get(gpe_ctx *c, int32_t kind, int32_t *variant, int32_t *zchunk, int32_t *nw, int32_t *rows) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (kind >= KP0 && kind < KP0 + N_PAIR_KINDS) {
    const int q = kind - KP0;
    // variant -1: this pair runs as its two stages
    if (variant) *variant = (c->fuse && c->pfuse[q]) ? c->pvid[q] : -1;
    if (zchunk) *zchunk = c->pzc[q];
    if (nw) *nw = kPairTiles[c->pvid[q]].nw;
    if (rows) *rows = kPairTiles[c->pvid[q]].p;
    return GPE_OK;
  }
  if (kind < 0 || kind >= K_COUNT) return fail(GPE_EINVAL, "kind out of range");
  if (variant) *variant = c->vid[kind];
  if (zchunk) *zchunk = c->zc[kind];
  if (nw) *nw = kTiles[c->vid[kind]].nw;
  if (rows) *rows = kTiles[c->vid[kind]].p;
  return GPE_OK;
}

int gpe_tune_key(gpe_ctx *c, char *buf, int64_t cap) {
  if (!c || !buf) return fail(GPE_EINVAL, "NULL argument");
  static const char *precs[] = {"fp64", "fp32", "fp32c"};
  const char *scheme = c->P.scheme == GPE_EULER ? "euler"
                       : c->P.scheme == GPE_RK2 ? "rk2"
                       : c->P.scheme == GPE_RK4 ? "rk4"
                                                : "rk45";
  const int n = std::snprintf(buf, cap > 0 ? static_cast<size_t>(cap) : 0, "%dx%dx%d/o%d/%s/%s/b%d%s", c->P.nx,
                              c->P.ny, c->P.nz, c->P.order, precs[c->prec], scheme, c->P.bcs,
                              c->fuse ? "/fused" : "");
  if (n < 0 || n >= cap) return fail(GPE_EINVAL, "key buffer too small (%d bytes needed)", n + 1);
  return GPE_OK;
}

int gpe_tune_load(gpe_ctx *c, const char *path, int32_t *applied) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (applied) *applied = 0;
  std::string file;
  if (path && *path) {
    file = path;
  } else if (const char *env = std::getenv("GPE_TUNING_FILE"); env && *env) {
    file = env;
  } else {
    // tuning.json in the directory of this library
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void *>(&gpe_tune_load), &info) || !info.dli_fname)
      return fail(GPE_EIO, "cannot locate the library to find tuning.json");
    file = info.dli_fname;
    const size_t slash = file.rfind('/');
    file = (slash == std::string::npos ? std::string(".") : file.substr(0, slash)) + "/tuning.json";
  }
  char key[128];
  if (int rc = gpe_tune_key(c, key, sizeof key)) return rc;
  std::vector<TuneItem> items;
  bool found = false;
  if (int rc = read_tune_table(file.c_str(), key, items, found)) return rc;
  if (!found) return GPE_OK;
  // names this library does not know (a table from another version) are skipped
  std::vector<int> kinds(items.size(), -1);
  for (size_t i = 0; i < items.size(); ++i)
    for (int k = 0; k < KP0 + N_PAIR_KINDS; ++k)
      if ((k < K_COUNT || k >= KP0) && items[i].name == kind_name(k)) kinds[i] = k;
  int vid[K_COUNT], zc[K_COUNT], pvid[N_PAIR_KINDS], pzc[N_PAIR_KINDS];
  bool pfuse[N_PAIR_KINDS];
  std::memcpy(vid, c->vid, sizeof vid);
  std::memcpy(zc, c->zc, sizeof zc);
  std::memcpy(pvid, c->pvid, sizeof pvid);
  std::memcpy(pzc, c->pzc, sizeof pzc);
  std::memcpy(pfuse, c->pfuse, sizeof pfuse);
  for (size_t i = 0; i < items.size(); ++i) {
    if (kinds[i] < 0) continue;
    if (gpe_tune_set(c, kinds[i], items[i].variant, items[i].zchunk) != GPE_OK) {
      // leave the configuration as it was
      std::memcpy(c->vid, vid, sizeof vid);
      std::memcpy(c->zc, zc, sizeof zc);
      std::memcpy(c->pvid, pvid, sizeof pvid);
      std::memcpy(c->pzc, pzc, sizeof pzc);
      std::memcpy(c->pfuse, pfuse, sizeof pfuse);
      const std::string why = g_err;
      return fail(GPE_EIO, "%s: entry %s, kind %s: %s", file.c_str(), key, items[i].name.c_str(), why.c_str());
    }
  }
  if (applied) *applied = 1;
  return GPE_OK;
}

// Time every tile variant and a few z-chunk lengths of each stage kind of
// the context's scheme on the resident field, keep the fastest.  Outputs go
// to the work fields (k, acc, next state), which every step rewrites before
// reading, so the state is unchanged.
int gpe_autotune(gpe_ctx *c, int32_t reps) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (!c->uploaded) return fail(GPE_ESTATE, "no field uploaded");
  if (reps < 1) reps = 1;
  std::vector<int> kinds;
  switch (c->P.scheme) {
    case GPE_EULER: kinds = {K_EULER}; break;
    case GPE_RK2: kinds = {K_RK2_1, K_RK2_2}; break;
    case GPE_RK4: kinds = {K_RK4_1, K_RK4_2, K_RK4_3, K_RK4_4}; break;
    case GPE_RK45: kinds = {K_CK_1, K_CK_2, K_CK_3, K_CK_4, K_CK_5, K_CK_6}; break;
  }
  std::vector<int> pairs;
  if (c->fuse && c->P.scheme == GPE_RK4) pairs = {P_RK4_12, P_RK4_34};
  if (c->fuse && c->P.scheme == GPE_RK45) pairs = {P_CK_12, P_CK_34, P_CK_56};
  // best time per stage kind and per fused pair; a pair runs fused only if
  // that beats its two stages one by one
  float kbest[K_COUNT], pbest[N_PAIR_KINDS];
  for (float &v : kbest) v = 1e30f;
  for (float &v : pbest) v = 1e30f;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const bool prof = c->prof;
  c->prof = false;
  int rc = GPE_OK;
  // fused pairs: outputs go to work fields every step rewrites before reading
  for (int pair : pairs) {
    StageArgs A = args_for(c), B = args_for(c);
    for (StageArgs *X : {&A, &B}) {
      X->sdtr = c->st.dt_re;
      X->sdti = c->st.dt_im;
      X->psi = c->psi;
      X->lo = c->lo;
      X->write_scale = 0;
      for (int m = 0; m < 5; ++m) X->kin[m] = c->k[m] ? c->k[m] : c->psi;
      X->psi_out = c->psi2;
      X->lo_out = c->lo2;
      X->acc_in = c->acc;
      X->scale_in = c->scale;
      X->scale_out = c->scale;
    }
    if (pair == P_RK4_12) {
      B.kout = c->k[1];
      B.acc_out = c->acc;
    } else if (pair == P_RK4_34) {
      A.kin[0] = c->k[1];
    } else if (pair == P_CK_12) {
      A.kout = c->k[2];
      B.kout = c->k[3];
    } else if (pair == P_CK_34) {
      A.kout = c->k[4];
      B.kout = c->k[4];
    }
    // z-chunks to try; the whole z extent gives one chunk per tile, which
    // on a large grid is a single wave of workgroups (no halo reloads
    // between waves)
    std::vector<int> zcs = {0, 32, 64, 128, 256, 512};
    if (c->P.nz > 512) zcs.push_back(c->P.nz);
    const int nv = c->prec == GPE_FP64 ? N_PAIR_VARIANTS_F64 : N_PAIR_VARIANTS;
    float best = 1e30f;
    int bv = c->pvid[pair], bz = c->pzc[pair];
    for (int v = 0; v < nv && rc == GPE_OK; ++v) {
      for (int zc : zcs) {
        if (zc > c->P.nz) continue;
        c->pvid[pair] = v;
        c->pzc[pair] = zc;
        if ((rc = run_pair(c, pair, A, B))) break;  // warm
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(GPE_EHIP, "event"); break; }
        for (int r = 0; r < reps && rc == GPE_OK; ++r) rc = run_pair(c, pair, A, B);
        if (rc) break;
        float ms = 0.f;
        hipError_t he = hipEventRecord(e1, c->stream);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) {
          // a failed trial must not be taken for the fastest
          rc = fail(GPE_EHIP, "autotune timing: %s", hipGetErrorString(he));
          break;
        }
        if (ms < best) {
          best = ms;
          bv = v;
          bz = zc;
        }
      }
    }
    c->pvid[pair] = bv;
    c->pzc[pair] = bz;
    pbest[pair] = best;
    if (rc) break;
  }
  for (int kind : kinds) {
    StageArgs A = args_for(c);
    A.sdtr = c->st.dt_re;
    A.sdti = c->st.dt_im;
    A.psi = c->psi;
    A.lo = c->lo;
    A.psi_out = c->psi2;
    A.lo_out = c->lo2;
    A.acc_in = c->acc;
    A.acc_out = c->acc;
    A.scale_in = c->scale;
    A.scale_out = c->scale;
    A.write_scale = 0;
    for (int m = 0; m < 5; ++m) A.kin[m] = c->k[m] ? c->k[m] : c->psi;
    // outputs into a k-field the kind does not read
    int nin = (kind >= K_CK_2 && kind <= K_CK_6) ? kind - K_CK_1 : (kind == K_RK2_2 || kind == K_RK4_2 || kind == K_RK4_3 || kind == K_RK4_4) ? 1 : 0;
    A.kout = c->k[nin < 5 ? nin : 0];
    if (kind == K_RK4_2) A.kout = c->k[1];
    if (kind == K_RK4_3) { A.kin[0] = c->k[1]; A.kout = c->k[0]; }
    std::vector<int> zcs = {0, 16, 32, 64, 128, 256, 512};  // as for the pairs
    if (c->P.nz > 512) zcs.push_back(c->P.nz);
    float best = 1e30f;
    int bv = c->vid[kind], bz = c->zc[kind];
    for (int v = 0; v < N_TILE_VARIANTS && rc == GPE_OK; ++v) {
      if (!variant_allowed(c->prec == GPE_FP64, kind, v)) continue;
      for (int zc : zcs) {
        if (zc > c->P.nz) continue;
        c->vid[kind] = v;
        c->zc[kind] = zc;
        if ((rc = run_kind(c, kind, A))) break;  // warm
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(GPE_EHIP, "event"); break; }
        for (int r = 0; r < reps && rc == GPE_OK; ++r) rc = run_kind(c, kind, A);
        if (rc) break;
        float ms = 0.f;
        hipError_t he = hipEventRecord(e1, c->stream);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) {
          // a failed trial must not be taken for the fastest
          rc = fail(GPE_EHIP, "autotune timing: %s", hipGetErrorString(he));
          break;
        }
        if (ms < best) {
          best = ms;
          bv = v;
          bz = zc;
        }
      }
    }
    c->vid[kind] = bv;
    c->zc[kind] = bz;
    kbest[kind] = best;
    if (rc) break;
  }
  {
    static const int first[N_PAIR_KINDS] = {K_RK4_1, K_RK4_3, K_CK_1, K_CK_3, K_CK_5};
    // a clear margin, so timing noise does not pick the slower form
    for (int pair : pairs) c->pfuse[pair] = pbest[pair] < 0.97f * (kbest[first[pair]] + kbest[first[pair] + 1]);
  }
  c->prof = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return GPE_OK;
}

int gpe_probe_bandwidth(int64_t bytes, int32_t reps, double *read_gbs, double *write_gbs, double *copy_gbs) {
  if (bytes < (1 << 20) || reps < 1) return fail(GPE_EINVAL, "bytes >= 1 MiB and reps >= 1 required");
  double g[3] = {0, 0, 0};
  hipError_t e = probe_bandwidth((size_t)bytes, reps, g);
  if (e != hipSuccess) return fail(GPE_EHIP, "bandwidth probe: %s", hipGetErrorString(e));
  if (read_gbs) *read_gbs = g[0];
  if (write_gbs) *write_gbs = g[1];
  if (copy_gbs) *copy_gbs = g[2];
  return GPE_OK;
}

int gpe_prof_enable(gpe_ctx *c, int32_t enable) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (c->prof) prof_drain(c);
  c->prof = enable != 0;
  for (int k = 0; k < GPE_PROF_KINDS; ++k) {
    c->prof_n[k] = 0;
    c->prof_ms[k] = 0.0;
  }
  return GPE_OK;
}

int gpe_prof_query(gpe_ctx *c, int32_t kind, int64_t *launches, double *total_ms, double *bytes_per_launch,
                   char *name) {
  if (!c) return fail(GPE_EINVAL, "NULL context");
  if (kind < 0 || kind >= GPE_PROF_KINDS) return fail(GPE_EINVAL, "kind out of range");
  prof_drain(c);
  if (launches) *launches = c->prof_n[kind];
  if (total_ms) *total_ms = c->prof_ms[kind];
  if (bytes_per_launch) {
    double passes = kind_passes(kind);
    double bytes = passes * (double)c->esz;
    if (c->comp && kind_final(kind)) bytes += 2.0 * c->esz;
    *bytes_per_launch = bytes * (double)c->n;
  }
  if (name) {
    std::snprintf(name, 64, "%s", kind_name(kind));
  }
  return GPE_OK;
}

}  // extern "C"
