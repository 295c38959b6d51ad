// gpe_batch.cpp -- many independent simulations advanced together.
//
// Each instance is its own solver() loop (solve.f90:25-53 called as in
// gpe.f90:230) with its own grid, boundary conditions, trap, dt and step
// count.  Instances of one (precision, order, scheme) form a group; per
// stage kind, a group is swept by ONE persistent launch over a work table
// of (instance, tile) items, and after the final stage one control kernel
// advances every instance (rk4 bookkeeping, or the accept/reject and dt
// update of rkqs, solve.f90:180-229) on the device.  One iteration (all
// groups' stages + control) is captured once in a hipGraph and replayed;
// the graph is captured again only when the set of instances changes.
//
// Per-site arithmetic is the single-simulation kernels' (stage_tile), so an
// instance's field does not depend on the batch it runs in; its reductions
// use only its own partial slots, in a fixed order, so its dt sequence and
// norm do not either.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/gpe_hip.h"
#include "gpe_ic.h"
#include "host_common.h"
#include "launch.h"

using namespace gpe;

namespace {

constexpr int kBatchDefaultVid = 0;
// the default tile variant of stage kind k in a group of precision prec
int default_vid(int prec, int k) { return variant_allowed(prec == GPE_FP64, k, kBatchDefaultVid) ? kBatchDefaultVid : kVidF64Ck; }
constexpr int kDefaultZc = 32;
constexpr int kMaxGrid = 2048;  // persistent workgroups per batched launch

// Stage kinds of one iteration of each scheme: [k0, k0 + nk).
void scheme_kinds(int scheme, int *k0, int *nk) {
  switch (scheme) {
    case GPE_EULER: *k0 = K_EULER; *nk = 1; break;
    case GPE_RK2: *k0 = K_RK2_1; *nk = 2; break;
    case GPE_RK4: *k0 = K_RK4_1; *nk = 4; break;
    default: *k0 = K_CK_1; *nk = 6; break;
  }
}

// Device fields of one instance (allocated as one block).
int n_work_fields(int scheme) {
  switch (scheme) {
    case GPE_EULER: return 0;
    case GPE_RK2: return 1;            // k1
    case GPE_RK4: return 3;            // k1/k3, k2, acc
    default: return 6;                 // kk1..kk5, error scale
  }
}

size_t footprint(const gpe_params *p) {
  const size_t n = (size_t)p->nx * p->ny * p->nz;
  const size_t esz = p->precision == GPE_FP64 ? 16 : 8;
  size_t fields = 2 + n_work_fields(p->scheme);  // psi, psi'
  if (p->precision == GPE_FP32C) fields += 2;     // lo, lo'
  // the compensated form keeps no rk45 error scale
  if (p->precision == GPE_FP32C && p->scheme == GPE_RK45) fields -= 1;
  // round each field to 256 bytes
  return fields * ((n * esz + 255) & ~(size_t)255);
}

struct Inst {
  int id = -1;
  gpe_params P{};
  int group = -1;
  void *block = nullptr;
  size_t bytes = 0;
  void *buf[2] = {nullptr, nullptr}, *lob[2] = {nullptr, nullptr};
  void *k[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  void *acc = nullptr, *scale = nullptr;
  long long steps_target = 0;
  bool reported = false;
};

struct KindTable {
  WorkItem *d = nullptr;
  int cap = 0, n = 0, grid = 0;
  std::vector<WorkItem> h;
};

// Fused stage pairs (fused.hpp) and their first stage kinds.
constexpr int kPairFirst[P_COUNT] = {K_RK4_1, K_RK4_3, K_CK_1, K_CK_3, K_CK_5};

int pair_at(int kind) {
  for (int q = 0; q < P_COUNT; ++q)
    if (kPairFirst[q] == kind) return q;
  return -1;
}

struct Group {
  int prec = GPE_FP64, order = 2, scheme = GPE_RK4;
  int R = 1, k0 = 0, nk = 0;
  int vid[K_COUNT], zc[K_COUNT];
  KindTable tab[K_COUNT];
  // pairs run fused where the autotuner found that faster
  bool pfuse[P_COUNT];
  int pvid[P_COUNT], pzc[P_COUNT];
  KindTable ptab[P_COUNT];
  std::vector<int> slots;  // instance slots in this group, ascending
  int part_off = 0;        // first partial slot of this group's final stage
};

// One launch of an iteration: a stage kind, or a fused pair (first kind k).
struct Unit {
  int kind, pair;  // pair = -1 for a single stage
  bool final_;
};

std::vector<Unit> units_of(const Group &G) {
  std::vector<Unit> u;
  for (int k = G.k0; k < G.k0 + G.nk; ++k) {
    const int q = pair_at(k);
    if (q >= 0 && G.pfuse[q] && k + 1 < G.k0 + G.nk) {
      u.push_back({k, q, k + 1 == G.k0 + G.nk - 1});
      ++k;
    } else {
      u.push_back({k, -1, k == G.k0 + G.nk - 1});
    }
  }
  return u;
}

}  // namespace

constexpr int kAuxStreams = 3;
constexpr int kMaxGroups = 16;  // 2 precisions x 2 orders x 4 schemes

struct gpe_batch {
  hipStream_t stream = nullptr;
  // captured iterations run the groups concurrently: forked onto these
  // streams and joined before the control kernel
  hipStream_t aux[kAuxStreams] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_done[kMaxGroups] = {};
  int cap = 0;                 // instance slots in the device tables
  StageArgs *d_args = nullptr; // [K_COUNT][cap]
  InstCtl *d_ctl = nullptr;    // [cap]
  int *d_active = nullptr;     // [cap]
  int *d_range = nullptr;      // [2 * cap]
  double *d_pn = nullptr, *d_pe = nullptr;
  int part_cap = 0;
  std::vector<Inst> slot;      // [cap]; id < 0 = free
  std::map<int, int> id2slot;
  std::vector<Group> groups;
  int next_id = 1;
  size_t budget = 0, in_use = 0, pooled = 0;
  std::multimap<size_t, void *> pool;  // freed instance blocks, by size
  bool dirty = true;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<InstCtl> hctl;
};

namespace {

int group_of(gpe_batch *b, int prec, int order, int scheme) {
  for (size_t g = 0; g < b->groups.size(); ++g) {
    const Group &G = b->groups[g];
    if (G.prec == prec && G.order == order && G.scheme == scheme) return (int)g;
  }
  Group G;
  G.prec = prec;
  G.order = order;
  G.scheme = scheme;
  G.R = order / 2;
  scheme_kinds(scheme, &G.k0, &G.nk);
  for (int k = 0; k < K_COUNT; ++k) {
    G.vid[k] = default_vid(prec, k);
    G.zc[k] = kDefaultZc;
  }
  for (int q = 0; q < P_COUNT; ++q) {
    G.pfuse[q] = false;
    G.pvid[q] = 0;
    G.pzc[q] = 0;
  }
  b->groups.push_back(G);
  return (int)b->groups.size() - 1;
}

void drop_graph(gpe_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Grow the per-slot device tables to hold at least `want` slots, keeping
// their contents.  Only called with the stream idle.
int grow_slots(gpe_batch *b, int want) {
  if (want <= b->cap) return GPE_OK;
  int ncap = b->cap ? b->cap : 16;
  while (ncap < want) ncap *= 2;
  StageArgs *na = nullptr;
  InstCtl *nc = nullptr;
  int *nact = nullptr, *nr = nullptr;
  if (hipMalloc((void **)&na, sizeof(StageArgs) * K_COUNT * (size_t)ncap) != hipSuccess ||
      hipMalloc((void **)&nc, sizeof(InstCtl) * (size_t)ncap) != hipSuccess ||
      hipMalloc((void **)&nact, sizeof(int) * (size_t)ncap) != hipSuccess ||
      hipMalloc((void **)&nr, 2 * sizeof(int) * (size_t)ncap) != hipSuccess) {
    if (na) (void)hipFree(na);
    if (nc) (void)hipFree(nc);
    if (nact) (void)hipFree(nact);
    if (nr) (void)hipFree(nr);
    return fail(GPE_ENOMEM, "batch tables for %d instances", ncap);
  }
  GPE_HIPCHK(hipMemset(nc, 0, sizeof(InstCtl) * (size_t)ncap));
  GPE_HIPCHK(hipMemset(nact, 0, sizeof(int) * (size_t)ncap));
  GPE_HIPCHK(hipMemset(nr, 0, 2 * sizeof(int) * (size_t)ncap));
  if (b->cap) {
    for (int k = 0; k < K_COUNT; ++k)
      GPE_HIPCHK(hipMemcpy(na + (size_t)k * ncap, b->d_args + (size_t)k * b->cap, sizeof(StageArgs) * b->cap,
                           hipMemcpyDeviceToDevice));
    GPE_HIPCHK(hipMemcpy(nc, b->d_ctl, sizeof(InstCtl) * b->cap, hipMemcpyDeviceToDevice));
    GPE_HIPCHK(hipMemcpy(nact, b->d_active, sizeof(int) * b->cap, hipMemcpyDeviceToDevice));
    (void)hipFree(b->d_args);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->d_active);
    (void)hipFree(b->d_range);
  }
  b->d_args = na;
  b->d_ctl = nc;
  b->d_active = nact;
  b->d_range = nr;
  b->cap = ncap;
  b->slot.resize(ncap);
  b->hctl.resize(ncap);
  b->dirty = true;
  return GPE_OK;
}

int take_block(gpe_batch *b, size_t bytes, void **out) {
  auto it = b->pool.find(bytes);
  if (it != b->pool.end()) {
    *out = it->second;
    b->pool.erase(it);
    b->pooled -= bytes;
    return GPE_OK;
  }
  // make room: release pooled blocks of other sizes first
  while (b->in_use + b->pooled + bytes > b->budget && !b->pool.empty()) {
    auto last = std::prev(b->pool.end());
    b->pooled -= last->first;
    (void)hipFree(last->second);
    b->pool.erase(last);
  }
  if (b->in_use + bytes > b->budget)
    return fail(GPE_EFULL, "instance needs %zu bytes; %zu of %zu in use", bytes, b->in_use, b->budget);
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) return fail(GPE_ENOMEM, "hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
  return GPE_OK;
}

void give_block(gpe_batch *b, void *p, size_t bytes) {
  b->pool.emplace(bytes, p);
  b->pooled += bytes;
}

// StageArgs of kind `kind` for instance `I` at the instance's initial
// state (the control kernel keeps dt and the state buffers current).
StageArgs inst_args(const Inst &I, int kind, const gpe_step_state &st) {
  StageArgs A{};
  fill_base(&I.P, I.P.precision, A);
  A.dtr = st.dt_re;
  A.dti = st.dt_im;
  A.hr = 0.5 * st.dt_re;
  A.hi = 0.5 * st.dt_im;
  A.sdtr = st.dt_re;
  A.sdti = st.dt_im;
  A.inv_absdt = 1.0 / std::hypot(st.dt_re, st.dt_im);
  A.write_scale = 1;
  A.psi = I.buf[0];
  A.psi_out = I.buf[1];
  A.lo = I.lob[0];
  A.lo_out = I.lob[1];
  switch (kind) {
    case K_RK2_1: A.kout = I.k[0]; break;
    case K_RK2_2: A.kin[0] = I.k[0]; break;
    // k[0]: k1 then k3 (k1 is folded into acc by stage 2), k[1]: k2
    case K_RK4_1: A.kout = I.k[0]; break;
    case K_RK4_2:
      A.kin[0] = I.k[0];
      A.kout = I.k[1];
      A.acc_out = I.acc;
      break;
    case K_RK4_3:
      A.kin[0] = I.k[1];
      A.kout = I.k[0];
      break;
    case K_RK4_4:
      A.kin[0] = I.k[0];
      A.acc_in = I.acc;
      break;
    default:
      if (kind >= K_CK_1 && kind <= K_CK_6) {
        const int s = kind - K_CK_1 + 1;
        for (int m = 0; m < s - 1; ++m) A.kin[m] = I.k[m];
        if (s < 6) A.kout = I.k[s - 1];
        if (s == 1) A.scale_out = I.scale;
        if (s == 6) A.scale_in = I.scale;
      }
  }
  return A;
}

// Work table of one kind of a group: every tile of every instance, each
// instance's partial slots contiguous in (z, y, x) tile order; the table
// itself is ordered largest tile first so the persistent grid ends evenly.
// Work table over output tiles of TX x TY sites (a stage kind: 64 x TY; a
// fused pair: the (64-2R) x (TY-2R) interiors).
void make_table(gpe_batch *b, const Group &G, int TX, int TY, int zc, std::vector<WorkItem> &out,
                std::vector<int> *range, int part_off) {
  out.clear();
  int slotc = part_off;
  std::vector<long long> cost;
  for (int s : G.slots) {
    const Inst &I = b->slot[s];
    const int nx = I.P.nx, ny = I.P.ny, nz = I.P.nz;
    const int zcl = zc > 0 ? std::min(zc, nz) : nz;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    const int beg = slotc;
    for (int z0 = 0; z0 < nz; z0 += zcl) {
      const int z1 = std::min(z0 + zcl, nz);
      for (int ty = 0; ty < nty; ++ty)
        for (int tx = 0; tx < ntx; ++tx) {
          WorkItem w;
          w.inst = s;
          w.tx = tx;
          w.ty = ty;
          w.z0 = z0;
          w.z1 = z1;
          w.slot = slotc++;
          out.push_back(w);
          const long long cx = std::min(TX, nx - TX * tx), cy = std::min(TY, ny - TY * ty);
          cost.push_back((long long)(cx * cy) * (z1 - z0 + 2 * G.R));
        }
    }
    if (range) {
      (*range)[2 * s] = beg;
      (*range)[2 * s + 1] = slotc;
    }
  }
  std::vector<int> ord(out.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  std::vector<WorkItem> sorted(out.size());
  for (size_t i = 0; i < ord.size(); ++i) sorted[i] = out[ord[i]];
  out.swap(sorted);
}

void kind_tiles(int vid, int *TX, int *TY) {
  *TX = 64;
  *TY = kTiles[vid].nw * kTiles[vid].p;
}

void pair_tiles(const Group &G, int vid, int *TX, int *TY) {
  *TX = 64 - 2 * G.R;
  *TY = kPairTiles[vid].nw * kPairTiles[vid].p - 2 * G.R;
}

int upload_table(KindTable &T) {
  const int n = (int)T.h.size();
  if (n > T.cap) {
    if (T.d) (void)hipFree(T.d);
    T.d = nullptr;
    int c = std::max(n, 2 * T.cap);
    if (hipMalloc((void **)&T.d, sizeof(WorkItem) * (size_t)c) != hipSuccess)
      return fail(GPE_ENOMEM, "work table of %d items", c);
    T.cap = c;
  }
  if (n) GPE_HIPCHK(hipMemcpy(T.d, T.h.data(), sizeof(WorkItem) * (size_t)n, hipMemcpyHostToDevice));
  T.n = n;
  T.grid = std::min(n, kMaxGrid);
  return GPE_OK;
}

int ensure_partials(gpe_batch *b, int need) {
  if (need <= b->part_cap) return GPE_OK;
  if (b->d_pn) (void)hipFree(b->d_pn);
  if (b->d_pe) (void)hipFree(b->d_pe);
  b->d_pn = b->d_pe = nullptr;
  int c = std::max(need, 2 * b->part_cap);
  if (hipMalloc((void **)&b->d_pn, sizeof(double) * (size_t)c) != hipSuccess ||
      hipMalloc((void **)&b->d_pe, sizeof(double) * (size_t)c) != hipSuccess) {
    b->part_cap = 0;
    return fail(GPE_ENOMEM, "partials for %d tiles", c);
  }
  b->part_cap = c;
  return GPE_OK;
}

// Regenerate every group's work tables and partial ranges.  Stream idle.
int rebuild(gpe_batch *b) {
  int rc;
  std::vector<int> range(2 * (size_t)b->cap, 0);
  int part = 0;
  for (Group &G : b->groups) {
    G.slots.clear();
    for (int s = 0; s < b->cap; ++s)
      if (b->slot[s].id >= 0 && b->slot[s].group == (int)(&G - &b->groups[0])) G.slots.push_back(s);
    G.part_off = part;
    for (const Unit &u : units_of(G)) {
      int TX, TY;
      KindTable &T = u.pair >= 0 ? G.ptab[u.pair] : G.tab[u.kind];
      if (u.pair >= 0)
        pair_tiles(G, G.pvid[u.pair], &TX, &TY);
      else
        kind_tiles(G.vid[u.kind], &TX, &TY);
      make_table(b, G, TX, TY, u.pair >= 0 ? G.pzc[u.pair] : G.zc[u.kind], T.h, u.final_ ? &range : nullptr,
                 part);
      if ((rc = upload_table(T))) return rc;
      if (u.final_) part += (int)T.h.size();
    }
  }
  if ((rc = ensure_partials(b, std::max(part, 1)))) return rc;
  if (b->cap) GPE_HIPCHK(hipMemcpy(b->d_range, range.data(), sizeof(int) * range.size(), hipMemcpyHostToDevice));
  drop_graph(b);
  b->dirty = false;
  return GPE_OK;
}

// One iteration: every group's stages, then the per-instance control.
int enqueue_iteration(gpe_batch *b, bool fork) {
  int ngroups = 0;
  for (Group &G : b->groups) ngroups += !G.slots.empty();
  fork = fork && ngroups > 1;
  if (fork) GPE_HIPCHK(hipEventRecord(b->ev_fork, b->stream));
  int j = 0;
  for (Group &G : b->groups) {
    if (G.slots.empty()) continue;
    hipStream_t s = b->stream;
    if (fork) {
      s = b->aux[j % kAuxStreams];
      GPE_HIPCHK(hipStreamWaitEvent(s, b->ev_fork, 0));
    }
    for (const Unit &u : units_of(G)) {
      const int k = u.kind;
      const KindTable &T = u.pair >= 0 ? G.ptab[u.pair] : G.tab[k];
      if (!T.n) continue;
      hipError_t e;
      if (u.pair >= 0) {
        PairBatchLaunch L;
        L.instA = b->d_args + (size_t)k * b->cap;
        L.instB = b->d_args + (size_t)(k + 1) * b->cap;
        L.work = T.d;
        L.nwork = T.n;
        L.active = b->d_active;
        L.pn = u.final_ ? b->d_pn : nullptr;
        L.pe = u.final_ ? b->d_pe : nullptr;
        L.grid = T.grid;
        e = launch_pair_batched(G.prec == GPE_FP64 ? 16 : 8, G.pvid[u.pair], u.pair, G.R, G.prec == GPE_FP32C, L, s);
      } else {
        BatchLaunch L;
        L.insts = b->d_args + (size_t)k * b->cap;
        L.work = T.d;
        L.nwork = T.n;
        L.active = b->d_active;
        L.pn = u.final_ ? b->d_pn : nullptr;
        L.pe = u.final_ ? b->d_pe : nullptr;
        L.grid = T.grid;
        e = launch_stage_batched(G.prec == GPE_FP64 ? 16 : 8, G.vid[k], k, G.R, G.prec == GPE_FP32C, L, s);
      }
      if (e != hipSuccess) return fail(GPE_EHIP, "batched stage launch: %s", hipGetErrorString(e));
    }
    if (fork) GPE_HIPCHK(hipEventRecord(b->ev_done[j], s));
    ++j;
  }
  if (fork)
    for (int q = 0; q < j; ++q) GPE_HIPCHK(hipStreamWaitEvent(b->stream, b->ev_done[q], 0));
  hipError_t e = launch_batch_control(b->d_ctl, b->d_args, b->cap, b->d_active, b->d_range, b->d_pn, b->d_pe,
                                      b->stream);
  if (e != hipSuccess) return fail(GPE_EHIP, "batch control launch: %s", hipGetErrorString(e));
  return GPE_OK;
}

int capture(gpe_batch *b) {
  GPE_HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_iteration(b, true);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(GPE_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return fail(GPE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return GPE_OK;
}

int prepare(gpe_batch *b) {
  int rc;
  if (b->dirty) {
    GPE_HIPCHK(hipStreamSynchronize(b->stream));
    if ((rc = rebuild(b))) return rc;
  }
  if (b->use_graph && !b->exec && (rc = capture(b))) return rc;
  return GPE_OK;
}

int check_batch_params(const gpe_params *p) {
  if (!p) return fail(GPE_EINVAL, "params is NULL");
  if (p->nx < 5 || p->ny < 5 || p->nz < 5) return fail(GPE_EINVAL, "grid must be at least 5 points per axis");
  if (p->scheme != GPE_EULER && p->scheme != GPE_RK2 && p->scheme != GPE_RK4 && p->scheme != GPE_RK45)
    return fail(GPE_EINVAL, "ERROR: Unrecognised time stepping scheme.");
  if (p->eqn < 1 || p->eqn > 4) return fail(GPE_EINVAL, "eqn_to_solve must be 1..4");
  if (p->bcs != 1 && p->bcs != 2) return fail(GPE_EINVAL, "bcs must be 1 (periodic) or 2 (reflective)");
  if (p->order != 2 && p->order != 4) return fail(GPE_EINVAL, "order must be 2 or 4");
  if (p->precision != GPE_FP64 && p->precision != GPE_FP32C)
    return fail(GPE_EINVAL, "batches run GPE_FP64 or GPE_FP32C instances");
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0)) return fail(GPE_EINVAL, "grid spacing must be positive");
  return GPE_OK;
}

int find(gpe_batch *b, int id) {
  auto it = b->id2slot.find(id);
  return it == b->id2slot.end() ? -1 : it->second;
}

int fetch_ctl(gpe_batch *b) {
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  if (b->cap)
    GPE_HIPCHK(hipMemcpy(b->hctl.data(), b->d_ctl, sizeof(InstCtl) * (size_t)b->cap, hipMemcpyDeviceToHost));
  return GPE_OK;
}

}  // namespace

extern "C" {

int gpe_batch_create(int64_t hbm_budget, gpe_batch **out) {
  if (!out) return fail(GPE_EINVAL, "out is NULL");
  *out = nullptr;
  gpe_batch *b = new (std::nothrow) gpe_batch();
  if (!b) return fail(GPE_ENOMEM, "host allocation failed");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(GPE_EHIP, "hipStreamCreate failed");
  }
  bool ok = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
  for (int q = 0; q < kAuxStreams && ok; ++q) ok = hipStreamCreateWithFlags(&b->aux[q], hipStreamNonBlocking) == hipSuccess;
  for (int q = 0; q < kMaxGroups && ok; ++q) ok = hipEventCreateWithFlags(&b->ev_done[q], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    gpe_batch_destroy(b);
    return fail(GPE_EHIP, "batch streams/events");
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
  // default: 90% of what is free now
  b->budget = hbm_budget > 0 ? (size_t)hbm_budget : (size_t)(0.9 * (double)fr);
  int rc = grow_slots(b, 16);
  if (!rc) rc = ensure_partials(b, 4096);
  if (rc) {
    gpe_batch_destroy(b);
    return rc;
  }
  *out = b;
  return GPE_OK;
}

void gpe_batch_destroy(gpe_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (Inst &I : b->slot)
    if (I.block) (void)hipFree(I.block);
  for (auto &kv : b->pool) (void)hipFree(kv.second);
  for (Group &G : b->groups)
    for (KindTable &T : G.tab)
      if (T.d) (void)hipFree(T.d);
  void *ptrs[] = {b->d_args, b->d_ctl, b->d_active, b->d_range, b->d_pn, b->d_pe};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (hipStream_t q : b->aux)
    if (q) (void)hipStreamDestroy(q);
  for (hipEvent_t ev : b->ev_done)
    if (ev) (void)hipEventDestroy(ev);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int gpe_batch_footprint(const gpe_params *params, int64_t *bytes) {
  int rc = check_batch_params(params);
  if (rc) return rc;
  if (!bytes) return fail(GPE_EINVAL, "bytes is NULL");
  *bytes = (int64_t)footprint(params);
  return GPE_OK;
}

namespace {
// The initial field of a new instance: a host array, or the random-phase
// state built on the device.
struct FieldSource {
  const void *psi = nullptr;
  int32_t host_type = 0;
  bool random_phase = false;
  uint64_t seed = 0;
  double kmax = 0.0;
};

int add_instance(gpe_batch *b, const gpe_params *params, const FieldSource &src, const gpe_step_state *st,
                 int64_t nsteps, int32_t *id);
}  // namespace

int gpe_batch_add(gpe_batch *b, const gpe_params *params, const void *psi, int32_t host_type,
                  const gpe_step_state *st, int64_t nsteps, int32_t *id) {
  if (!b || !psi || !st || !id) return fail(GPE_EINVAL, "NULL argument");
  FieldSource src;
  src.psi = psi;
  src.host_type = host_type;
  return add_instance(b, params, src, st, nsteps, id);
}

int gpe_batch_add_random_phase(gpe_batch *b, const gpe_params *params, uint64_t seed, double kmax,
                               const gpe_step_state *st, int64_t nsteps, int32_t *id) {
  if (!b || !st || !id) return fail(GPE_EINVAL, "NULL argument");
  FieldSource src;
  src.random_phase = true;
  src.seed = seed;
  src.kmax = kmax;
  return add_instance(b, params, src, st, nsteps, id);
}

namespace {
int add_instance(gpe_batch *b, const gpe_params *params, const FieldSource &src, const gpe_step_state *st,
                 int64_t nsteps, int32_t *id) {
  int rc = check_batch_params(params);
  if (rc) return rc;
  if (nsteps < 1) return fail(GPE_EINVAL, "nsteps must be >= 1");
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  const size_t bytes = footprint(params);
  int s = -1;
  for (int i = 0; i < b->cap; ++i)
    if (b->slot[i].id < 0) {
      s = i;
      break;
    }
  if (s < 0) {
    s = b->cap;
    if ((rc = grow_slots(b, b->cap + 1))) return rc;
  }
  Inst I;
  I.P = *params;
  I.bytes = bytes;
  if ((rc = take_block(b, bytes, &I.block))) return rc;
  {
    const size_t n = (size_t)params->nx * params->ny * params->nz;
    const size_t esz = params->precision == GPE_FP64 ? 16 : 8;
    const size_t fb = (n * esz + 255) & ~(size_t)255;
    char *p = (char *)I.block;
    I.buf[0] = p;
    I.buf[1] = p + fb;
    p += 2 * fb;
    if (params->precision == GPE_FP32C) {
      I.lob[0] = p;
      I.lob[1] = p + fb;
      p += 2 * fb;
    }
    if (params->scheme == GPE_RK2) {
      I.k[0] = p;
    } else if (params->scheme == GPE_RK4) {
      I.k[0] = p;
      I.k[1] = p + fb;
      I.acc = p + 2 * fb;
    } else if (params->scheme == GPE_RK45) {
      for (int m = 0; m < 5; ++m) I.k[m] = p + m * fb;
      if (params->precision != GPE_FP32C) I.scale = p + 5 * fb;
    }
    if (src.random_phase)
      rc = random_phase_device(params->nx, params->ny, params->nz, params->dx, params->dy, params->dz, src.kmax,
                               src.seed, params->precision, I.buf[0], I.lob[0], b->stream);
    else
      rc = field_upload(params->precision, (size_t)params->nx * params->ny, params->nz, src.psi, src.host_type,
                        (size_t)params->nx * params->ny, 0, I.buf[0], I.lob[0]);
    if (rc) {
      give_block(b, I.block, bytes);
      return rc;
    }
  }
  I.id = b->next_id++;
  I.group = group_of(b, params->precision, params->order, params->scheme);
  I.steps_target = nsteps;
  const Group &G = b->groups[I.group];
  // per-kind launch descriptors of this slot
  for (int k = G.k0; k < G.k0 + G.nk; ++k) {
    StageArgs A = inst_args(I, k, *st);
    GPE_HIPCHK(hipMemcpy(b->d_args + (size_t)k * b->cap + s, &A, sizeof A, hipMemcpyHostToDevice));
  }
  InstCtl C{};
  C.active = 1;
  C.status = 0;
  C.scheme = params->scheme;
  C.real_time = params->real_time;
  C.cur = 0;
  C.first_attempt = 1;
  C.steps_done = 0;
  C.steps_target = nsteps;
  C.k0 = G.k0;
  C.nk = G.nk;
  C.t = st->t;
  C.im_t = st->im_t;
  C.dt_re = st->dt_re;
  C.dt_im = st->dt_im;
  C.sdt_re = st->dt_re;
  C.sdt_im = st->dt_im;
  C.safety = params->safety;
  C.eps = params->eps;
  C.dt_inc = params->dt_increase;
  C.dt_dec = params->dt_decrease;
  // errcon = (5/safety)**(1/dt_increase) (solve.f90:173)
  C.errcon = params->scheme == GPE_RK45 ? std::pow(5.0 / params->safety, 1.0 / params->dt_increase) : 0.0;
  C.dv = params->dx * params->dy * params->dz;
  for (int q = 0; q < 2; ++q) {
    C.buf[q] = I.buf[q];
    C.lob[q] = I.lob[q];
  }
  GPE_HIPCHK(hipMemcpy(b->d_ctl + s, &C, sizeof C, hipMemcpyHostToDevice));
  const int one = 1;
  GPE_HIPCHK(hipMemcpy(b->d_active + s, &one, sizeof one, hipMemcpyHostToDevice));
  b->in_use += bytes;
  b->slot[s] = I;
  b->id2slot[I.id] = s;
  b->dirty = true;
  *id = I.id;
  return GPE_OK;
}
}  // namespace

int gpe_batch_remove(gpe_batch *b, int32_t id) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  const int s = find(b, id);
  if (s < 0) return fail(GPE_EINVAL, "no instance %d in the batch", id);
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  const int zero = 0;
  GPE_HIPCHK(hipMemcpy(b->d_active + s, &zero, sizeof zero, hipMemcpyHostToDevice));
  GPE_HIPCHK(hipMemcpy(&b->d_ctl[s].active, &zero, sizeof zero, hipMemcpyHostToDevice));
  Inst &I = b->slot[s];
  give_block(b, I.block, I.bytes);
  b->in_use -= I.bytes;
  I = Inst();
  b->id2slot.erase(id);
  b->dirty = true;
  return GPE_OK;
}

int gpe_batch_run(gpe_batch *b, int32_t iterations) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  Range range("batch run");
  if (iterations < 0) return fail(GPE_EINVAL, "iterations must be >= 0");
  if (b->id2slot.empty()) return GPE_OK;
  int rc = prepare(b);  // iterations == 0: only rebuild the tables and capture
  if (rc) return rc;
  for (int i = 0; i < iterations; ++i) {
    if (b->use_graph) {
      GPE_HIPCHK(hipGraphLaunch(b->exec, b->stream));
    } else if ((rc = enqueue_iteration(b, false))) {
      return rc;
    }
  }
  return GPE_OK;
}

int gpe_batch_sync(gpe_batch *b) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  return GPE_OK;
}

int gpe_batch_poll(gpe_batch *b, int32_t *ids, int32_t *status, int32_t cap, int32_t *n) {
  if (!b || !n) return fail(GPE_EINVAL, "NULL argument");
  int rc = fetch_ctl(b);
  if (rc) return rc;
  int m = 0;
  for (int s = 0; s < b->cap && m < cap; ++s) {
    Inst &I = b->slot[s];
    if (I.id < 0 || I.reported) continue;
    const InstCtl &C = b->hctl[s];
    if (C.active) continue;
    I.reported = true;
    if (ids) ids[m] = I.id;
    if (status) status[m] = C.status;
    ++m;
  }
  *n = m;
  return GPE_OK;
}

int gpe_batch_query(gpe_batch *b, int32_t id, gpe_step_state *st, int64_t *steps_done, int32_t *status,
                    double *norm, int32_t *pending_rejects) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  const int s = find(b, id);
  if (s < 0) return fail(GPE_EINVAL, "no instance %d in the batch", id);
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  InstCtl C;
  GPE_HIPCHK(hipMemcpy(&C, b->d_ctl + s, sizeof C, hipMemcpyDeviceToHost));
  if (st) {
    st->dt_re = C.dt_re;
    st->dt_im = C.dt_im;
    st->t = C.t;
    st->im_t = C.im_t;
    st->errmax = C.errmax;
    st->n_rejects = C.last_rejects;
    st->p = (int32_t)C.steps_done;
  }
  if (steps_done) *steps_done = C.steps_done;
  if (status) *status = C.status;
  if (norm) *norm = C.norm;
  if (pending_rejects) *pending_rejects = C.n_rejects;
  return GPE_OK;
}

int gpe_batch_download(gpe_batch *b, int32_t id, void *psi, int32_t host_type) {
  if (!b || !psi) return fail(GPE_EINVAL, "NULL argument");
  const int s = find(b, id);
  if (s < 0) return fail(GPE_EINVAL, "no instance %d in the batch", id);
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  int cur = 0;
  GPE_HIPCHK(hipMemcpy(&cur, &b->d_ctl[s].cur, sizeof cur, hipMemcpyDeviceToHost));
  const Inst &I = b->slot[s];
  return field_download(I.P.precision, (size_t)I.P.nx * I.P.ny, I.P.nz, I.buf[cur & 1], I.lob[cur & 1], psi,
                        host_type);
}

int gpe_batch_stats(gpe_batch *b, int32_t *instances, int32_t *running, int64_t *bytes_in_use,
                    int64_t *budget) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  int rc = fetch_ctl(b);
  if (rc) return rc;
  int n = 0, r = 0;
  for (int s = 0; s < b->cap; ++s)
    if (b->slot[s].id >= 0) {
      ++n;
      if (b->hctl[s].active) ++r;
    }
  if (instances) *instances = n;
  if (running) *running = r;
  if (bytes_in_use) *bytes_in_use = (int64_t)b->in_use;
  if (budget) *budget = (int64_t)b->budget;
  return GPE_OK;
}

int gpe_batch_use_graph(gpe_batch *b, int32_t enable) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  b->use_graph = enable != 0;
  drop_graph(b);
  return GPE_OK;
}

int gpe_batch_tune_set(gpe_batch *b, int32_t precision, int32_t order, int32_t scheme, int32_t kind,
                       int32_t variant, int32_t zchunk) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  if (kind < 16 && (variant < 0 || variant >= N_TILE_VARIANTS)) return fail(GPE_EINVAL, "variant out of range");
  if (zchunk < 0 || (zchunk > 0 && zchunk < 8)) return fail(GPE_EINVAL, "zchunk must be 0 or >= 8");
  if ((precision != GPE_FP64 && precision != GPE_FP32C) || (order != 2 && order != 4))
    return fail(GPE_EINVAL, "bad precision or order");
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  Group &G = b->groups[group_of(b, precision, order, scheme)];
  if (kind >= 16 && kind < 16 + P_COUNT) {
    // a fused pair: variant -1 runs it as its two stages
    const int q = kind - 16;
    if (kPairFirst[q] < G.k0 || kPairFirst[q] + 1 >= G.k0 + G.nk)
      return fail(GPE_EINVAL, "pair %d is not in scheme %d", kind, scheme);
    const int nv = precision == GPE_FP64 ? N_PAIR_VARIANTS_F64 : N_PAIR_VARIANTS;
    if (variant >= nv) return fail(GPE_EINVAL, "pair variant out of range (FP64 has no fused pairs)");
    G.pfuse[q] = variant >= 0;
    if (variant >= 0) G.pvid[q] = variant;
    G.pzc[q] = zchunk;
    b->dirty = true;
    return GPE_OK;
  }
  if (kind < G.k0 || kind >= G.k0 + G.nk) return fail(GPE_EINVAL, "kind %d is not a stage of scheme %d", kind, scheme);
  if (!variant_allowed(precision == GPE_FP64, kind, variant))
    return fail(GPE_EINVAL, "FP64 Cash-Karp stages do not run tile variant %d", variant);
  G.vid[kind] = variant;
  G.zc[kind] = zchunk;
  b->dirty = true;
  return GPE_OK;
}

int gpe_batch_tune_get(gpe_batch *b, int32_t precision, int32_t order, int32_t scheme, int32_t kind,
                       int32_t *variant, int32_t *zchunk) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  for (const Group &G : b->groups)
    if (G.prec == precision && G.order == order && G.scheme == scheme) {
      if (kind >= 16 && kind < 16 + P_COUNT) {
        const int q = kind - 16;
        if (variant) *variant = G.pfuse[q] ? G.pvid[q] : -1;
        if (zchunk) *zchunk = G.pzc[q];
        return GPE_OK;
      }
      if (kind < G.k0 || kind >= G.k0 + G.nk) return fail(GPE_EINVAL, "kind %d is not a stage of scheme %d", kind, scheme);
      if (variant) *variant = G.vid[kind];
      if (zchunk) *zchunk = G.zc[kind];
      return GPE_OK;
    }
  if (variant) *variant = kind >= 16 ? -1 : default_vid(precision, kind);
  if (zchunk) *zchunk = kind >= 16 ? 0 : kDefaultZc;
  return GPE_OK;
}

// Time every tile variant and z-chunk for each stage kind of every group
// over the current instances (reps launches each, stage kernels only: the
// control kernel does not run, so no state is committed) and keep the
// fastest.
int gpe_batch_autotune(gpe_batch *b, int32_t reps) {
  if (!b) return fail(GPE_EINVAL, "batch is NULL");
  if (reps < 1) reps = 1;
  int rc = prepare(b);
  if (rc) return rc;
  GPE_HIPCHK(hipStreamSynchronize(b->stream));
  hipEvent_t e0, e1;
  GPE_HIPCHK(hipEventCreate(&e0));
  GPE_HIPCHK(hipEventCreate(&e1));
  static const int zcs[] = {8, 16, 32, 64, 0};
  for (Group &G : b->groups) {
    if (G.slots.empty()) continue;
    float kbest[K_COUNT];
    for (float &v : kbest) v = 1e30f;
    for (int k = G.k0; k < G.k0 + G.nk; ++k) {
      const bool fin = k == G.k0 + G.nk - 1;
      float best = 1e30f;
      int bv = G.vid[k], bz = G.zc[k];
      KindTable T;
      for (int v = 0; v < N_TILE_VARIANTS; ++v)
        for (int z : zcs) {
          if (!variant_allowed(G.prec == GPE_FP64, k, v)) continue;
          int TX, TY;
          kind_tiles(v, &TX, &TY);
          make_table(b, G, TX, TY, z, T.h, nullptr, 0);
          if ((rc = upload_table(T)) || (rc = ensure_partials(b, T.n))) break;
          BatchLaunch L;
          L.insts = b->d_args + (size_t)k * b->cap;
          L.work = T.d;
          L.nwork = T.n;
          L.active = b->d_active;
          L.pn = fin ? b->d_pn : nullptr;
          L.pe = fin ? b->d_pe : nullptr;
          L.grid = T.grid;
          const int pb = G.prec == GPE_FP64 ? 16 : 8;
          const bool comp = G.prec == GPE_FP32C;
          hipError_t e = launch_stage_batched(pb, v, k, G.R, comp, L, b->stream);  // warm
          e = e != hipSuccess ? e : hipEventRecord(e0, b->stream);
          for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stage_batched(pb, v, k, G.R, comp, L, b->stream);
          e = e != hipSuccess ? e : hipEventRecord(e1, b->stream);
          e = e != hipSuccess ? e : hipEventSynchronize(e1);
          if (e != hipSuccess) {
            rc = fail(GPE_EHIP, "autotune launch: %s", hipGetErrorString(e));
            break;
          }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (ms < best) {
            best = ms;
            bv = v;
            bz = z;
          }
        }
      if (T.d) (void)hipFree(T.d);
      if (rc) break;
      G.vid[k] = bv;
      G.zc[k] = bz;
      kbest[k] = best;
    }
    if (rc) break;
    // fused pairs (FP32 forms): kept where >= 3% faster than their two stages
    for (int k = G.k0; k + 1 < G.k0 + G.nk && rc == GPE_OK && G.prec != GPE_FP64; ++k) {
      const int q = pair_at(k);
      if (q < 0) continue;
      const bool fin = k + 1 == G.k0 + G.nk - 1;
      const int nv = G.prec == GPE_FP64 ? N_PAIR_VARIANTS_F64 : N_PAIR_VARIANTS;
      float best = 1e30f;
      int bv = G.pvid[q], bz = G.pzc[q];
      KindTable T;
      for (int v = 0; v < nv && rc == GPE_OK; ++v)
        for (int z : {0, 32, 64}) {
          int TX, TY;
          pair_tiles(G, v, &TX, &TY);
          make_table(b, G, TX, TY, z, T.h, nullptr, 0);
          if ((rc = upload_table(T)) || (rc = ensure_partials(b, T.n))) break;
          PairBatchLaunch L;
          L.instA = b->d_args + (size_t)k * b->cap;
          L.instB = b->d_args + (size_t)(k + 1) * b->cap;
          L.work = T.d;
          L.nwork = T.n;
          L.active = b->d_active;
          L.pn = fin ? b->d_pn : nullptr;
          L.pe = fin ? b->d_pe : nullptr;
          L.grid = T.grid;
          const int pb = G.prec == GPE_FP64 ? 16 : 8;
          const bool comp = G.prec == GPE_FP32C;
          hipError_t e = launch_pair_batched(pb, v, q, G.R, comp, L, b->stream);  // warm
          e = e != hipSuccess ? e : hipEventRecord(e0, b->stream);
          for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_pair_batched(pb, v, q, G.R, comp, L, b->stream);
          e = e != hipSuccess ? e : hipEventRecord(e1, b->stream);
          e = e != hipSuccess ? e : hipEventSynchronize(e1);
          if (e != hipSuccess) {
            rc = fail(GPE_EHIP, "autotune launch: %s", hipGetErrorString(e));
            break;
          }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (ms < best) {
            best = ms;
            bv = v;
            bz = z;
          }
        }
      if (T.d) (void)hipFree(T.d);
      G.pvid[q] = bv;
      G.pzc[q] = bz;
      G.pfuse[q] = best < 0.97f * (kbest[k] + kbest[k + 1]);
      ++k;  // pairs do not overlap
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->dirty = true;
  return rc;
}

}  // extern "C"
