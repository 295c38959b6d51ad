// gpe_serve.cpp -- a request queue in front of a batch (gpe_serve_* in
// include/gpe_hip.h): continuous batching for C and Fortran callers.
//
// No reference counterpart: the reference runs one simulation per job.
// Requests (a solver() loop of nsteps steps from a field and a state) are
// queued in submission order.  Each gpe_serve_step admits queued requests,
// first come first served, while their device footprint fits the batch's
// HBM budget and the resident grid points stay under the cap; then it runs
// the requested number of batch iterations, in which every resident request
// advances one step (rk45: one attempt), and retires the finished ones: their
// field and state are copied to host memory kept for gpe_serve_result and
// their device memory returns to the batch's pool.  A request's result does
// not depend on what else is resident (the batch's invariance), so it equals
// the same request run alone.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gpe_hip.h"
#include "host_common.h"

using gpe::fail;

namespace {

struct Request {
  int64_t ticket = 0;
  gpe_params P{};
  gpe_step_state st{};
  int64_t nsteps = 0;
  std::vector<unsigned char> field;  // host copy of the initial field (empty: random phase)
  int32_t host_type = GPE_HOST_C128;
  uint64_t seed = 0;
  double kmax = 0.0;
  int64_t sites = 0;
};

struct Result {
  int32_t status = 0;
  gpe_step_state st{};
  gpe_params P{};
  std::vector<unsigned char> field;  // complex128
};

size_t sites_of(const gpe_params &P) { return (size_t)P.nx * (size_t)P.ny * (size_t)P.nz; }

}  // namespace

struct gpe_serve {
  gpe_batch *batch = nullptr;
  int64_t budget = 0;
  int64_t max_sites = 0;  // 0: no cap
  int64_t next_ticket = 1;
  int64_t resident_sites = 0;
  std::deque<Request> queue;
  std::map<int32_t, Request> running;  // batch id -> request (its field released)
  std::map<int64_t, Result> finished;  // ticket -> result, until collected
  std::deque<std::pair<int64_t, int32_t>> unreported;  // finished (ticket, status) not yet polled
};

extern "C" {

int gpe_serve_create(int64_t hbm_budget, int64_t max_sites, gpe_serve **out) {
  if (!out) return fail(GPE_EINVAL, "out is NULL");
  *out = nullptr;
  if (max_sites < 0) return fail(GPE_EINVAL, "max_sites < 0");
  gpe_serve *s = new (std::nothrow) gpe_serve();
  if (!s) return fail(GPE_ENOMEM, "host allocation failed");
  int rc = gpe_batch_create(hbm_budget, &s->batch);
  if (rc) {
    delete s;
    return rc;
  }
  int64_t budget = 0;
  if ((rc = gpe_batch_stats(s->batch, nullptr, nullptr, nullptr, &budget))) {
    gpe_batch_destroy(s->batch);
    delete s;
    return rc;
  }
  s->budget = budget;
  s->max_sites = max_sites;
  *out = s;
  return GPE_OK;
}

void gpe_serve_destroy(gpe_serve *s) {
  if (!s) return;
  gpe_batch_destroy(s->batch);
  delete s;
}

int gpe_serve_submit(gpe_serve *s, const gpe_params *params, const void *psi, int32_t host_type, uint64_t seed,
                     double kmax, const gpe_step_state *st, int64_t nsteps, int64_t *ticket) {
  if (!s || !params || !st || !ticket) return fail(GPE_EINVAL, "NULL argument");
  if (nsteps < 1) return fail(GPE_EINVAL, "nsteps must be >= 1");
  int64_t bytes = 0;
  int rc = gpe_batch_footprint(params, &bytes);
  if (rc) return rc;
  if (bytes > s->budget)
    return fail(GPE_EFULL, "request needs %lld device bytes, more than the budget of %lld", (long long)bytes,
                (long long)s->budget);
  Request r;
  r.P = *params;
  r.st = *st;
  r.nsteps = nsteps;
  r.sites = (int64_t)sites_of(*params);
  if (psi) {
    if (host_type != GPE_HOST_C128 && host_type != GPE_HOST_C64) return fail(GPE_EINVAL, "host_type");
    const size_t nb = (size_t)r.sites * (host_type == GPE_HOST_C128 ? 16 : 8);
    try {
      r.field.assign(static_cast<const unsigned char *>(psi), static_cast<const unsigned char *>(psi) + nb);
    } catch (const std::bad_alloc &) {
      return fail(GPE_ENOMEM, "host copy of the initial field (%zu bytes)", nb);
    }
    r.host_type = host_type;
  } else {
    if (!(kmax > 0.0)) return fail(GPE_EINVAL, "random-phase request needs kmax > 0");
    r.seed = seed;
    r.kmax = kmax;
  }
  r.ticket = s->next_ticket++;
  *ticket = r.ticket;
  s->queue.push_back(std::move(r));
  return GPE_OK;
}

int gpe_serve_step(gpe_serve *s, int32_t iterations, int32_t *running, int32_t *queued) {
  if (!s) return fail(GPE_EINVAL, "serve is NULL");
  if (iterations < 0) return fail(GPE_EINVAL, "iterations must be >= 0");
  int rc;
  // admission, first come first served: the first request that does not fit
  // waits, and so do the ones behind it
  while (!s->queue.empty()) {
    Request &r = s->queue.front();
    if (s->max_sites > 0 && !s->running.empty() && s->resident_sites + r.sites > s->max_sites) break;
    int32_t id = -1;
    if (r.field.empty())
      rc = gpe_batch_add_random_phase(s->batch, &r.P, r.seed, r.kmax, &r.st, r.nsteps, &id);
    else
      rc = gpe_batch_add(s->batch, &r.P, r.field.data(), r.host_type, &r.st, r.nsteps, &id);
    if (rc == GPE_EFULL) break;  // waits for memory
    if (rc) {
      // cannot be admitted at all (e.g. a device error building its state):
      // it leaves the queue as finished with status -3, so the requests
      // behind it are not held up, and the call reports the error
      Result res;
      res.status = -3;
      res.P = r.P;
      res.st = r.st;
      const std::string msg = gpe::g_err;
      s->unreported.emplace_back(r.ticket, res.status);
      s->finished.emplace(r.ticket, std::move(res));
      s->queue.pop_front();
      return fail(rc, "request %lld not admitted: %s", (long long)r.ticket, msg.c_str());
    }
    s->resident_sites += r.sites;
    r.field.clear();
    r.field.shrink_to_fit();
    s->running.emplace(id, std::move(r));
    s->queue.pop_front();
  }
  if (!s->running.empty() && iterations > 0) {
    if ((rc = gpe_batch_run(s->batch, iterations))) return rc;
  }
  // retire
  std::vector<int32_t> ids(s->running.size() + 1), status(s->running.size() + 1);
  int32_t n = 0;
  if ((rc = gpe_batch_poll(s->batch, ids.data(), status.data(), (int32_t)ids.size(), &n))) return rc;
  // Every finished request leaves `running` (the batch reports it once); one
  // whose result cannot be read back finishes with status -3 and no field,
  // and the first such error is returned after the others are retired.
  int first_rc = GPE_OK;
  std::string first_msg;
  for (int32_t i = 0; i < n; ++i) {
    auto it = s->running.find(ids[i]);
    if (it == s->running.end()) continue;
    Result res;
    res.status = status[i];
    res.P = it->second.P;
    res.st = it->second.st;
    rc = GPE_OK;
    try {
      res.field.resize(sites_of(res.P) * 16);
    } catch (const std::bad_alloc &) {
      rc = fail(GPE_ENOMEM, "host copy of a result");
    }
    if (!rc) rc = gpe_batch_query(s->batch, ids[i], &res.st, nullptr, nullptr, nullptr, nullptr);
    if (!rc) rc = gpe_batch_download(s->batch, ids[i], res.field.data(), GPE_HOST_C128);
    const int rc_remove = gpe_batch_remove(s->batch, ids[i]);
    if (!rc) rc = rc_remove;
    if (rc) {
      res.status = -3;
      res.field.clear();
      if (!first_rc) {
        first_rc = rc;
        first_msg = "result of request " + std::to_string(it->second.ticket) + ": " + gpe::g_err;
      }
    }
    s->resident_sites -= it->second.sites;
    s->unreported.emplace_back(it->second.ticket, res.status);
    s->finished.emplace(it->second.ticket, std::move(res));
    s->running.erase(it);
  }
  if (running) *running = (int32_t)s->running.size();
  if (queued) *queued = (int32_t)s->queue.size();
  if (first_rc) return fail(first_rc, "%s", first_msg.c_str());
  return GPE_OK;
}

int gpe_serve_poll(gpe_serve *s, int64_t *tickets, int32_t *status, int32_t cap, int32_t *n) {
  if (!s || !n) return fail(GPE_EINVAL, "NULL argument");
  if (cap < 0) return fail(GPE_EINVAL, "cap < 0");
  int32_t m = 0;
  while (m < cap && !s->unreported.empty()) {
    const std::pair<int64_t, int32_t> t = s->unreported.front();
    s->unreported.pop_front();
    if (tickets) tickets[m] = t.first;
    if (status) status[m] = t.second;
    ++m;
  }
  *n = m;
  return GPE_OK;
}

int gpe_serve_result(gpe_serve *s, int64_t ticket, void *psi, int32_t host_type, gpe_step_state *st) {
  if (!s || !psi) return fail(GPE_EINVAL, "NULL argument");
  auto it = s->finished.find(ticket);
  if (it == s->finished.end()) return fail(GPE_ESTATE, "request %lld has not finished (or was collected)", (long long)ticket);
  const Result &r = it->second;
  if (r.field.empty()) {
    const int status = r.status;
    s->finished.erase(it);
    return fail(GPE_ESTATE, "request %lld has no field (status %d)", (long long)ticket, status);
  }
  const size_t n = sites_of(r.P);
  const double *src = reinterpret_cast<const double *>(r.field.data());
  if (host_type == GPE_HOST_C128) {
    std::memcpy(psi, src, n * 16);
  } else if (host_type == GPE_HOST_C64) {
    float *d = static_cast<float *>(psi);
    for (size_t i = 0; i < 2 * n; ++i) d[i] = (float)src[i];
  } else {
    return fail(GPE_EINVAL, "host_type");
  }
  if (st) *st = r.st;
  s->finished.erase(it);
  return GPE_OK;
}

}  // extern "C"
