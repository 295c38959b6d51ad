"""Continuous-batching service for GP simulations on one GPU.

Requests (grid, initial state, trap, steps, scheme, precision) arrive over
time.  The server keeps every admitted simulation's field and its RK work
fields resident in HBM inside one `Batch`, admits queued requests first come
first served while their device footprint fits the HBM budget (and the
resident site count stays under a cap that keeps the persistent sweeps full
without stretching every iteration), and runs the batch in chunks of
iterations; between chunks it retires finished requests (result copied to
the host, memory returned to the pool) and admits new ones, so new work
joins the running step instead of waiting for a drain.

The reference has no service mode; each request's result equals the same
simulation run alone (see tests/test_serve.py).

    python -m gpe_amd.serve --rate 40 --duration 120
"""
import argparse
import json
import os
import sys
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import batch as _batch
from .batch import Batch
from .sim import random_phase_prepare
from .workloads import Request


# Algorithmic bytes per site-update of the fused-sweep model (RK4 13 field
# passes per step, RKF45 27 per attempt, counted once per accepted step; the
# compensated form adds its low word's read and write).
_BYTES = {("rk4", "fp64"): 208.0, ("rk4", "fp32c"): 120.0, ("rk45", "fp64"): 432.0, ("rk45", "fp32c"): 232.0}


def _model_bytes(cfg):
    return _BYTES.get((cfg.scheme, cfg.precision), 0.0)


class _Ready:
    """A computed initial state kept for a request still in the queue."""

    def __init__(self, v):
        self.v = v

    def result(self):
        return self.v

    def cancel(self):
        return False


class Server:
    def __init__(self, hbm_budget=0, max_sites=None, target_chunk_s=0.15, ic_workers=4, prefetch=8,
                 deliver=None, device_ic=True, autotune_every_s=None, backfill_s=None):
        """hbm_budget: device bytes for resident simulations (0: 90% of free).
        max_sites: cap on resident grid points (None: no cap).
        target_chunk_s: device time per chunk of iterations between
        retirements and admissions.
        deliver(request, psi, info): called with each finished result.
        device_ic: build random-phase initial states on the device (hipFFT)
        instead of on the host (the other states are built on the host by
        ic_workers threads, `prefetch` requests ahead).
        autotune_every_s: when requests of a (precision, order, scheme)
        group not yet tuned are resident, and at least this many seconds
        after the last tuning, time the tile variants and z-chunks of every
        group on the resident instances (gpe_batch_autotune: no step is
        committed); None: keep the default launch configuration.
        backfill_s: let queued requests that fit the resident-site cap go
        ahead of one that does not, until the queue's head has waited this
        long (None: strict first come first served)."""
        self.batch = Batch(hbm_budget)
        self.budget = self.batch.stats()["budget"]
        self.max_sites = max_sites
        self.target_chunk_s = target_chunk_s
        self.pool = ThreadPoolExecutor(max_workers=ic_workers)
        self.prefetch = prefetch
        self.deliver = deliver
        self.device_ic = device_ic
        self.autotune_every_s = autotune_every_s
        self.backfill_s = backfill_s

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def serve(self, requests, duration=None, drain=True, clock=None, log=None, checkpoint_dir=None,
              resume_dir=None):
        """Serve `requests` (sorted by .arrival, seconds from the start).

        duration: stop admitting at this time (None: admit all).  drain: keep
        running after the last admission until every admitted request is
        done.  clock: callable giving seconds since the start (default: wall
        clock).  checkpoint_dir: when the window ends with requests still
        running, save them there (fields in the end_state format, at step
        boundaries) instead of dropping them.  resume_dir: re-admit the
        requests of such a checkpoint first.  Returns the metrics dict."""
        t0 = time.perf_counter()
        if clock is None:
            clock = lambda: time.perf_counter() - t0  # noqa: E731
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        futures = {}
        running = {}      # batch id -> (request, admit time, sites)
        resident_sites = 0
        if resume_dir is not None:
            with open(os.path.join(resume_dir, "requests.json")) as fh:
                meta = json.load(fh)
            for old, new in self.batch.restore(resume_dir).items():
                if new is None:
                    raise RuntimeError("a checkpointed request does not fit the HBM budget")
                m = meta[str(old)]
                r = Request(m["rid"], m["arrival"], self.batch.cfgs[new], m["steps"], m["ic_kind"], m["seed"])
                running[new] = (r, clock(), r.cfg.nx * r.cfg.ny * r.cfg.nz)
                resident_sites += r.cfg.nx * r.cfg.ny * r.cfg.nz
        done = []         # (request, arrival, admitted, finished, status, steps)
        chunk = 4
        tuned, last_tune = set(), -1e30
        busy_s = 0.0
        su_done = 0.0
        model_bytes = 0.0
        last_log = 0.0

        def on_device(r):
            return self.device_ic and r.ic_kind == "random_phase"

        def prefetch():
            for r in list(queue)[: self.prefetch]:
                if r.rid not in futures:
                    # host states are built ahead; for device-built ones the
                    # FFT plan of the grid is made ahead
                    job = (lambda c=r.cfg: random_phase_prepare(c)) if on_device(r) else r.initial_state
                    futures[r.rid] = self.pool.submit(job)

        while True:
            now = clock()
            while pending and pending[0].arrival <= now and (duration is None or pending[0].arrival < duration):
                queue.append(pending.popleft())
            prefetch()
            # admission, first come first served.  With backfill_s set, a
            # request that does not fit the resident-site cap lets the ones
            # behind it that do fit go ahead, until the queue's head has
            # waited backfill_s; then nothing passes it.  The HBM budget
            # (memory) stops admission outright.
            i = 0
            while i < len(queue):
                r = queue[i]
                sites = r.cfg.nx * r.cfg.ny * r.cfg.nz
                if self.max_sites is not None and running and resident_sites + sites > self.max_sites:
                    head_waited = clock() - queue[0].arrival
                    if self.backfill_s is None or head_waited >= self.backfill_s:
                        break
                    i += 1
                    continue
                if _batch.footprint(r.cfg) > self.budget:
                    raise ValueError("request %d can never fit the HBM budget" % r.rid)
                if on_device(r):
                    # the random-phase state is built in the instance's memory
                    f = futures.pop(r.rid, None)
                    if f is not None:
                        f.result()  # its FFT plan is ready
                    iid = r.add_to(self.batch, device_ic=True)
                    if iid is None:  # does not fit now: wait for memory
                        if not running:
                            raise RuntimeError("request %d does not fit in device memory with nothing resident"
                                               % r.rid)
                        break
                else:
                    f = futures.pop(r.rid, None)
                    psi = f.result() if f is not None else r.initial_state()
                    iid = self.batch.add(r.cfg, psi, r.steps)
                    if iid is None:  # does not fit now: wait for memory
                        if not running:
                            raise RuntimeError("request %d does not fit in device memory with nothing resident"
                                               % r.rid)
                        futures[r.rid] = _Ready(psi)
                        break
                del queue[i]
                running[iid] = (r, clock(), sites)
                resident_sites += sites
                prefetch()
            no_more = not pending or (duration is not None and pending[0].arrival >= duration)
            if not running:
                if no_more and not queue:
                    break
                # idle until the next arrival
                if pending:
                    wait = pending[0].arrival - clock()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            if duration is not None and not drain and clock() >= duration:
                break
            if self.autotune_every_s is not None:
                keys = {(r.cfg.precision, r.cfg.order, r.cfg.scheme) for r, _, _ in running.values()}
                if keys - tuned and clock() - last_tune >= self.autotune_every_s:
                    self.batch.autotune(reps=1)
                    tuned |= keys
                    last_tune = clock()
            ts = time.perf_counter()
            self.batch.run(chunk)
            finished = self.batch.poll()
            el = time.perf_counter() - ts
            busy_s += el
            # size the next chunk for ~target_chunk_s of device time
            per_it = el / chunk
            chunk = int(max(1, min(256, round(self.target_chunk_s / max(per_it, 1e-6)))))
            tnow = clock()
            for iid, status in finished:
                r, tadm, sites = running.pop(iid)
                q = self.batch.query(iid)
                psi = self.batch.download(iid) if self.deliver is not None else None
                self.batch.remove(iid)
                resident_sites -= sites
                assert resident_sites >= 0, resident_sites
                su_done += float(sites) * q["steps"]
                model_bytes += float(sites) * q["steps"] * _model_bytes(r.cfg)
                done.append((r, r.arrival, tadm, tnow, status, q["steps"]))
                if self.deliver is not None:
                    self.deliver(r, psi, q)
            if log is not None and tnow - last_log >= 5.0:
                last_log = tnow
                log("t=%.1fs running %d queued %d done %d chunk %d" % (tnow, len(running), len(queue), len(done),
                                                                      chunk))
        end = clock()
        for f in futures.values():
            f.cancel()
        if checkpoint_dir is not None and running:
            self.batch.checkpoint(checkpoint_dir, ids=list(running))
            meta = {str(i): {"rid": r.rid, "arrival": r.arrival, "steps": r.steps, "ic_kind": r.ic_kind,
                             "seed": r.seed} for i, (r, _, _) in running.items()}
            with open(os.path.join(checkpoint_dir, "requests.json"), "w") as fh:
                json.dump(meta, fh)
        # work done on requests still running counts toward throughput
        su_partial = 0.0
        for iid, (r, _, sites) in running.items():
            k = float(sites) * self.batch.query(iid)["steps"]
            su_partial += k
            model_bytes += k * _model_bytes(r.cfg)
        m = self._metrics(done, queue, pending, running, end, busy_s, su_done + su_partial)
        # for the caller: the finished requests (request, arrival, admitted,
        # finished, status, steps) and the fused-sweep model bytes of all
        # work done (accepted steps)
        self.last_done = done
        self.last_model_bytes = model_bytes
        m["site_updates_per_s_completed_only"] = su_done / end if end > 0 else 0.0
        return m

    @staticmethod
    def _metrics(done, queue, pending, running, end, busy_s, su_done):
        lat = np.array([d[3] - d[1] for d in done]) if done else np.zeros(0)
        wait = np.array([d[2] - d[1] for d in done]) if done else np.zeros(0)
        ok = sum(1 for d in done if d[4] == 1)
        # requests that arrived but did not finish: their latency is at least
        # end - arrival (censored); folded in for a lower bound on the tail
        unfinished = [end - r.arrival for r in list(queue) + [v[0] for v in running.values()] if r.arrival <= end]
        allat = np.concatenate([lat, np.array(unfinished)]) if unfinished else lat

        def pct(a, q):
            return float(np.percentile(a, q)) if a.size else None

        failures = [{"rid": d[0].rid, "status": d[4], "scheme": d[0].cfg.scheme, "precision": d[0].cfg.precision,
                     "n": d[0].cfg.nx, "order": d[0].cfg.order, "bcs": d[0].cfg.bcs, "ic": d[0].ic_kind,
                     "steps_done": d[5]} for d in done if d[4] != 1][:20]
        return {
            "completed": len(done), "ok": ok, "failed": len(done) - ok, "failures": failures,
            "unfinished": len(unfinished), "not_arrived": len(pending),
            "elapsed_s": end, "solves_per_s": ok / end if end > 0 else 0.0,
            "site_updates_per_s": su_done / end if end > 0 else 0.0,
            "latency_p50_s": pct(lat, 50), "latency_p99_s": pct(lat, 99),
            "queue_wait_p50_s": pct(wait, 50), "queue_wait_p99_s": pct(wait, 99),
            "latency_all_p50_s_lower_bound": pct(allat, 50), "latency_all_p99_s_lower_bound": pct(allat, 99),
            "device_busy_frac": busy_s / end if end > 0 else 0.0,
        }


class NativeServer:
    """The library's own serving loop (gpe_serve_* in the C ABI): the same
    first-come-first-served admission by HBM footprint and resident-site cap
    over one batch, with the scheduling in C++ so C and Fortran callers have
    it too.  Python only submits, steps and collects."""

    def __init__(self, hbm_budget=0, max_sites=0):
        import ctypes
        from . import _lib
        self._c = ctypes
        self._L = _lib.lib()
        self._h = ctypes.c_void_p()
        _lib.check(self._L.gpe_serve_create(int(hbm_budget), int(max_sites), ctypes.byref(self._h)))
        self._req = {}

    def close(self):
        if self._h:
            self._L.gpe_serve_destroy(self._h)
            self._h = self._c.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, r, device_ic=True):
        """Queue request r (a workloads.Request); returns its ticket.
        Random-phase states are built on the device at admission."""
        from . import _lib
        from .sim import make_params
        c = self._c
        st = _lib.GpeStepState()
        dt = complex(r.cfg.dt)
        st.dt_re, st.dt_im = dt.real, dt.imag
        t = c.c_int64()
        if device_ic and r.ic_kind == "random_phase":
            rc = self._L.gpe_serve_submit(self._h, c.byref(make_params(r.cfg)), None, _lib.GPE_HOST_C128,
                                          int(r.seed), 0.5, c.byref(st), int(r.steps), c.byref(t))
        else:
            psi = np.ascontiguousarray(r.initial_state(), dtype=np.complex128)
            rc = self._L.gpe_serve_submit(self._h, c.byref(make_params(r.cfg)), psi.ctypes.data_as(c.c_void_p),
                                          _lib.GPE_HOST_C128, 0, 0.0, c.byref(st), int(r.steps), c.byref(t))
        _lib.check(rc)
        self._req[t.value] = r
        return t.value

    def step(self, iterations):
        """One round: admit, run `iterations`, retire.  Returns (running, queued)."""
        from . import _lib
        c = self._c
        run, q = c.c_int32(), c.c_int32()
        _lib.check(self._L.gpe_serve_step(self._h, int(iterations), c.byref(run), c.byref(q)))
        return run.value, q.value

    def poll(self):
        """[(ticket, status)] of requests finished since the last poll."""
        from . import _lib
        c = self._c
        cap = max(1, len(self._req))
        tk = (c.c_int64 * cap)()
        stt = (c.c_int32 * cap)()
        n = c.c_int32()
        _lib.check(self._L.gpe_serve_poll(self._h, tk, stt, cap, c.byref(n)))
        return [(tk[i], stt[i]) for i in range(n.value)]

    def result(self, ticket):
        """(field as complex128, state dict) of a finished request."""
        from . import _lib
        c = self._c
        r = self._req.pop(ticket)
        out = np.empty(r.cfg.shape, dtype=np.complex128)
        st = _lib.GpeStepState()
        _lib.check(self._L.gpe_serve_result(self._h, int(ticket), out.ctypes.data_as(c.c_void_p),
                                            _lib.GPE_HOST_C128, c.byref(st)))
        return out, {"t": st.t, "dt": complex(st.dt_re, st.dt_im), "im_t": st.im_t, "steps": st.p}


def main():
    from .workloads import request_stream
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=40.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--budget-gib", type=float, default=0.0)
    ap.add_argument("--max-msites", type=float, default=64.0, help="resident grid points cap, millions")
    ap.add_argument("--no-drain", action="store_true")
    ap.add_argument("--host-ic", action="store_true", help="build random-phase states on the host")
    ap.add_argument("--chunk-s", type=float, default=0.15, help="device time per chunk of iterations, seconds")
    ap.add_argument("--autotune-every-s", type=float, default=3.0,
                    help="tune the launch configuration of new groups, at most this often (seconds; <= 0: never)")
    ap.add_argument("--backfill-s", type=float, default=None,
                    help="let smaller requests pass one that does not fit, until it has waited this long (seconds)")
    a = ap.parse_args()
    reqs = request_stream(a.rate, a.duration, a.seed)
    with Server(hbm_budget=int(a.budget_gib * 2 ** 30), max_sites=a.max_msites * 1e6, device_ic=not a.host_ic,
                target_chunk_s=a.chunk_s,
                autotune_every_s=a.autotune_every_s if a.autotune_every_s > 0 else None,
                backfill_s=a.backfill_s) as s:
        m = s.serve(reqs, duration=a.duration, drain=not a.no_drain,
                    log=lambda *x: print("[serve]", *x, file=sys.stderr, flush=True))
    m.update({"rate": a.rate, "duration": a.duration, "requests": len(reqs)})
    print(json.dumps(m), flush=True)


if __name__ == "__main__":
    main()
