"""Benchmark of the GP time-stepping path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the configuration the headline metric is quoted on): random-phase
field on a 512^3 periodic grid, dx = 0.5, 4th-order FD, Cash-Karp RKF45 with
error tolerance 1e-6, V = 0, g = mu = 1.  A step is one accepted rk45 step
(one solver() call, all attempts included) of the whole field.  The
compensated FP32 form (FP32 storage and stencil, FP64 reductions, step
control and psi update) is the headline `value`; the plain FP64 path runs
on the same input beside it.

`value` is site-updates per second over the K timed steps, field resident
in HBM.  `roofline` is for the dominant kernel, timed with HIP events
recorded on the stream it is launched on over the same K steps; its
algorithmic bytes are the stage's share of the fused-sweep traffic model
(RKF45: 27 field passes per attempt, + the low word read and write of the
compensated form); `traffic` is the kernel's HBM bytes per launch from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary.  `long_run` advances the
FP32C field 1000 accepted steps from t = 0 and reports its reject ratio.
`cpu_baseline` is the C oracle (a port of the reference's FP64 path) timed
on this host on the same 512^3 field, 3 steps minus 1.  The batch leg runs
64 uneven simulations against the same ones run one after another in plain
FP64; the serve leg feeds the 40 requests/s Poisson stream for 120 s and
then runs a bounded sample of the completed requests one by one in plain
FP64.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site-updates/s + % HBM roofline at 512^3 FP32 vs FP64; batched solves/s, p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(precision, n=512):
    from gpe_amd.config import RunConfig, xr_for_spacing
    return RunConfig(nx=n, ny=n, nz=n, xr=xr_for_spacing(0.5, n), bcs=1, order=4, scheme="rk45",
                     eqn=4, g=1.0, mu=1.0, tau=0.05, precision=precision)


def run_path(precision, psi, steps, warmup, tune=True):
    from gpe_amd import tuning
    from gpe_amd.sim import Simulation
    cfg = workload(precision)
    sim = Simulation(cfg, psi)
    if tune:
        # launch configuration from the autotuner (setup, not timed)
        if not tuning.apply(sim):
            tuning.tune(sim, reps=3)
        sim.set_state(cfg.dt)
        sim.upload(psi)
    log(precision, "warmup", warmup)
    sim.advance(warmup)
    s0 = sim.state()
    # timed region, with HIP events recorded around every stage launch on
    # the context's stream (the kernel times of the roofline); the advance
    # call synchronises the stream before returning
    sim.prof_enable(True)
    t0 = time.perf_counter()
    sim.advance(steps)
    t1 = time.perf_counter()
    prof = sim.prof_query()
    s1 = sim.state()
    elapsed = t1 - t0
    tiles = {k: v for k, v in sim.tune_get().items() if k in prof}
    sim.close()
    sites = cfg.nx * cfg.ny * cfg.nz
    return {
        "elapsed_s": elapsed,
        "su_per_s": sites * steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "t_sim": s1["t"] - s0["t"],
        "dt_last": s1["dt"].real,
        "prof": prof,
        "sites": sites,
        "tiles": tiles,
        "attempts_per_step": _attempts(prof) / steps,
    }


def long_run(precision, psi, steps):
    """The same field advanced `steps` accepted RKF45 steps from t = 0 (the
    stored launch configuration): site-updates/s over the whole run and its
    rejected-attempt ratio, which a short window from the initial field
    does not show (dt first grows for a few hundred steps)."""
    from gpe_amd.sim import Simulation
    cfg = workload(precision)
    with Simulation(cfg, psi) as sim:
        sim.prof_enable(True)
        t0 = time.perf_counter()
        sim.advance(steps)
        el = time.perf_counter() - t0
        attempts = _attempts(sim.prof_query())
        st = sim.state()
    sites = cfg.nx * cfg.ny * cfg.nz
    return {"precision": precision, "steps": steps, "attempts": attempts,
            "reject_ratio": round((attempts - steps) / attempts, 4) if attempts else None,
            "su_per_s": round(sites * steps / el, 1), "ms_per_step": round(1e3 * el / steps, 4),
            "t_end": st["t"]}


def _attempts(prof):
    # one final-stage launch per rk45 attempt (fused: the ck_56 pair)
    return prof.get("ck_6", prof.get("ck_56", (0,)))[0]


def measured_traffic():
    """HBM bytes per launch per kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on copy kernels)."""
    path = os.path.join(ROOT, "profiles", "rk45_512cube_pmc_traffic.md")
    if not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return json.loads(line.split(":", 1)[1])
    return {}


def roofline_of(res, precision):
    # dominant kernel = largest total event time
    kind, (n, ms, by) = max(res["prof"].items(), key=lambda kv: kv[1][1])
    avg_s = ms / n / 1e3
    achieved = by / avg_s / 1e9
    traffic = measured_traffic().get("%s/%s" % (precision, kind))
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_ms": round(avg_s * 1e3, 4), "bytes_per_launch": by,
            "traffic_gbs": round(traffic / avg_s / 1e9, 1) if traffic else None}


def step_roofline(res, bytes_per_site_attempt):
    # whole-step achieved bandwidth over all stage kernels
    tot_ms = sum(v[1] for v in res["prof"].values())
    attempts = _attempts(res["prof"])
    if not attempts:
        return None
    bytes_ = bytes_per_site_attempt * res["sites"] * attempts
    return round(bytes_ / (tot_ms / 1e3) / 1e9, 1)


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(psi):
    """The oracle (a C port of the reference's FP64 path) on this host's
    cores: the same 512^3 field advanced 1 and then 3 accepted steps; the
    difference of the two times is two steps without setup."""
    from oracle import pyoracle
    cfg = workload("fp64")
    nt = min(16, os.cpu_count() or 1)
    times = []
    for k in (1, 3):
        st = pyoracle.make_state(cfg.dt)
        t0 = time.perf_counter()
        out, _ = pyoracle.run(cfg, st, psi, k, nt)
        times.append(time.perf_counter() - t0)
        del out
    el = times[1] - times[0]
    sites = cfg.nx * cfg.ny * cfg.nz
    return {"value": round(2 * sites / el, 1), "unit": "site-updates/s", "cores": nt, "kind": "port",
            "host_cpu": host_cpu(), "host_threads": os.cpu_count(),
            "sample": "the same 512^3 field, FP64: 3 accepted rk45 steps minus 1 (%.1f s - %.1f s = %.1f s "
                      "for 2 steps, setup differenced out)" % (times[1], times[0], el)}


def batch_leg(seed, steps=None, log_=log):
    """64 uneven simulations (N in {32..128}, RK4, 500 steps, FP32C) in one
    batch, against the same simulations run one after another on the plain
    FP64 single-simulation path.  Fields are resident before each timed
    region; the batch's launch configuration is autotuned first (commits no
    step)."""
    from gpe_amd.batch import Batch
    from gpe_amd.sim import Simulation
    from gpe_amd.workloads import batch_mix
    mix = batch_mix(seed)
    if steps:
        for m in mix:
            m.steps = steps
    psis = [m.initial_state() for m in mix]
    sites = sum(m.cfg.nx * m.cfg.ny * m.cfg.nz for m in mix)
    su = float(sum(m.cfg.nx * m.cfg.ny * m.cfg.nz * m.steps for m in mix))
    with Batch() as b:
        ids = [b.add(m.cfg, p, m.steps) for m, p in zip(mix, psis)]
        b.autotune(reps=2)
        b.run(0)  # work tables and graph capture, outside the timed region
        b.sync()
        t0 = time.perf_counter()
        b.run(mix[0].steps)
        b.sync()
        el_total = time.perf_counter() - t0
        done = dict(b.poll())
        assert len(done) == len(ids) and all(v == 1 for v in done.values()), done
        got = [b.download(i) for i in ids]
        tiles = {k: b.tune_get("fp32c", 2, "rk4", k) for k in ("rk4_1", "rk4_2", "rk4_3", "rk4_4", "rk4_12", "rk4_34")}
    log_("batch fp32c: %d solves in %.3f s" % (len(ids), el_total))
    # sequential plain FP64, one simulation after another
    seq = 0.0
    errs = []
    for m, p, g in zip(mix, psis, got):
        c = m.cfg.with_(precision="fp64")
        with Simulation(c, p) as sim:
            sim.advance(1)
            sim.set_state(c.dt)
            sim.upload(p)
            t0 = time.perf_counter()
            sim.advance(m.steps)
            seq += time.perf_counter() - t0
            ref = sim.download()
        errs.append(float(np.linalg.norm(g - ref) / np.linalg.norm(ref)))
    log_("sequential fp64: %.3f s" % seq)
    return {"instances": len(mix), "sites": sites, "steps": mix[0].steps,
            "solves_per_s": round(len(mix) / el_total, 3),
            "seq_fp64_solves_per_s": round(len(mix) / seq, 3),
            "speedup_vs_seq_fp64": round(seq / el_total, 3),
            "site_updates_per_s": round(su / el_total, 1),
            "achieved_gbs": round(su * 120.0 / el_total / 1e9, 1),
            "psi_rel_l2_vs_fp64_max": max(errs), "tiles": tiles}


def seq_fp64(reqs, budget_s, log_=log):
    """The requests run one after another on the plain FP64 single-simulation
    path (the comparator of the served mix), until budget_s of timed work."""
    from gpe_amd.sim import Simulation
    el, su, n = 0.0, 0.0, 0
    for r in reqs:
        c = r.cfg.with_(precision="fp64")
        with Simulation(c) as sim:
            if r.ic_kind == "random_phase":
                sim.random_phase(r.seed, kmax=0.5)
            else:
                sim.upload(r.initial_state())
            t0 = time.perf_counter()
            sim.advance(r.steps)
            el += time.perf_counter() - t0
        su += float(c.nx * c.ny * c.nz) * r.steps
        n += 1
        if el >= budget_s:
            break
    log_("sequential fp64: %d of the completed requests in %.1f s" % (n, el))
    return {"requests": n, "elapsed_s": round(el, 3), "solves_per_s": round(n / el, 4) if el else None,
            "site_updates_per_s": round(su / el, 1) if el else None}


def serve_leg(rate, duration, seed, seq_budget_s=30.0, log_=log):
    """Poisson request stream into the continuous-batching server, then the
    completed requests one by one on the plain FP64 path (bounded sample)."""
    from gpe_amd.serve import Server
    from gpe_amd.workloads import request_stream
    reqs = request_stream(rate, duration, seed)
    # launch configuration of each (precision, order, scheme) group tuned on
    # the resident requests as groups appear (at most every 3 s; no step is
    # committed by tuning); requests that fit the resident-site cap may pass
    # one that does not until it has waited 5 s
    with Server(max_sites=64e6, autotune_every_s=3.0, backfill_s=5.0) as s:
        m = s.serve(reqs, duration=duration, drain=False, log=lambda *x: log_("serve", *x))
        done = [d[0] for d in s.last_done if d[4] == 1]
        model_bytes = s.last_model_bytes
    m.update({"rate": rate, "duration_s": duration, "requests": len(reqs)})
    # bytes of the fused-sweep model over all resident work (accepted steps;
    # rk45 rejections not counted, so a lower bound on the bytes moved)
    m["achieved_gbs_model"] = round(model_bytes / m["elapsed_s"] / 1e9, 1) if m["elapsed_s"] else None
    if seq_budget_s > 0 and done:
        seq = seq_fp64(done, seq_budget_s, log_)
        m["seq_fp64"] = seq
        if seq["site_updates_per_s"]:
            m["served_over_seq_fp64_site_updates"] = round(m["site_updates_per_s"] / seq["site_updates_per_s"], 3)
            m["served_over_seq_fp64_solves"] = round(m["solves_per_s"] / seq["solves_per_s"], 3)
    return m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-simulation batch leg")
    ap.add_argument("--serve-seconds", type=float, default=120.0,
                    help="length of the Poisson serving leg (0 to skip)")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batch leg and print its JSON (for profiling)")
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="accepted steps of the FP32C long run (reject ratio; 0 to skip)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from gpe_amd import _lib, ic
    if a.only_batch:
        print(json.dumps({"batch": batch_leg(a.seed)}), flush=True)
        return
    probe = _lib.probe_bandwidth(1 << 31, 20)
    log("HBM probe GB/s: read %.0f write %.0f copy %.0f" % (probe["read"], probe["write"], probe["copy"]))
    t0 = time.perf_counter()
    psi = ic.random_phase(workload("fp64"), a.seed)
    log("IC ready in %.1f s" % (time.perf_counter() - t0))

    r32 = run_path("fp32c", psi, a.steps, a.warmup)
    log("fp32c %.3e su/s, %.3f ms/step" % (r32["su_per_s"], r32["ms_per_step"]))
    r64 = run_path("fp64", psi, a.steps, a.warmup)
    log("fp64  %.3e su/s, %.3f ms/step" % (r64["su_per_s"], r64["ms_per_step"]))
    lr = None
    if a.long_steps > 0:
        lr = long_run("fp32c", psi, a.long_steps)
        log("fp32c long run: %d steps, %d attempts, %.3e su/s" % (lr["steps"], lr["attempts"], lr["su_per_s"]))

    del psi
    bat = None if a.no_batch else batch_leg(a.seed)
    srv = serve_leg(40.0, a.serve_seconds, a.seed) if a.serve_seconds > 0 else None
    psi = ic.random_phase(workload("fp64"), a.seed)
    cpu = None if a.no_cpu else cpu_baseline(psi)
    if cpu:
        log("cpu baseline %.3e su/s on %d threads" % (cpu["value"], cpu["cores"]))

    out = {
        "metric": METRIC,
        "value": round(r32["su_per_s"], 1),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "random-phase 512^3 periodic, dx=0.5, 4th-order FD, RKF45 tol 1e-6, "
                               "FP32 storage + compensated psi update (FP64 beside)",
                   "grid": [512, 512, 512], "scheme": "rk45", "order": 4, "bcs": "periodic",
                   "seed": a.seed},
        "roofline": roofline_of(r32, "fp32c"),
        "cpu_baseline": cpu,
        "fp64": {"value": round(r64["su_per_s"], 1), "ms_per_step": round(r64["ms_per_step"], 4),
                 "roofline": roofline_of(r64, "fp64")},
        "fp32c_over_fp64": round(r32["su_per_s"] / r64["su_per_s"], 3),
        "hbm_probe_gbs": {k: round(v, 1) for k, v in probe.items()},
        "sim_time": {"fp32c": round(r32["t_sim"], 6), "fp64": round(r64["t_sim"], 6)},
        "attempts_per_step": {"fp32c": r32["attempts_per_step"], "fp64": r64["attempts_per_step"]},
        "long_run": lr,
        "tiles": {"fp32c": r32["tiles"], "fp64": r64["tiles"]},
        "step_gbs": {"fp32c": step_roofline(r32, 232.0), "fp64": step_roofline(r64, 432.0)},
        "kernels": {"fp32c": {k: [v[0], round(v[1], 3)] for k, v in r32["prof"].items()},
                    "fp64": {k: [v[0], round(v[1], 3)] for k, v in r64["prof"].items()}},
        "batch": bat,
        "serve": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
