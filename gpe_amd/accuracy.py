"""Accuracy of the compensated FP32 form against the FP64 path on the GPU.

    python -m gpe_amd.accuracy [--n 512] [--steps 1000]

Runs the random-phase periodic field (dx = 0.5, 4th-order FD, RKF45 with
tolerance 1e-6, V = 0, g = mu = 1) for `steps` accepted steps in FP64 and
in FP32C from the same initial state, and reports for each: simulated time,
the reference's energy functional (energy, variables.f90:580-604), the
eqn-4 Hamiltonian H = K + (g/2)int|psi|^4 - mu N that the dynamics conserve,
the norm, and the vortex-line length (linelength, variables.f90:896-1162);
then the relative L2 distance of the two fields, the differences of those
quantities, and the vortex-core position difference: the distance from
each core point of one run (the reference's zeros file, io.f90:886-1236)
to the nearest core point of the other, in units of dx.

H is gpe_sim_hamiltonian: with periodic walls int Re(-lap(psi) psi*) = 2K,
so the reference functional E_ref = 2K + (1/2) int |psi|^4 gives K, and one
device pass sums the quartic, trap and mu N terms.

The FP32C run takes exactly the FP64 run's accepted time steps (its error
tolerance is set so high that every first attempt is accepted, and each
step's dt is set to the FP64 step's), so the two fields are compared at the
same times and the difference is the storage/arithmetic precision alone.
(With its own step control the FP32C run accepts or rejects differently on
a few percent of steps, which shifts its clock; that difference is a time
offset, not an accuracy loss.)  A control run measures how much any
perturbation of this chaotic tangle grows: FP64 again, from the initial
field rounded to complex64 (a relative change of ~3e-8), on the same time
sequence.

A second report covers the 64^3 vortex ring (dx = 0.4, reflective, 2nd
order, RK4 dt = 0.01, 1000 steps): FP32C against FP64 after 1000 RK4 steps,
with the ring's radius and height from its core points.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import ic
from .config import RunConfig, xr_for_spacing
from .sim import Simulation


def workload(precision, n):
    return RunConfig(nx=n, ny=n, nz=n, xr=xr_for_spacing(0.5, n), bcs=1, order=4, scheme="rk45", eqn=4, g=1.0,
                     mu=1.0, tau=0.05, precision=precision)


def _observables(sim, cfg):
    psi = sim.download(np.complex128)
    return psi, {"t": sim.state()["t"], "energy_ref": sim.energy(), "hamiltonian": sim.hamiltonian(),
                 "norm": sim.norm(), "mass": sim.mass(), "linelength": sim.linelength()}


def core_points(pts):
    """(n, 3) coordinates of the finite points of a zeros pass (the
    reference's face formula gives NaN where one component's two face
    crossings share a coordinate)."""
    xyz = np.stack([pts["x"], pts["y"], pts["z"]], axis=1)
    return xyz[np.all(np.isfinite(xyz), axis=1)]


def core_offsets(pa, pb, dx):
    """Vortex-core position difference of two fields: the distance from
    every core point (zeros file, get_zeros + get_extra_zeros) of one field
    to the nearest core point of the other, both ways, in units of dx."""
    from scipy.spatial import cKDTree
    a, b = core_points(pa), core_points(pb)
    if len(a) == 0 or len(b) == 0:
        return {"points": [len(a), len(b)]}
    d = np.concatenate([cKDTree(b).query(a)[0], cKDTree(a).query(b)[0]]) / dx
    return {"points": [len(a), len(b)], "mean_dx": float(d.mean()), "p99_dx": float(np.quantile(d, 0.99)),
            "max_dx": float(d.max()), "frac_over_tenth_dx": float((d > 0.1).mean())}


def run(n=512, steps=1000, seed=20261202, log=None):
    log = log or (lambda *a: None)
    psi0 = ic.random_phase(workload("fp64", n), seed)
    out = {"grid": n, "steps": steps, "seed": seed}
    c64, c32 = workload("fp64", n), workload("fp32c", n)
    follow = 1e30  # tolerance of the runs that follow the FP64 run's steps
    with Simulation(c64, psi0) as s64, Simulation(c32.with_(eps=follow), psi0) as s32, \
            Simulation(c64.with_(eps=follow), psi0.astype(np.complex64).astype(np.complex128)) as sct:
        del psi0
        _, o64 = _observables(s64, c64)
        _, o32 = _observables(s32, c32)
        start = {"fp64": o64, "fp32c": o32}
        rejected = 0
        t0 = time.perf_counter()
        for i in range(steps):
            before = s64.state()
            s64.advance(1)
            after = s64.state()
            rejected += after["n_rejects"]
            dt = complex(after["t"] - before["t"], 0.0) if c64.real_time else complex(0.0, before["im_t"] - after["im_t"])
            for s in (s32, sct):
                s.set_state(dt, t=before["t"], im_t=before["im_t"], p=i)
                s.advance(1)
            if (i + 1) % 100 == 0:
                log("step %d, t = %.6f, %.1f s" % (i + 1, after["t"], time.perf_counter() - t0))
        p64, e64 = _observables(s64, c64)
        p32, e32 = _observables(s32, c32)
        z64, z32 = s64.zeros(), s32.zeros()
        pct = sct.download()
    out["psi_rel_l2_control_fp64_from_rounded_ic"] = float(np.linalg.norm((pct - p64).ravel()) /
                                                           np.linalg.norm(p64.ravel()))
    del pct
    for prec, e in (("fp64", e64), ("fp32c", e32)):
        o = start[prec]
        out[prec] = {"start": o, "end": e,
                     "energy_ref_drift": (e["energy_ref"] - o["energy_ref"]) / abs(o["energy_ref"]),
                     "hamiltonian_drift": (e["hamiltonian"] - o["hamiltonian"]) / abs(o["hamiltonian"]),
                     "norm_drift": (e["norm"] - o["norm"]) / o["norm"]}
    out["fp64_rejected_attempts"] = rejected
    out["psi_rel_l2"] = float(np.linalg.norm((p32 - p64).ravel()) / np.linalg.norm(p64.ravel()))
    out["linelength_rel_diff"] = (e32["linelength"] - e64["linelength"]) / e64["linelength"]
    out["core_position"] = core_offsets(z32, z64, c64.dx)
    out["hamiltonian_rel_diff"] = (e32["hamiltonian"] - e64["hamiltonian"]) / abs(e64["hamiltonian"])
    out["energy_ref_rel_diff"] = (e32["energy_ref"] - e64["energy_ref"]) / abs(e64["energy_ref"])
    return out


def ring_rk4(steps=1000):
    """64^3 ring, RK4: FP32C against FP64 after `steps` steps."""
    cfg = RunConfig(nx=64, ny=64, nz=64, xr=xr_for_spacing(0.4, 64), bcs=2, order=2, scheme="rk4", tau=0.01,
                    g=1.0, mu=1.0)
    psi0 = ic.vortex_ring(cfg, 4.0)
    res = {}
    for prec in ("fp64", "fp32c"):
        c = cfg.with_(precision=prec)
        with Simulation(c, psi0) as sim:
            e0 = sim.energy()
            sim.advance(steps)
            res[prec] = (sim.download(np.complex128), sim.energy(), e0, sim.linelength(), sim.zeros())
    a, b = res["fp32c"], res["fp64"]
    ring = {}
    for prec, r in res.items():
        q = core_points(r[4])
        ring[prec] = {"radius": float(np.hypot(q[:, 0], q[:, 1]).mean()), "z": float(q[:, 2].mean())}
    return {"steps": steps, "psi_rel_l2": float(np.linalg.norm((a[0] - b[0]).ravel()) / np.linalg.norm(b[0].ravel())),
            "energy_ref_rel_diff": (a[1] - b[1]) / abs(b[1]),
            "energy_ref_drift": {"fp64": (b[1] - b[2]) / abs(b[2]), "fp32c": (a[1] - a[2]) / abs(a[2])},
            "linelength": {"fp64": b[3], "fp32c": a[3]},
            "core_position": core_offsets(a[4], b[4], cfg.dx), "ring": ring}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20261202)
    a = ap.parse_args()
    res = {"ring_64_rk4": ring_rk4()}
    print("[accuracy] ring:", json.dumps(res["ring_64_rk4"]), file=sys.stderr, flush=True)
    res["random_phase_rk45"] = run(a.n, a.steps, a.seed,
                                   log=lambda *x: print("[accuracy]", *x, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
