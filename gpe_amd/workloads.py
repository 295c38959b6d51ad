"""Synthetic workloads of the batch and serving forms (host side).

batch_mix(): 64 independent simulations with seeded uneven cubic grids
N in {32, 48, 64, 96, 128}, dx = 0.4, 2nd-order FD, reflective walls, RK4
with dt = 0.01 for 500 steps, complex64 storage with the compensated
update; each in a trap omx = omy ~ U[0.05, 0.2], omz = 2*omx, starting from
a centred vortex ring of radius R ~ U[0.2, 0.35]*N*dx times
sqrt(max(0, 5 - V)).

request_stream(): Poisson arrivals at a fixed rate; each request draws
N in {48, 64, ..., 256}, periodic or reflective walls, order 2 or 4, a step
count in [200, 5000], RK4 (dt = 0.01) or RKF45 (tolerance 1e-6), FP32C or
FP64, and a ring, line-pair or random-phase initial state (the parameters
of the 64^3 ring, 256^3 line-pair and 512^3 random-phase cases), seeded per
request.

All draws come from SplitMix64 (ic.uniform) so a seed names the same
workload everywhere.
"""
from dataclasses import dataclass

import numpy as np

from . import ic
from .config import RunConfig

DX = 0.4


def _cube(n, dx, **kw):
    xr = dx * (n - 1) / 2.0
    return RunConfig(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, **kw)


@dataclass
class Instance:
    cfg: RunConfig
    steps: int
    radius: float
    seed: int

    def add_to(self, batch, device_ic=True):
        """Admit this request into a Batch: the random-phase state is built
        on the device (no host FFT, no upload) unless device_ic is False;
        the others are built on the host and uploaded.  Returns the
        instance id, or None when it does not fit the HBM budget."""
        if device_ic and self.ic_kind == "random_phase":
            return batch.add_random_phase(self.cfg, self.seed, self.steps, kmax=0.5)
        return batch.add(self.cfg, self.initial_state(), self.steps)

    def initial_state(self):
        return ic.trapped_ring(self.cfg, self.radius, mu_ic=5.0)


def batch_mix(seed=20261202, count=64, steps=500, precision="fp32c"):
    sizes = (32, 48, 64, 96, 128)
    out = []
    for i in range(count):
        u = ic.uniform(seed + i, 3)
        n = sizes[min(int(u[0] * len(sizes)), len(sizes) - 1)]
        om = 0.05 + 0.15 * u[1]
        cfg = _cube(n, DX, scheme="rk4", order=2, bcs=2, tau=0.01, eqn=4, g=1.0, mu=1.0, omx=om, omy=om,
                    omz=2 * om, precision=precision)
        radius = (0.2 + 0.15 * u[2]) * n * DX
        out.append(Instance(cfg, steps, radius, seed + i))
    return out


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the stream starts
    cfg: RunConfig
    steps: int
    ic_kind: str        # ring | line_pair | random_phase
    seed: int

    def add_to(self, batch, device_ic=True):
        """Admit this request into a Batch: the random-phase state is built
        on the device (no host FFT, no upload) unless device_ic is False;
        the others are built on the host and uploaded.  Returns the
        instance id, or None when it does not fit the HBM budget."""
        if device_ic and self.ic_kind == "random_phase":
            return batch.add_random_phase(self.cfg, self.seed, self.steps, kmax=0.5)
        return batch.add(self.cfg, self.initial_state(), self.steps)

    def initial_state(self):
        c = self.cfg
        if self.ic_kind == "ring":
            # the 64^3 ring case: R = 4 in the z = 0 plane
            return ic.vortex_ring(c, 4.0)
        if self.ic_kind == "line_pair":
            return ic.line_pair(c, sep=3.0, mu_ic=10.0)
        return ic.random_phase(c, self.seed, kmax=0.5, workers=4)


# grid spacing and trap of the case each initial state comes from
_IC_CASE = {
    "ring": (0.4, (0.0, 0.0, 0.0)),                          # 64^3 ring, V = 0
    "line_pair": (0.25, (0.1, 0.1, float(np.sqrt(0.02)))),   # 256^3 line pair in a trap
    "random_phase": (0.5, (0.0, 0.0, 0.0)),                  # 512^3 random phase, V = 0
}


def _request_cfg(n, bcs, order, scheme, precision, kind):
    dx, (ox, oy, oz) = _IC_CASE[kind]
    # rk45 starts from dt = 0.01 and adapts (tolerance eps = 1e-6)
    return _cube(n, dx, scheme=scheme, order=order, bcs=bcs, tau=0.01, eqn=4, g=1.0, mu=1.0, omx=ox, omy=oy,
                 omz=oz, eps=1e-6, precision=precision)


def request_stream(rate=40.0, duration=120.0, seed=20261202):
    """Requests of a Poisson stream of `rate` per second over `duration` s."""
    reqs = []
    t = 0.0
    i = 0
    while True:
        u = ic.uniform(seed + 7919 * (i + 1), 9)
        t += -np.log(1.0 - u[0]) / rate
        if t >= duration:
            break
        n = 48 + 16 * min(int(u[1] * 14), 13)
        bcs = 1 if u[2] < 0.5 else 2
        order = 2 if u[3] < 0.5 else 4
        steps = 200 + min(int(u[4] * 4801), 4800)
        scheme = "rk4" if u[5] < 0.5 else "rk45"
        precision = "fp32c" if u[6] < 0.5 else "fp64"
        kind = ("ring", "line_pair", "random_phase")[min(int(u[7] * 3), 2)]
        cfg = _request_cfg(n, bcs, order, scheme, precision, kind)
        reqs.append(Request(i, t, cfg, steps, kind, seed + i))
        i += 1
    return reqs
