"""A seeded sweep of FP64 parity over launch configurations.

Every case draws a grid shape (ragged in x against the 64-wide tiles, odd
and even y/z), order, boundary condition, scheme, tile variant and z-chunk
length, runs three solver() calls on the GPU and on the oracle, and asks
for 1e-12 relative L2 in psi and equal t and dt.  FP64 results do not
depend on the launch configuration, so any case outside the bar means an
indexing, chunking or synchronisation fault in one configuration.  The draws
are fixed by the seed; the sweep is ordinary regression coverage."""
import os
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def _cases(n=40, seed=20261202):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n):
        nx, ny, nz = (int(v) for v in rng.integers([18, 14, 12], [140, 70, 60]))
        out.append(dict(case=c, shape=(nx, ny, nz), order=int(rng.choice([2, 4])), bcs=int(rng.choice([1, 2])),
                        scheme=str(rng.choice(["euler", "rk2", "rk4", "rk45"])), variant=int(rng.integers(0, 5)),
                        zchunk=int(rng.choice([0, 8, 16, 24, 64]))))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: "c%d" % c["case"])
def test_fp64_parity_across_launch_configurations(oracle, case):
    from gpe_amd import _lib
    from gpe_amd.sim import Simulation
    nx, ny, nz = case["shape"]
    dx = 0.4
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2, eqn=4,
                    g=1.0, mu=1.0, omx=0.2, omy=0.25, omz=0.3, order=case["order"], bcs=case["bcs"],
                    scheme=case["scheme"], tau=0.002 if case["scheme"] == "euler" else 0.01)
    psi = ic.vortex_ring(cfg, 0.2 * nx * dx) * (1 + 0.03 * np.random.default_rng(case["case"]).standard_normal(cfg.shape))
    print("[case]", case, flush=True)
    sys.stdout.flush()
    st = oracle.make_state(cfg.dt)
    want = psi
    for _ in range(3):
        want = oracle.solver(cfg, st, want, NT)
    with Simulation(cfg, psi) as sim:
        for k in _lib.STAGE_KINDS:
            sim.tune_set(k, case["variant"], case["zchunk"])
        sim.advance(3)
        got = sim.download()
        s = sim.state()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12
    # rk45 forms dt from the error estimate, a difference of nearly equal
    # Cash-Karp sums: the one-ulp hypot differences of |psi^2| (device math
    # library against glibc) reach it at ~1e-13, so t and dt get 1e-12
    assert s["t"] == pytest.approx(st.t, rel=1e-12, abs=1e-15)
    assert s["dt"].real == pytest.approx(st.dt_re, rel=1e-12)
