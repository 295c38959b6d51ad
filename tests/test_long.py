"""Full-length parity runs (minutes of oracle time on 16 host threads);
skipped unless GPE_LONG=1.

The 256^3 line pair in a trap (dx = 0.25, 4th-order FD, reflective, RK4
dt = 0.005) for its full 2000 steps: FP64 psi, norm and energy within
1e-12 of the oracle, and the vortex-core points of the two lines within
1e-9 dx of the oracle's."""
import os
import sys
import time

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig, xr_for_spacing

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(os.environ.get("GPE_LONG") != "1",
                                                  reason="full-length run; set GPE_LONG=1")]
NT = min(16, os.cpu_count() or 1)


def test_line_pair_256cube_2000_steps(oracle):
    from gpe_amd.sim import Simulation
    n = int(os.environ.get("GPE_LONG_STEPS", "2000"))
    cfg = RunConfig(nx=256, ny=256, nz=256, xr=xr_for_spacing(0.25, 256), bcs=2, order=4, scheme="rk4",
                    tau=0.005, g=1.0, mu=1.0, omx=0.1, omy=0.1, omz=float(np.sqrt(0.02)))
    psi = ic.line_pair(cfg)
    print("[case] 256^3 line pair rk4 x%d" % n, flush=True)
    with Simulation(cfg, psi) as sim:
        t0 = time.perf_counter()
        gn = sim.advance(n, record_norm=True)
        got = sim.download()
        ge = sim.energy()
        gpu_s = time.perf_counter() - t0
    print("[case] gpu %.1f s; oracle running" % gpu_s, flush=True)
    t0 = time.perf_counter()
    st = oracle.make_state(cfg.dt)
    want, wn = psi, None
    done = 0
    while done < n:
        m = min(50, n - done)
        want, wn = oracle.run(cfg, st, want, m, NT, with_norms=True)
        done += m
        print("[case] oracle %d/%d steps, %.0f s" % (done, n, time.perf_counter() - t0), flush=True)
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    we = oracle.energy(cfg, want, NT)
    print("[result] psi rel L2 %.3e, norm rel %.3e, energy rel %.3e" %
          (err, abs(gn[-1] - wn[-1]) / wn[-1], abs(ge - we) / abs(we)), flush=True)
    sys.stdout.flush()
    assert err < 1e-12
    assert abs(gn[-1] - wn[-1]) < 1e-12 * wn[-1]
    assert ge == pytest.approx(we, rel=1e-12)
    # vortex-core positions of the two lines (the reference's zeros file)
    from gpe_amd.accuracy import core_offsets
    with Simulation(cfg, got) as sim:
        zg = sim.zeros()
    d = core_offsets(zg, oracle.zeros(cfg, want, 0), cfg.dx)
    print("[result] core offsets %s" % d, flush=True)
    assert d["points"][0] == d["points"][1] > 100
    assert d["max_dx"] < 1e-9
