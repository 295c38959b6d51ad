"""The Fortran drop-in at the line-pair workload's real size.

oracle/_ref/gpe_dropin_256_d is the drop-in harness (the reference's
parameters.f90 and error.f90 unchanged, the drop-in module solve of
gpe_amd/fortran/solve.f90, and a driver running gpe.f90's time loop) built
for a 256^3 grid in double precision.  It advances the 256^3 line pair
(dx = 0.25, reflective, 4th order, RK4 dt = 0.005) 30 solver() calls, once
with psi copied to and from the host every step (the plain drop-in) and once
device-resident (GPE_RESIDENT=1); both end states must equal the library's
own 30 steps bit for bit (whose 2000-step run equals the oracle's, see
tests/test_line_pair_full_gpu.py); the resident run uploads once and
downloads only where the driver reads psi (its I/O every 10 steps).  The
wall time of both runs is printed.
"""
import os
import sys
import time

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.state_io import read_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "gpe_dropin_256_d")
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in harness not built (make -C oracle ref)")]


def test_dropin_256_line_pair_plain_and_resident_equal_library(tmp_path):
    import make_line_pair_256 as mk
    from test_dropin import _run_mode
    from gpe_amd.sim import Simulation
    cfg = mk.workload()
    psi = ic.line_pair(cfg)
    n = 30
    print("[case] 256^3 drop-in, line pair, rk4, %d steps" % n, flush=True)
    with Simulation(cfg, psi) as sim:
        sim.advance(n)
        want = sim.download()
    fb = cfg.nx * cfg.ny * cfg.nz * 16
    # energy and the other I/O every 10 steps: the only steps after which
    # the driver reads psi (besides the final state)
    io = {"save_rate": 10, "save_rate2": 1e9, "save_rate3": 1e9, "p_save": 1e9}
    moved = {}
    for resident in (False, True):
        t0 = time.perf_counter()
        d, h2d, d2h = _run_mode(tmp_path, BIN, cfg, psi, n, io, resident)
        el = time.perf_counter() - t0
        got, p, t, _, _ = read_state(str(d / "out.dat"))
        print("[result] %s: %.2f s wall for %d steps, %d B to the device, %d B back" %
              ("resident" if resident else "plain", el, n, h2d, d2h), flush=True)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        moved[resident] = (h2d, d2h)
    # plain: the field goes up and comes back every step; resident: one
    # upload, and downloads only where the driver reads psi
    assert moved[False][0] >= n * fb and moved[False][1] >= n * fb
    assert moved[True][0] == fb
    assert moved[True][1] <= (n // 10 + 1) * fb
