"""The 256^3 line pair for its full 2000 RK4 steps against the oracle's
stored run (tests/golden/line_pair_256_rk4_2000.npz, written by
tests/golden/make_line_pair_256.py: the oracle takes about an hour on 8 host
threads, the GPU about two seconds).

Workload: 256^3 grid, dx = 0.25, reflective walls, 4th-order FD, RK4 with
dt = 0.005 (solve.f90:109-154) for 2000 solver() calls (the time loop of
gpe.f90:136), eqn 4 with g = mu = 1 in the trap omx = omy = 0.1,
omz = sqrt(0.02); two straight vortex lines along z at (+-3, 0) in the
Thomas-Fermi background sqrt(max(0, 10 - V)).

Checked against the fixture:
  * get_norm after every step (gpe.f90:243-249): within 1e-12 relative;
  * energy() every 250 steps: within 1e-12 relative;
  * psi at 2^18 sites spread over the whole box: bit for bit (the FP64
    kernels round every operation as the oracle does), so also within the
    1e-12 relative L2 bar;
  * t after the run;
  * the zeros-file core points of the final field (io.f90:886-1393): the
    same count, and the points of every 8th plane equal bit for bit.
"""
import os
import sys
import time

import numpy as np
import pytest

from gpe_amd import ic

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "line_pair_256_rk4_2000.npz")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def test_line_pair_256cube_2000_steps_against_stored_oracle_run(gpu):
    import make_line_pair_256 as mk
    from gpe_amd.sim import Simulation
    ref = np.load(FIXTURE)
    cfg = mk.workload()
    psi = ic.line_pair(cfg)
    print("[case] 256^3 line pair, rk4, %d steps, fp64" % mk.STEPS, flush=True)
    norms, energy = [], []
    t0 = time.perf_counter()
    with Simulation(cfg, psi) as sim:
        energy.append(sim.energy())
        done = 0
        while done < mk.STEPS:
            norms.append(sim.advance(mk.ENERGY_EVERY, record_norm=True))
            energy.append(sim.energy())
            done += mk.ENERGY_EVERY
        got = sim.download()
        t = sim.state()["t"]
        zg = sim.zeros()
    print("[case] gpu %.1f s" % (time.perf_counter() - t0), flush=True)
    norms = np.concatenate(norms)
    energy = np.array(energy)
    sample = got.reshape(-1)[mk.sample_sites()]
    want = ref["sample"]
    rel = float(np.linalg.norm(sample - want) / np.linalg.norm(want))
    nerr = float(np.max(np.abs(norms - ref["norm"]) / ref["norm"]))
    eerr = float(np.max(np.abs(energy - ref["energy"]) / np.abs(ref["energy"])))
    print("[result] sampled psi rel L2 %.3e (bitwise equal sites %.6f), norm history max rel %.3e, "
          "energy max rel %.3e, t %.17g" % (rel, np.mean(sample == want), nerr, eerr, t), flush=True)
    assert rel < 1e-12
    assert nerr < 1e-12
    assert eerr < 1e-12
    assert np.array_equal(sample.view(np.uint64), want.view(np.uint64))
    assert t == float(ref["t"][0])
    zs = mk.core_subset(zg)
    print("[result] core points %d (oracle %d), stored plane subset %d" %
          (len(zg), int(ref["core_count"][0]), len(zs)), flush=True)
    assert len(zg) == int(ref["core_count"][0])
    assert np.array_equal(zs["section"], ref["core_section"]) and np.array_equal(zs["cell"], ref["core_cell"])
    for c in "xyz":
        assert np.array_equal(zs[c], ref["core_" + c], equal_nan=True)
