"""Accuracy of the compensated FP32 form at the headline size, in the
driver-run suite: the 512^3 random-phase field (dx = 0.5, periodic,
4th-order FD, RKF45 with tolerance 1e-6) for 1000 accepted steps in FP64 and
in FP32C from the same initial state (gpe_amd.accuracy.run).

The FP32C run takes the FP64 run's accepted steps (each step's dt set to the
FP64 step's), so the two fields are compared at the same times and the
difference is the storage/arithmetic precision alone; a control FP64 run
from the initial field rounded to complex64 shows how much the tangle grows
a perturbation of that size.  Bars: psi relative L2 <= 1e-5, vortex-line
length and the eqn-4 Hamiltonian within 1e-5, mean vortex-core offset below
1e-3 dx; FP64 norm drift and Hamiltonian drift are printed.
"""
import json
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_fp32c_against_fp64_random_phase_512_1000_rk45_steps(gpu):
    from gpe_amd import accuracy
    print("[case] 512^3 random phase, rk45, 1000 accepted steps, fp32c vs fp64", flush=True)
    out = accuracy.run(512, 1000, log=lambda *a: print("[case]", *a, flush=True))
    brief = {k: out[k] for k in ("psi_rel_l2", "psi_rel_l2_control_fp64_from_rounded_ic", "linelength_rel_diff",
                                 "hamiltonian_rel_diff", "energy_ref_rel_diff", "fp64_rejected_attempts",
                                 "core_position")}
    brief["t_end"] = out["fp64"]["end"]["t"]
    brief["hamiltonian_drift"] = {p: out[p]["hamiltonian_drift"] for p in ("fp64", "fp32c")}
    print("[result]", json.dumps(brief), flush=True)
    sys.stdout.flush()
    assert out["psi_rel_l2"] < 1e-5
    assert abs(out["linelength_rel_diff"]) < 1e-5
    assert abs(out["hamiltonian_rel_diff"]) < 1e-5
    assert out["core_position"]["mean_dx"] < 1e-3
    assert out["core_position"]["points"][0] > 1000
