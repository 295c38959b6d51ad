"""GPU imprint_vortex_line (gpe_sim_imprint_vortex_line) against the
oracle's restatement of variables.f90:1166-1180.

Per site the arithmetic is the reference's; sin, cos, atan2, exp and pow
come from the device math library rather than glibc, so values agree to a
few ulps, not bit for bit (1e-13 of the field's scale in double precision,
1e-5 in single).  Every direction, with and without the core amplitude, in
every storage form; then the imaginary-time loop of gpe.f90:230-272 (solver,
get_norm, renormalise, imprint) for a few steps against the oracle."""
import os
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)
LINE = dict(x0=0.6, y0=-0.4, z0=0.2, amp1=0.5, amp2=0.3, ll=3.2, sgn=1.0)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _cfg(**kw):
    base = dict(nx=70, ny=21, nz=19, xr=7.0, yr=4.0, zr=3.6, eqn=4, g=1.0, mu=1.0, omx=0.3, omy=0.2, omz=0.4,
                tau=0.005)
    base.update(kw)
    return RunConfig(**base)


def _field(cfg, seed=4):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.standard_normal(cfg.shape) + 1j * rng.standard_normal(cfg.shape))


def _gpu(cfg, psi, line, scal=1.0):
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi) as sim:
        sim.imprint_vortex_line(line, scal)
        return sim.download(np.complex64 if cfg.precision == "fp32" else np.complex128)


@pytest.mark.parametrize("direction", ["x", "y", "z"])
@pytest.mark.parametrize("phase_only", [True, False])
def test_fp64_matches_oracle(oracle, direction, phase_only):
    c = _cfg()
    psi = _field(c)
    line = dict(LINE, dir=direction, imprint_phase=phase_only)
    say("imprint fp64", direction, phase_only)
    got = _gpu(c, psi, line, scal=1.3)
    want = oracle.imprint_vortex_line(c, psi, line, scal=1.3)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_fp32_and_fp32c(oracle):
    c = _cfg()
    psi = _field(c, 6)
    line = dict(LINE, dir="z", imprint_phase=False)
    say("imprint fp32 / fp32c")
    p32 = psi.astype(np.complex64)
    got = _gpu(c.with_(precision="fp32"), p32, line)
    want = oracle.imprint_vortex_line(c.with_(precision="fp32"), p32, line)
    assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()
    hi = psi.astype(np.complex64)
    recon = hi.astype(np.complex128) + (psi - hi.astype(np.complex128)).astype(np.complex64).astype(np.complex128)
    got = _gpu(c.with_(precision="fp32c"), psi, line)
    want = oracle.imprint_vortex_line(c, recon, line)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_zero_wavelength_with_disturbance_is_refused():
    from gpe_amd import _lib
    c = _cfg()
    with pytest.raises(_lib.GpeError, match="wavelength is zero"):
        _gpu(c, _field(c), dict(amp1=0.5, ll=0.0, dir="z"))


def test_imaginary_time_loop_with_imprint(oracle):
    """The trap_vortex setting: rk2 in imaginary time, renormalised to the
    norm and imprinted with a straight line after every step."""
    from gpe_amd.sim import Simulation
    c = _cfg(scheme="rk2", order=2, bcs=2, real_time=False, tau=0.001, g=5.0, mu=8.0)
    psi = np.ascontiguousarray(np.sqrt(np.maximum(0.0, 8.0 - ic.trap_potential(c))) + 0j)
    line = dict(x0=-1.5, y0=0.0, z0=0.0, sgn=1.0, dir="z", imprint_phase=True)  # vl1 of parameters.in
    say("imaginary time rk2 + renormalise + imprint")
    st = oracle.make_state(c.dt)
    cur = psi
    with Simulation(c, psi) as sim:
        for _ in range(5):
            cur = oracle.solver(c, st, cur, NT)
            nrm = oracle.norm(c, cur)
            cur = oracle.renormalise(c, cur, 1.0, nrm)
            cur = oracle.imprint_vortex_line(c, cur, line)
            sim.advance(1)
            gn = sim.norm()
            assert gn == pytest.approx(nrm, rel=1e-12)
            sim.renormalise(1.0, gn)
            sim.imprint_vortex_line(line)
        got = sim.download()
    assert np.linalg.norm(got - cur) / np.linalg.norm(cur) < 1e-12
