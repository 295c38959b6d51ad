"""GPU vortex-line length (gpe_sim_linelength) against the oracle's
restatement of linelength (variables.f90:896-1162).

Per-cell arithmetic is the reference's, one rounding per operation, so in
FP64 every cell's contribution equals the oracle's; only the order of the
final sum differs (1e-12 relative).  Random fields exercise every branch,
including the cells that find one point and reuse the previous cell's
second point, and overflow the device point lists (which then grow)."""
import math
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _cfg(n, dx=0.5, precision="fp64"):
    nx, ny, nz = n
    return RunConfig(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2,
                     precision=precision)


def _gpu_ll(cfg, psi):
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi) as sim:
        return sim.linelength()


@pytest.mark.parametrize("n", [(40, 33, 29), (96, 96, 96)])
def test_random_field_fp64(oracle, n):
    c = _cfg(n)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    say("linelength fp64 random", n)
    got = _gpu_ll(c, psi)
    want = oracle.linelength(c, psi)
    assert got == pytest.approx(want, rel=1e-12)


def test_ring_and_line_fp64(oracle):
    c = _cfg((48, 44, 40), dx=0.4)
    psi = ic.vortex_ring(c, 4.0)
    say("linelength ring")
    got = _gpu_ll(c, psi)
    assert got == pytest.approx(oracle.linelength(c, psi), rel=1e-12)
    assert got == pytest.approx(2 * math.pi * 4.0, rel=0.05)
    X = c.coords()[0][None, None, :]
    Y = c.coords()[1][None, :, None]
    line = np.ascontiguousarray(np.broadcast_to((X - 0.13) + 1j * (Y + 0.07), c.shape))
    assert _gpu_ll(c, line) == pytest.approx((c.nz - 2) * c.dx, rel=1e-15)


def test_random_field_fp32_and_fp32c(oracle):
    c = _cfg((40, 33, 29))
    rng = np.random.default_rng(12)
    psi = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    say("linelength fp32 / fp32c")
    # plain FP32: the reference built in single precision
    p32 = psi.astype(np.complex64)
    got32 = _gpu_ll(c.with_(precision="fp32"), p32)
    assert got32 == pytest.approx(oracle.linelength(c.with_(precision="fp32"), p32), rel=1e-5)
    # compensated: hi + lo in double equals the FP64 oracle on that field
    hi = psi.astype(np.complex64)
    lo = (psi - hi.astype(np.complex128)).astype(np.complex64)
    recon = hi.astype(np.complex128) + lo.astype(np.complex128)
    gotc = _gpu_ll(c.with_(precision="fp32c"), psi)
    assert gotc == pytest.approx(oracle.linelength(c, recon), rel=1e-12)
