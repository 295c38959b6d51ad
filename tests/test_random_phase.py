"""The random-phase initial state built on the device (gpe_sim_random_phase,
gpe_batch_add_random_phase) against the host generator ic.random_phase.

The mode list is formed on the host in the host generator's arithmetic:
the same flat indices, bit for bit, and the same phasors to one rounding
(cos/sin of the same phases).  The device then scatters it, runs hipFFT's
inverse transform and normalises to <|psi|^2> = 1: against the host's
pocketfft result that is a relative L2 difference at the FFTs' rounding
level (1e-12 is asserted), in every storage form, for cubic, non-cubic and
odd grids.  The batched instance starts from exactly the field
gpe_sim_random_phase builds."""
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

SHAPES = [(64, 48, 40, 0.5), (33, 31, 29, 0.7), (96, 96, 96, 0.5)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _cfg(nx, ny, nz, dx, **kw):
    return RunConfig(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2,
                     scheme="rk45", order=4, bcs=1, **kw)


def _host_modes(c, seed, kmax):
    kx = 2 * np.pi * np.fft.fftfreq(c.nx, d=c.dx)
    ky = 2 * np.pi * np.fft.fftfreq(c.ny, d=c.dy)
    kz = 2 * np.pi * np.fft.fftfreq(c.nz, d=c.dz)
    mx, my, mz = (np.nonzero(np.abs(k) <= kmax)[0] for k in (kx, ky, kz))
    K2 = kz[mz][:, None, None] ** 2 + ky[my][None, :, None] ** 2 + kx[mx][None, None, :] ** 2
    sel = np.nonzero(K2 <= kmax * kmax)
    flat = np.sort((mz[sel[0]].astype(np.int64) * c.ny + my[sel[1]]) * c.nx + mx[sel[2]])
    return flat, np.exp(1j * (2 * np.pi * ic.uniform(seed, flat.size)))


@pytest.mark.parametrize("shape", SHAPES + [(512, 512, 512, 0.5)])
def test_mode_list_matches_host_generator(shape):
    from gpe_amd.sim import random_phase_modes
    c = _cfg(*shape)
    idx, ph = random_phase_modes(c, 20261202, 0.5)
    want_idx, want_ph = _host_modes(c, 20261202, 0.5)
    assert len(idx) > 10
    np.testing.assert_array_equal(idx, want_idx)
    assert np.abs(ph - want_ph).max() <= 2.3e-16


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32"])
def test_device_field_matches_host(shape, precision):
    from gpe_amd.sim import Simulation
    c = _cfg(*shape, precision=precision)
    say("random phase on device", shape, precision)
    want = ic.random_phase(c, 7, kmax=0.5)
    with Simulation(c) as sim:
        sim.random_phase(7, 0.5)
        got = sim.download(np.complex128)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < (1e-6 if precision == "fp32" else 1e-12), err
    assert np.mean(np.abs(got) ** 2) == pytest.approx(1.0, rel=1e-6 if precision == "fp32" else 1e-12)


@pytest.mark.gpu
def test_batched_instance_starts_from_the_same_field():
    from gpe_amd.batch import Batch
    from gpe_amd.sim import Simulation
    c = _cfg(64, 48, 40, 0.5, precision="fp32c", tau=0.05, eps=1e-6)
    say("random phase batched")
    with Simulation(c) as sim:
        sim.random_phase(11)
        want = sim.download(np.complex128)
    with Batch() as b:
        i = b.add_random_phase(c, 11, 4)
        start = b.download(i)
        j = b.add(c, want, 4)
        st = b.run_to_completion(chunk=4)
        a, bb = b.download(i), b.download(j)
    np.testing.assert_array_equal(start, want)
    assert st[i] == st[j] == 1
    np.testing.assert_array_equal(a, bb)
