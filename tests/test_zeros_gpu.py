"""GPU vortex-core points (gpe_sim_zeros) against the oracle's restatement
of get_zeros, get_extra_zeros and get_re_im_zeros (io.f90:886-1393).

Every point is one cell's own arithmetic, one rounding per operation in the
reference's order, so the GPU lists equal the oracle's bit for bit: the
same sections, cells and coordinates, in the same file order, NaN where
the reference's face formula divides by zero.  Random fields hit every
test in every section and overflow the first device list (which then
grows); plain FP32 follows the single-precision oracle, the compensated
form the FP64 oracle on hi + lo."""
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


def _gpu(cfg, psi, which):
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi) as sim:
        return sim.zeros(which)


def _same(got, want):
    assert len(got) == len(want)
    for k in ("section", "cell"):
        np.testing.assert_array_equal(got[k], want[k])
    for k in ("x", "y", "z"):
        np.testing.assert_array_equal(got[k], want[k])  # NaN equals NaN here


@pytest.mark.parametrize("n", [(40, 33, 29), (96, 96, 96)])
def test_random_field_fp64(oracle, n):
    c = _cfg(n)
    rng = np.random.default_rng(21)
    psi = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    for w, name in enumerate(("zeros", "re_zeros", "im_zeros")):
        say("zeros fp64 random", n, name)
        got = _gpu(c, psi, name)
        want = oracle.zeros(c, psi, w)
        assert len(want) > 1000
        _same(got, want)


def test_ring_and_exact_zeros_fp64(oracle):
    c = _cfg((48, 44, 40), dx=0.4)
    psi = ic.vortex_ring(c, 4.0) * np.exp(0.3j)
    # exact zeros at some grid points: the == 0 branches
    flat = psi.reshape(-1)
    flat[::997] = 0.0
    flat.real[5::1009] = 0.0
    flat.imag[7::1013] = 0.0
    assert (psi == 0).sum() > 50
    for w, name in enumerate(("zeros", "re_zeros", "im_zeros")):
        say("zeros fp64 ring + exact zeros", name)
        _same(_gpu(c, psi, name), oracle.zeros(c, psi, w))


def test_fp32_and_fp32c(oracle):
    c = _cfg((40, 33, 29))
    rng = np.random.default_rng(22)
    psi = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    say("zeros fp32 / fp32c")
    p32 = psi.astype(np.complex64)
    c32 = c.with_(precision="fp32")
    for w, name in enumerate(("zeros", "re_zeros", "im_zeros")):
        _same(_gpu(c32, p32, name), oracle.zeros(c32, p32, w))
    hi = psi.astype(np.complex64)
    lo = (psi - hi.astype(np.complex128)).astype(np.complex64)
    recon = hi.astype(np.complex128) + lo.astype(np.complex128)
    for w, name in enumerate(("zeros", "re_zeros", "im_zeros")):
        _same(_gpu(c.with_(precision="fp32c"), psi, name), oracle.zeros(c, recon, w))


def test_zeros_files(tmp_path, oracle):
    from gpe_amd.sim import Simulation
    from gpe_amd.state_io import read_zeros
    c = _cfg((32, 30, 28), dx=0.4)
    psi = ic.vortex_ring(c, 3.0) * np.exp(0.5j)
    say("zeros files")
    with Simulation(c, psi) as sim:
        paths = sim.save_zeros(str(tmp_path / "proc0000"))
    assert [p.rsplit("/", 1)[1] for p in paths] == ["zeros0000000.dat", "re_zeros0000000.dat", "im_zeros0000000.dat"]
    for w, p in enumerate(paths):
        want = oracle.zeros(c, psi, w)
        secs = read_zeros(p)
        assert [len(s) for s in secs] == [int((want["section"] == q).sum()) for q in range(len(secs))]
        for q, s in enumerate(secs):
            m = want["section"] == q
            ref = np.stack([want["x"][m], want["y"][m], want["z"][m]], axis=1)
            np.testing.assert_allclose(s, ref, rtol=1e-8, atol=1e-12)
