"""FFT-side diagnostics on the GPU against the oracle (oracle/spectral.py):
condensed_particles, spectrum and filtered_surface (io.f90:304-395,
524-716), for every storage form, on ragged and even grids.  The device
transforms in FP64 whatever the storage form, so each is compared with the
oracle on the field the GPU holds (downloaded in complex128)."""
import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig
from oracle import pyoracle
from oracle import spectral as S

pytestmark = pytest.mark.gpu


def _cfg(shape, precision, bcs=1):
    nz, ny, nx = shape
    d = 0.5
    return RunConfig(nx=nx, ny=ny, nz=nz, xr=d * (nx - 1) / 2, yr=d * (ny - 1) / 2, zr=d * (nz - 1) / 2, eqn=4,
                     g=1.0, mu=1.0, bcs=bcs, order=4, scheme="rk4", tau=0.01, precision=precision)


def _field(cfg, seed):
    rng = np.random.default_rng(seed)
    psi = ic.vortex_ring(cfg, 0.25 * cfg.nx * cfg.dx)
    return psi * (1 + 0.05 * rng.standard_normal(cfg.shape)) + 0.05 * rng.standard_normal(cfg.shape) * 1j


CASES = [((18, 20, 24), "fp64"), ((18, 20, 24), "fp32c"), ((18, 20, 24), "fp32"), ((32, 32, 32), "fp64"),
         ((40, 36, 48), "fp32c")]


@pytest.mark.parametrize("shape,precision", CASES, ids=lambda v: str(v))
def test_spectral_diagnostics_match_oracle(shape, precision):
    from gpe_amd.sim import Simulation
    cfg = _cfg(shape, precision)
    print("[case] spectral", shape, precision, flush=True)
    psi0 = _field(cfg, 7)
    if precision == "fp32":
        psi0 = psi0.astype(np.complex64)
    with Simulation(cfg, psi0) as sim:
        sim.advance(2)
        held = sim.download().astype(np.complex128)
        rec = sim.condensed_particles()
        nh, eta, idf = sim.spectrum()
        f, ln = sim.filtered(0.6 * cfg.nx / 2, m=1.0, with_length=True)
        f2 = sim.filtered(3.0, m=2.0)
        f32 = sim.filtered(3.0, m=2.0, single=True)
        st = sim.state()
        after = sim.download().astype(np.complex128)
    assert np.array_equal(after, held)  # the resident field is untouched

    a = S.forward(held)
    want = S.condensed_particles(cfg.shape, cfg.dx, cfg.dy, cfg.dz, cfg.xr, cfg.yr, cfg.zr, held,
                                 pyoracle.mass(cfg, held), st["t"], st["im_t"])
    for k, v in want.items():
        assert rec[k] == pytest.approx(v, rel=1e-12, abs=1e-300), k
    wn, we, wf = S.spectrum(a)
    assert list(nh) == list(wn)
    assert np.allclose(eta, we, rtol=1e-12, atol=0)
    assert np.allclose(idf, wf, rtol=1e-12, atol=0)
    assert len(idf) == cfg.nx // 2
    ff = S.filtered(a, 0.6 * cfg.nx / 2, 1.0)
    scale = np.max(np.abs(ff))
    assert np.max(np.abs(f - ff)) < 1e-13 * scale
    assert np.max(np.abs(f2 - S.filtered(a, 3.0, 2.0))) < 1e-13 * scale
    assert f32.dtype == np.complex64 and np.array_equal(f32, f2.astype(np.complex64))
    assert ln == pytest.approx(pyoracle.linelength(cfg, ff), rel=1e-10)


def test_uniform_field_is_all_condensate():
    from gpe_amd.sim import Simulation
    cfg = _cfg((16, 18, 20), "fp64")
    c = 0.8 - 0.3j
    print("[case] spectral uniform", flush=True)
    with Simulation(cfg, np.full(cfg.shape, c)) as sim:
        rec = sim.condensed_particles()
        nh, eta, idf = sim.spectrum()
    V = float(np.prod(cfg.shape))
    dV = cfg.dx * cfg.dy * cfg.dz
    assert rec["n0"] == pytest.approx(V * abs(c) ** 2, rel=1e-13)
    assert rec["mass"] == pytest.approx(V * abs(c) ** 2 * dV, rel=1e-13)
    assert np.max(idf) < 1e-20 * V and np.max(eta) < 1e-20 * V


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_parseval_and_identity_filter_at_256cube(precision):
    # size-independent properties on a random-phase field whose modes all
    # lie within |k| <= nx/2: F(nx/2) + n0 = sum |psi|^2 = M / dV, and a
    # filter with an infinite cut-off returns the field
    from gpe_amd.sim import Simulation
    n = 256
    cfg = _cfg((n, n, n), precision)
    print("[case] spectral 256^3", precision, flush=True)
    with Simulation(cfg, None) as sim:
        sim.random_phase(seed=20261202, kmax=0.5)
        rec = sim.condensed_particles()
        nh, eta, idf = sim.spectrum()
        held = sim.download()
        f = sim.filtered(1e12)
    dV = cfg.dx * cfg.dy * cfg.dz
    assert idf[-1] + rec["n0"] == pytest.approx(rec["mass"] / dV, rel=1e-12)
    assert np.all(np.diff(idf) >= 0)
    assert nh.sum() <= n ** 3 - 1
    assert np.max(np.abs(f - held)) < 1e-13 * np.max(np.abs(held))
