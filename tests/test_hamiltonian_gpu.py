"""The eqn-4 Hamiltonian on the device (gpe_sim_hamiltonian).

Pinned two ways: against the oracle's energy functional plus host sums of
the same field, and by the physics, since H is what the dynamics conserve.
Over 200 RK4 steps in a trap with g != 1 and mu != 0, H changes by less
than 1e-9 while its parts trade far more.  Combinations with any
term's weight doubled do not stay constant."""
import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _cfg(precision="fp64"):
    n = 32
    d = 0.4
    return RunConfig(nx=n, ny=n, nz=n, xr=d * (n - 1) / 2, yr=d * (n - 1) / 2, zr=d * (n - 1) / 2, eqn=4, g=0.7,
                     mu=0.3, omx=0.2, omy=0.25, omz=0.3, bcs=1, order=4, scheme="rk4", tau=0.0005,
                     precision=precision)


def _parts(cfg, psi):
    """(2K, quartic sum, trap sum, |psi|^2 sum), each times dV, on the host."""
    psi = np.asarray(psi, dtype=np.complex128)
    dv = cfg.dx * cfg.dy * cfg.dz
    x, y, z = cfg.coords()
    V = 0.5 * ((cfg.omx ** 2 * x[None, None, :] ** 2 + cfg.omy ** 2 * y[None, :, None] ** 2)
               + cfg.omz ** 2 * z[:, None, None] ** 2)
    m2 = psi.real ** 2 + psi.imag ** 2
    s4 = float((m2 * m2).sum()) * dv
    E = pyoracle.energy(cfg, psi)
    return E - 0.5 * s4, s4, float((V * m2).sum()) * dv, float(m2.sum()) * dv


def _field(cfg):
    rng = np.random.default_rng(11)
    psi = ic.vortex_ring(cfg, 2.0) * np.exp(0.3j * rng.standard_normal(cfg.shape) * 0.1)
    return psi * (1 + 0.05 * rng.standard_normal(cfg.shape))


@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32"])
def test_hamiltonian_matches_oracle_sums(precision):
    from gpe_amd.sim import Simulation
    cfg = _cfg(precision)
    print("[case] hamiltonian", precision, flush=True)
    psi0 = _field(cfg)
    if precision == "fp32":
        psi0 = psi0.astype(np.complex64)
    with Simulation(cfg, psi0) as sim:
        sim.advance(3)
        H = sim.hamiltonian()
        psi = sim.download()
    k2, s4, sv, s2 = _parts(cfg, psi)
    want = 0.5 * k2 + 0.5 * cfg.g * s4 + sv - cfg.mu * s2
    print("[result] hamiltonian %s rel err %.3e" % (precision, abs(H - want) / abs(want)), flush=True)
    # FP32C forms the Laplacian in single precision, FP32 also sums energy() in it
    assert H == pytest.approx(want, rel={"fp64": 1e-12, "fp32c": 1e-5, "fp32": 1e-4}[precision])


def test_hamiltonian_is_conserved_and_its_terms_are_not():
    from gpe_amd.sim import Simulation
    cfg = _cfg()
    print("[case] hamiltonian conservation", flush=True)
    with Simulation(cfg, _field(cfg)) as sim:
        H0 = sim.hamiltonian()
        p0 = _parts(cfg, sim.download())
        sim.advance(200)
        H1 = sim.hamiltonian()
        p1 = _parts(cfg, sim.download())

    def comb(p, wk=1.0, wg=1.0, wv=1.0):
        return 0.5 * wk * p[0] + 0.5 * wg * cfg.g * p[1] + wv * p[2] - cfg.mu * p[3]

    drift = abs(H1 - H0) / abs(H0)
    assert drift < 1e-9
    assert comb(p0) == pytest.approx(H0, rel=1e-12)
    # the parts exchange energy, so a wrong weight on any of them shows up
    for w in ({"wk": 2.0}, {"wg": 2.0}, {"wv": 2.0}):
        assert abs(comb(p1, **w) - comb(p0, **w)) / abs(comb(p0, **w)) > 1e3 * max(drift, 1e-15)


def test_hamiltonian_needs_eqn_4():
    from gpe_amd import _lib
    from gpe_amd.sim import Simulation
    cfg = _cfg()
    cfg.eqn = 2
    with Simulation(cfg, _field(cfg)) as sim:
        with pytest.raises(_lib.GpeError):
            sim.hamiltonian()
