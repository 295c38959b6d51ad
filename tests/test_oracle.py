"""Pins of the C oracle (CPU): reference vectors, closed forms, invariants.

Each check is chosen so that a plausible slip in the oracle -- a dropped
term, a wrong sign or index, a transposed axis, a wrong tableau entry --
fails it:
  * stencils and boundary maps: bit-for-bit against the reference's own
    derivs module (golden vectors from tests/golden/make_golden.py, and the
    live harness when oracle/_ref is built);
  * Laplacian: plane-wave eigenvalues (periodic) and cosine-mode eigenvalues
    (reflective) in closed form, on a non-cubic grid;
  * get_rhs / integrators: a uniform state is stationary bit for bit; a
    plane wave reduces every integrator to the same integrator on a scalar
    ODE, and for g = 0 to the closed-form amplification factor R(z)^N;
  * Cash-Karp: convergence orders 5 (solution) and 4 (embedded estimate)
    against the exact plane-wave solution; step-size control decisions;
  * get_norm: box volume for |psi| = 1 and exactness on cubics (the
    extended closed rule); energy and mass of a plane wave in closed form.
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from gpe_amd.config import RunConfig
from gpe_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_derivs.npz")


def _cfg(**kw):
    base = dict(nx=12, ny=10, nz=9, xr=3.3, yr=2.1, zr=1.7, bcs=1, order=2, eqn=4, g=1.0, mu=1.0)
    base.update(kw)
    return RunConfig(**base)


# --------------------------------------------------------------------------
# reference vectors

@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("bcs", [1, 2])
@pytest.mark.parametrize("order", [2, 4])
def test_stencils_match_reference_golden(oracle, prec, bcs, order):
    g = np.load(GOLD)
    key = "%s_b%d_o%d" % (prec, bcs, order)
    cfg = _cfg(bcs=bcs, order=order, precision=prec)
    f = g[key + "_in"]
    lap = oracle.laplacian(cfg, f)
    fx = oracle.deriv_x(cfg, f)
    assert np.array_equal(lap.view(np.uint8), g[key + "_lap"].view(np.uint8)), "laplacian differs from reference"
    assert np.array_equal(fx.view(np.uint8), g[key + "_dx"].view(np.uint8)), "deriv_x differs from reference"


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_derivs_d")),
                    reason="reference harness not built (make -C oracle ref)")
@pytest.mark.parametrize("bcs", [1, 2])
@pytest.mark.parametrize("order", [2, 4])
def test_stencils_match_reference_live(oracle, bcs, order):
    rng = np.random.default_rng(bcs * 10 + order)
    cfg = _cfg(bcs=bcs, order=order)
    f = rng.standard_normal(cfg.shape) * 10.0 ** rng.uniform(-3, 3, cfg.shape) + 1j * rng.standard_normal(cfg.shape)
    with tempfile.TemporaryDirectory() as d:
        f.tofile(os.path.join(d, "in.bin"))
        subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_derivs_d"), str(bcs), str(order),
                               repr(cfg.dx), repr(cfg.dy), repr(cfg.dz), os.path.join(d, "in.bin"),
                               os.path.join(d, "out.bin")])
        res = np.fromfile(os.path.join(d, "out.bin"), dtype=np.complex128)
    assert np.array_equal(oracle.laplacian(cfg, f), res[: f.size].reshape(cfg.shape))
    assert np.array_equal(oracle.deriv_x(cfg, f), res[f.size:].reshape(cfg.shape))


# --------------------------------------------------------------------------
# closed forms for the stencils

def _pw_eig(cfg, m):
    """Discrete Laplacian eigenvalue of the periodic plane wave with mode m."""
    lam = 0.0
    for ma, n, d in ((m[0], cfg.nx, cfg.dx), (m[1], cfg.ny, cfg.dy), (m[2], cfg.nz, cfg.dz)):
        th = 2 * math.pi * ma / n
        if cfg.order == 2:
            lam += (2 * math.cos(th) - 2) / d ** 2
        else:
            lam += (-2 * math.cos(2 * th) + 32 * math.cos(th) - 30) / (12 * d ** 2)
    return lam


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("m", [(1, 0, 0), (0, 2, 0), (0, 0, 3), (2, -3, 1)])
def test_laplacian_plane_wave_eigen(oracle, order, m):
    cfg = _cfg(order=order, bcs=1)
    f = ic.plane_wave(cfg, m)
    lap = oracle.laplacian(cfg, f)
    lam = _pw_eig(cfg, m)
    assert np.max(np.abs(lap - lam * f)) <= 1e-12 * max(1.0, abs(lam))


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("m", [(1, 0, 0), (0, 2, 0), (0, 0, 3), (2, 1, 4)])
def test_laplacian_reflective_cosine_eigen(oracle, order, m):
    cfg = _cfg(order=order, bcs=2)
    nz, ny, nx = cfg.shape
    f = np.ones(cfg.shape, dtype=np.complex128)
    lam = 0.0
    for ax, (ma, n, d) in enumerate(((m[0], nx, cfg.dx), (m[1], ny, cfg.dy), (m[2], nz, cfg.dz))):
        th = math.pi * ma / (n - 1)
        idx = np.cos(th * np.arange(n))
        shape = [1, 1, 1]
        shape[2 - ax] = n
        f = f * idx.reshape(shape)
        if order == 2:
            lam += (2 * math.cos(th) - 2) / d ** 2
        else:
            lam += (-2 * math.cos(2 * th) + 32 * math.cos(th) - 30) / (12 * d ** 2)
    lap = oracle.laplacian(cfg, f)
    assert np.max(np.abs(lap - lam * f)) <= 1e-12 * max(1.0, abs(lam))


def test_laplacian_axes_not_transposed(oracle):
    # a mode along one axis only must not leak into the others' spacing
    cfg = _cfg(order=4, bcs=1, xr=1.0, yr=5.0, zr=9.0)
    for m in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        f = ic.plane_wave(cfg, m)
        lam = _pw_eig(cfg, m)
        assert np.allclose(oracle.laplacian(cfg, f), lam * f, rtol=0, atol=1e-12 * abs(lam))


# --------------------------------------------------------------------------
# right-hand side and integrators

@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4", "rk45"])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_uniform_state_is_stationary_bitwise(oracle, scheme, order, bcs):
    cfg = _cfg(scheme=scheme, order=order, bcs=bcs, tau=0.01)
    psi = np.ones(cfg.shape, dtype=np.complex128)
    assert not np.any(oracle.rhs(cfg, psi))
    st = oracle.make_state(cfg.dt)
    out = oracle.solver(cfg, st, psi)
    assert np.array_equal(out, psi)


def _scalar_rk(scheme, f, a, h):
    """Textbook explicit schemes on a scalar ODE a' = f(a)."""
    if scheme == "euler":
        return a + h * f(a)
    if scheme == "rk2":
        return a + h * f(a + 0.5 * h * f(a))
    k1 = f(a)
    k2 = f(a + 0.5 * h * k1)
    k3 = f(a + 0.5 * h * k2)
    k4 = f(a + h * k3)
    return a + h * (k1 + 2 * k2 + 2 * k3 + k4) / 6


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4"])
@pytest.mark.parametrize("order", [2, 4])
def test_plane_wave_reduces_to_scalar_ode(oracle, scheme, order):
    g, mu, a0, dt, n = 0.7, 1.3, 0.8 + 0.3j, 0.01, 25
    cfg = _cfg(scheme=scheme, order=order, bcs=1, g=g, mu=mu, tau=dt)
    m = (1, 2, 1)
    lam = _pw_eig(cfg, m)
    f = lambda a: -1j * (-0.5 * lam + g * abs(a) ** 2 - mu) * a  # noqa: E731
    a = a0
    for _ in range(n):
        a = _scalar_rk(scheme, f, a, dt)
    psi0 = a0 * ic.plane_wave(cfg, m)
    out, _ = oracle.run(cfg, oracle.make_state(cfg.dt), psi0, n)
    assert np.max(np.abs(out - a * ic.plane_wave(cfg, m))) < 1e-12


@pytest.mark.parametrize("scheme,R", [
    ("euler", lambda z: 1 + z),
    ("rk2", lambda z: 1 + z + z * z / 2),
    ("rk4", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24),
])
def test_linear_amplification_closed_form(oracle, scheme, R):
    mu, dt, n = 0.9, 0.02, 40
    cfg = _cfg(scheme=scheme, order=4, bcs=1, g=0.0, mu=mu, tau=dt)
    m = (0, 1, 2)
    omega = 0.5 * abs(_pw_eig(cfg, m)) - mu
    amp = R(-1j * omega * dt) ** n
    psi0 = ic.plane_wave(cfg, m)
    out, _ = oracle.run(cfg, oracle.make_state(cfg.dt), psi0, n)
    assert np.max(np.abs(out - amp * psi0)) < 1e-13


def test_imaginary_time_linear_decay(oracle):
    # dt = -i tau: f = -i w a with w real gives a' = -w tau a per step.
    mu, tau, n = 0.4, 0.01, 30
    cfg = _cfg(scheme="rk4", order=2, bcs=1, g=0.0, mu=mu, tau=tau, real_time=False)
    m = (1, 1, 0)
    omega = 0.5 * abs(_pw_eig(cfg, m)) - mu
    z = -1j * omega * (-1j * tau)
    amp = (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) ** n
    st = oracle.make_state(cfg.dt)
    out, _ = oracle.run(cfg, st, ic.plane_wave(cfg, m), n)
    assert np.max(np.abs(out - amp * ic.plane_wave(cfg, m))) < 1e-13
    assert st.t == 0.0 and abs(st.im_t - n * tau) < 1e-14


# Cash-Karp tableau (numerics.tex table tab:cash_karp), written out for the
# scalar reference used below.
_CK_B = [[], [1 / 5], [3 / 40, 9 / 40], [3 / 10, -9 / 10, 6 / 5], [-11 / 54, 5 / 2, -70 / 27, 35 / 27],
         [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]]
_CK_C = [37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771]
_CK_CS = [2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 1 / 4]


def _ck_scalar(f, a, h):
    k = []
    for s in range(6):
        u = a + sum(b * kk for b, kk in zip(_CK_B[s], k))
        k.append(h * f(u))
    y5 = a + sum(c * kk for c, kk in zip(_CK_C, k))
    err = sum((c - cs) * kk for c, cs, kk in zip(_CK_C, _CK_CS, k))
    return y5, err


def test_cash_karp_tableau_orders():
    # independent of the oracle: the tableau has order 5 (solution) and the
    # embedded difference is O(h^5), on the nonlinear plane-wave ODE
    g, mu, lam, a0 = 0.8, 1.1, -2.3, 0.9 + 0.2j
    w = -0.5 * lam + g * abs(a0) ** 2 - mu
    f = lambda a: -1j * (-0.5 * lam + g * abs(a) ** 2 - mu) * a  # noqa: E731
    errs, ests = [], []
    for h in (0.2, 0.1, 0.05):
        y5, e = _ck_scalar(f, a0, h)
        errs.append(abs(y5 - a0 * np.exp(-1j * w * h)))
        ests.append(abs(e))
    assert 40 < errs[0] / errs[1] < 90 and 40 < errs[1] / errs[2] < 90
    assert 20 < ests[0] / ests[1] < 45 and 20 < ests[1] / ests[2] < 45


def test_rk45_plane_wave_matches_scalar_cash_karp(oracle):
    g, mu, a0 = 0.6, 1.2, 0.7 - 0.4j
    cfg = _cfg(scheme="rk45", order=2, bcs=1, g=g, mu=mu, tau=1e-3)
    m = (1, 0, 1)
    lam = _pw_eig(cfg, m)
    f = lambda a: -1j * (-0.5 * lam + g * abs(a) ** 2 - mu) * a  # noqa: E731
    st = oracle.make_state(cfg.dt)
    out = oracle.solver(cfg, st, a0 * ic.plane_wave(cfg, m))
    y5, e = _ck_scalar(f, a0, 1e-3)
    assert np.max(np.abs(out - y5 * ic.plane_wave(cfg, m))) < 1e-13
    # errmax = max|err/scale|/eps with scale = |a| + dt|f(a)| + 1e-30
    scale = abs(a0) + 1e-3 * abs(f(a0)) + 1e-30
    assert st.errmax == pytest.approx(abs(e) / scale / 1e-6, rel=1e-6)
    # tiny error -> dt grows by 5 (solve.f90:214-220), t advances by dt
    assert st.errmax < (5 / 0.9) ** (1 / -0.2)
    assert st.dt_re == pytest.approx(5e-3, rel=1e-15) and st.t == pytest.approx(1e-3, rel=1e-15)


def test_rk45_rejects_then_accepts(oracle):
    # a step far too large is rejected and shrunk until errmax <= 1
    g, mu, a0, h = 1.0, 1.0, 1.0 + 0.5j, 0.8
    cfg = _cfg(scheme="rk45", order=4, bcs=1, g=g, mu=mu, tau=h)
    m = (2, 1, 1)
    lam = _pw_eig(cfg, m)
    f = lambda a: -1j * (-0.5 * lam + g * abs(a) ** 2 - mu) * a  # noqa: E731
    # scalar restatement of the control loop of rkqs (Numerical Recipes 16.2)
    scale = abs(a0) + h * abs(f(a0)) + 1e-30
    dt, rej = h, 0
    while True:
        y5, e = _ck_scalar(f, a0, dt)
        em = abs(e) / scale / 1e-6
        if em <= 1:
            break
        rej += 1
        dt = max(abs(0.9 * dt * em ** -0.25), 0.1 * abs(dt))
    st = oracle.make_state(cfg.dt)
    out = oracle.solver(cfg, st, a0 * ic.plane_wave(cfg, m))
    assert rej > 0 and st.n_rejects == rej
    assert st.t == pytest.approx(dt, rel=1e-12)
    assert np.max(np.abs(out - y5 * ic.plane_wave(cfg, m))) < 1e-12


# --------------------------------------------------------------------------
# reductions

def test_norm_of_unit_field_is_box_volume(oracle):
    cfg = _cfg(nx=17, ny=11, nz=9, xr=3.0, yr=2.0, zr=1.5)
    psi = np.exp(1j * np.linspace(0, 7, 17 * 11 * 9)).reshape(cfg.shape)
    assert oracle.norm(cfg, psi) == pytest.approx(6.0 * 4.0 * 3.0, rel=1e-14)


def test_norm_exact_on_cubics(oracle):
    # the extended closed rule integrates cubics exactly along every axis
    cfg = _cfg(nx=13, ny=11, nz=10, xr=2.0, yr=1.5, zr=1.0)
    x, y, z = (a.astype(np.float64) for a in cfg.coords())
    px = 2.0 + x + 0.3 * x ** 2 + 0.1 * x ** 3
    py = 3.0 - 0.5 * y + 0.2 * y ** 3
    pz = 1.5 + 0.25 * z ** 2 - 0.4 * z ** 3
    dens = pz[:, None, None] * py[None, :, None] * px[None, None, :]
    psi = np.sqrt(dens) * np.exp(0.3j)

    def integ(c, r):  # exact integral of c0 + c1 t + c2 t^2 + c3 t^3 over [-r, r]
        return 2 * c[0] * r + 2 * c[2] * r ** 3 / 3

    exact = integ((2.0, 1, 0.3, 0.1), 2.0) * integ((3.0, -0.5, 0, 0.2), 1.5) * integ((1.5, 0, 0.25, -0.4), 1.0)
    assert oracle.norm(cfg, psi) == pytest.approx(exact, rel=1e-13)


@pytest.mark.parametrize("order", [2, 4])
def test_energy_and_mass_of_plane_wave(oracle, order):
    cfg = _cfg(order=order, bcs=1)
    m = (1, 1, 2)
    lam = _pw_eig(cfg, m)
    psi = ic.plane_wave(cfg, m, amp=1.3)
    n, dv = psi.size, cfg.dx * cfg.dy * cfg.dz
    a2 = 1.3 ** 2
    assert oracle.energy(cfg, psi) == pytest.approx((abs(lam) * a2 + 0.5 * a2 * a2) * n * dv, rel=1e-12)
    assert oracle.mass(cfg, psi) == pytest.approx(a2 * n * dv, rel=1e-13)


def test_renormalise(oracle):
    cfg = _cfg()
    psi = ic.plane_wave(cfg, (1, 0, 0), amp=2.0)
    out = oracle.renormalise(cfg, psi, 3.0, 4.0)
    assert np.allclose(out, psi * math.sqrt(3.0) / 2.0, rtol=1e-15, atol=0)


def test_fp32_oracle_tracks_fp64(oracle):
    cfg64 = _cfg(scheme="rk4", order=4, bcs=2, tau=0.005)
    cfg32 = _cfg(scheme="rk4", order=4, bcs=2, tau=0.005, precision="fp32")
    psi = ic.vortex_ring(cfg64, 1.0)
    a, _ = oracle.run(cfg64, oracle.make_state(cfg64.dt), psi, 10)
    b, _ = oracle.run(cfg32, oracle.make_state(cfg32.dt), psi.astype(np.complex64), 10)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-5
