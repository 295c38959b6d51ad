"""Closed-form pins of the oracle's eqn 1-3 branches (CPU).

get_rhs (solve.f90:382-440) for eqns 1-3 adds three things eqn 4 does not
have: the frame term Urhs * deriv_x (solve.f90:411, 415-423), the sponge
diss (solve.f90:397-409) and, for eqn 3, the sphere() mask
(ic.f90:640-662).  On a plane wave every term is a known multiple of the
field, so each branch reduces to a scalar formula:

  deriv_x e^{i theta i}        = i s_x e^{i theta i},
      s_x = sin(theta)/dx (order 2), (8 sin theta - sin 2 theta)/(6 dx) (order 4)
  laplacian                    = lam * psi (the eigenvalue of test_oracle.py)
  eqn 1: f = 0.5 (i + D) (lam + 1 - |a|^2) psi + U i s_x psi
  eqn 2: f = i (lam - |a|^2) psi + U i s_x psi
  eqn 3: f = eqn 1 * 0.5 (1 + tanh(r - 10 - 2)),  r = |(x, y, z)|
  D = diss_amp * sum_axes (1 + tanh(x - a) tanh(x + a)),  a = x(nx-3)

A cosine mode on a reflective grid checks the mirrored deriv_x the same
way (its derivative is a sine mode).  Each formula is evaluated here with
numpy from the grid alone; a dropped term, a wrong sign, a wrong ∂x order,
a sponge edge taken at x(nx-2) or a sphere radius off by one fails it.
Eqn 2 keeps |a| fixed on a plane wave, so every integrator reduces to the
same integrator on the scalar ODE, checked over many steps.
"""
import math

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig


def _cfg(**kw):
    base = dict(nx=16, ny=12, nz=10, xr=3.3, yr=2.6, zr=2.1, bcs=1, order=2, eqn=1, g=1.0, mu=1.0)
    base.update(kw)
    return RunConfig(**base)


def _lam_axis(order, th, d):
    if order == 2:
        return (2 * math.cos(th) - 2) / d ** 2
    return (-2 * math.cos(2 * th) + 32 * math.cos(th) - 30) / (12 * d ** 2)


def _sx(order, th, d):
    if order == 2:
        return math.sin(th) / d
    return (8 * math.sin(th) - math.sin(2 * th)) / (6 * d)


def _pw(cfg, m, amp):
    nz, ny, nx = cfg.shape
    lam = (_lam_axis(cfg.order, 2 * math.pi * m[0] / nx, cfg.dx) + _lam_axis(cfg.order, 2 * math.pi * m[1] / ny, cfg.dy)
           + _lam_axis(cfg.order, 2 * math.pi * m[2] / nz, cfg.dz))
    return ic.plane_wave(cfg, m, amp=amp), lam, _sx(cfg.order, 2 * math.pi * m[0] / nx, cfg.dx)


def _mesh(cfg):
    x, y, z = (a.astype(np.float64) for a in cfg.coords())
    return x[None, None, :], y[None, :, None], z[:, None, None]


def _sponge(cfg):
    x, y, z = (a.astype(np.float64) for a in cfg.coords())
    a, b, c = x[cfg.nx - 3], y[cfg.ny - 3], z[cfg.nz - 3]
    X, Y, Z = _mesh(cfg)
    return cfg.diss_amp * ((1 + np.tanh(X - a) * np.tanh(X + a)) + (1 + np.tanh(Y - b) * np.tanh(Y + b))
                           + (1 + np.tanh(Z - c) * np.tanh(Z + c)))


def _close(got, want, tol=1e-12):
    scale = max(1.0, float(np.max(np.abs(want))))
    err = float(np.max(np.abs(got - want)))
    assert err <= tol * scale, err


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("m", [(1, 0, 0), (2, 1, 1), (-3, 2, 0)])
def test_deriv_x_plane_wave_symbol(oracle, order, m):
    cfg = _cfg(order=order)
    psi, _, sx = _pw(cfg, m, 1.0)
    _close(oracle.deriv_x(cfg, psi), 1j * sx * psi)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("mx", [1, 2, 5])
def test_deriv_x_reflective_cosine_mode(oracle, order, mx):
    # cos(theta i), theta = pi mx/(nx-1): the even mirror about both walls
    # is the mode's own extension, so d/dx gives -s_x sin(theta i) everywhere
    cfg = _cfg(order=order, bcs=2)
    nz, ny, nx = cfg.shape
    th = math.pi * mx / (nx - 1)
    i = np.arange(nx)
    f = np.ascontiguousarray(np.broadcast_to(np.cos(th * i)[None, None, :], cfg.shape) * (0.6 - 0.8j))
    want = np.broadcast_to(-_sx(order, th, cfg.dx) * np.sin(th * i)[None, None, :], cfg.shape) * (0.6 - 0.8j)
    _close(oracle.deriv_x(cfg, f), want)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("urhs", [0.0, 0.35])
def test_eqn1_plane_wave(oracle, order, urhs):
    amp = 0.9 - 0.4j
    cfg = _cfg(eqn=1, order=order, urhs=urhs)
    psi, lam, sx = _pw(cfg, (1, 2, 1), amp)
    want = (0.5j * (lam + 1 - abs(amp) ** 2) + urhs * 1j * sx) * psi
    _close(oracle.rhs(cfg, psi), want)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("urhs", [0.0, -0.25])
def test_eqn2_plane_wave(oracle, order, urhs):
    amp = 0.7 + 0.5j
    cfg = _cfg(eqn=2, order=order, urhs=urhs)
    psi, lam, sx = _pw(cfg, (2, 0, 1), amp)
    want = (1j * (lam - abs(amp) ** 2) + urhs * 1j * sx) * psi
    _close(oracle.rhs(cfg, psi), want)


@pytest.mark.parametrize("order", [2, 4])
def test_eqn1_sponge(oracle, order):
    # the sponge is ~diss_amp near the walls and ~0 inside; with a small
    # box every site sees a different value of it
    amp = 1.1 + 0.2j
    cfg = _cfg(eqn=1, order=order, urhs=0.2, diss_amp=0.3)
    psi, lam, sx = _pw(cfg, (1, 1, 1), amp)
    D = _sponge(cfg)
    assert float(D.max() - D.min()) > 0.1
    want = 0.5 * (1j + D) * (lam + 1 - abs(amp) ** 2) * psi + 0.2 * 1j * sx * psi
    _close(oracle.rhs(cfg, psi), want)


@pytest.mark.parametrize("diss", [0.0, 0.3])
def test_eqn3_is_eqn1_times_sphere(oracle, diss):
    # a box reaching past r = 12, so the mask runs from ~0 to ~1
    amp = 0.8 + 0.3j
    cfg = _cfg(eqn=3, order=4, urhs=0.15, diss_amp=diss, nx=20, ny=18, nz=16, xr=11.0, yr=10.0, zr=9.0)
    psi, lam, sx = _pw(cfg, (1, 0, 2), amp)
    X, Y, Z = _mesh(cfg)
    mask = 0.5 * (1 + np.tanh(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 10.0 - 2.0))
    assert mask.min() < 1e-3 and mask.max() > 0.9
    D = _sponge(cfg)
    want = (0.5 * (1j + D) * (lam + 1 - abs(amp) ** 2) * psi + 0.15 * 1j * sx * psi) * mask
    _close(oracle.rhs(cfg, psi), want)


def _scalar_rk(scheme, f, a, h):
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
@pytest.mark.parametrize("eqn", [1, 2])
def test_integrators_on_eqn12_plane_wave(oracle, scheme, eqn):
    a0, dt, n, u = 0.8 + 0.3j, 0.004, 30, 0.4
    cfg = _cfg(eqn=eqn, order=4, urhs=u, scheme=scheme, tau=dt)
    psi0, lam, sx = _pw(cfg, (1, 1, 0), 1.0)
    if eqn == 1:
        f = lambda a: (0.5j * (lam + 1 - abs(a) ** 2) + u * 1j * sx) * a  # noqa: E731
    else:
        f = lambda a: (1j * (lam - abs(a) ** 2) + u * 1j * sx) * a  # noqa: E731
    a = a0
    for _ in range(n):
        a = _scalar_rk(scheme, f, a, dt)
    out, _ = oracle.run(cfg, oracle.make_state(cfg.dt), a0 * psi0, n)
    _close(out, a * psi0)
