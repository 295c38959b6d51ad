"""The oracle's imprint_vortex_line (variables.f90:1166-1180, with
vortex_line/amp/get_theta/get_r of ic.f90) pinned on the CPU.

- A straight line along z (no disturbance, so the reference's safe
  wavelength applies) imprints the phase sgn*atan2(y - y0, x - x0), the
  library routine, and keeps |psi|.
- sgn = -1 gives the complex conjugate, bit for bit.
- dir = 'x' and dir = 'y' are dir = 'z' on cyclically relabelled axes, bit
  for bit, on a cube.
- A disturbed line with its core (imprint_phase false, |psi| = 1): the
  vortex-core points of the result (the reference's zeros) trace the
  displaced centre x0 + a1 sin(2 pi (z - z0)/ll), y0 + a2 cos(...) plane by
  plane; the core amplitude rises from 0 at the line towards 1.
- A zero wavelength with a disturbance is the reference's emergency stop."""
import numpy as np
import pytest

from gpe_amd.config import RunConfig


def _cube(n=24, dx=0.5, **kw):
    xr = dx * (n - 1) / 2
    return RunConfig(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, **kw)


def _field(cfg, seed=3):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.standard_normal(cfg.shape) + 1j * rng.standard_normal(cfg.shape))


def test_straight_line_phase_is_atan2(oracle):
    c = _cube(n=20)
    x, y, z = c.coords()
    psi = _field(c)
    for sgn in (1.0, -1.0):
        line = dict(x0=-1.3, y0=0.4, z0=0.0, sgn=sgn, dir="z", imprint_phase=True)
        out = oracle.imprint_vortex_line(c, psi, line)
        th = sgn * np.arctan2((y - 0.4)[None, :, None] + 0 * z[:, None, None], (x + 1.3)[None, None, :])
        np.testing.assert_allclose(np.angle(out * np.exp(-1j * th)), 0.0, atol=1e-14)
        np.testing.assert_allclose(np.abs(out), np.abs(psi), rtol=1e-15)


def test_sign_conjugates_and_axes_permute(oracle):
    c = _cube(n=18)
    psi = _field(c, 5)
    base = dict(x0=0.7, y0=-0.9, z0=0.3, amp1=0.4, amp2=0.25, ll=3.0, imprint_phase=False)
    a = oracle.imprint_vortex_line(c, psi, dict(base, sgn=1.0, dir="z"))
    b = oracle.imprint_vortex_line(c, psi, dict(base, sgn=-1.0, dir="z"))
    np.testing.assert_array_equal(b, np.conj(a))
    # arrays are [z, y, x].  A line along x with origin x0' through (y0', z0')
    # is the z-line of a field relabelled (x, y, z) -> (y, z, x):
    # g[k, j, i] = psi[i, k, j], line origins (z0, x0, y0) of the z-line
    g = np.ascontiguousarray(np.transpose(psi, (1, 2, 0)))
    gx = oracle.imprint_vortex_line(c, g, dict(base, sgn=1.0, dir="x", x0=base["z0"], y0=base["x0"],
                                               z0=base["y0"]))
    np.testing.assert_array_equal(gx, np.transpose(a, (1, 2, 0)))
    # a line along y: h[k, j, i] = psi[j, i, k]
    h = np.ascontiguousarray(np.transpose(psi, (2, 0, 1)))
    hy = oracle.imprint_vortex_line(c, h, dict(base, sgn=1.0, dir="y", x0=base["y0"], y0=base["z0"],
                                               z0=base["x0"]))
    np.testing.assert_array_equal(hy, np.transpose(a, (2, 0, 1)))


def test_disturbed_line_core_follows_the_displacement(oracle):
    c = _cube(n=40, dx=0.4)
    x, y, z = c.coords()
    x0, y0, a1, a2, ll = 0.23, -0.31, 1.2, 0.8, 6.4
    ones = np.ascontiguousarray(np.ones(c.shape, dtype=np.complex128))
    out = oracle.imprint_vortex_line(c, ones, dict(x0=x0, y0=y0, z0=0.0, amp1=a1, amp2=a2, ll=ll, sgn=1.0, dir="z",
                                                   imprint_phase=False))
    pts = oracle.zeros(c, out, 0)
    pts = pts[np.isfinite(pts["x"]) & np.isfinite(pts["y"])]
    assert len(pts) >= c.nz - 2
    cx = x0 + a1 * np.sin(2 * np.pi * pts["z"] / ll)
    cy = y0 + a2 * np.cos(2 * np.pi * pts["z"] / ll)
    assert np.max(np.hypot(pts["x"] - cx, pts["y"] - cy)) < 0.5 * c.dx
    # the core amplitude: 0 at the line, 1 - exp(-0.7 r^1.15) -> 1 away from it
    a = np.abs(out)
    assert a.min() < 0.35 and a.max() > 0.999 and np.all(a < 1.0)


def test_zero_wavelength_with_disturbance_stops(oracle):
    c = _cube(n=12)
    with pytest.raises(ValueError, match="wavelength is zero"):
        oracle.imprint_vortex_line(c, _field(c), dict(amp1=0.5, ll=0.0, dir="z"))


def test_single_precision(oracle):
    c = _cube(n=16)
    psi = _field(c, 9)
    line = dict(x0=0.5, y0=0.2, z0=-0.1, amp1=0.3, amp2=0.2, ll=2.5, sgn=1.0, dir="y", imprint_phase=False)
    a = oracle.imprint_vortex_line(c, psi, line)
    b = oracle.imprint_vortex_line(c.with_(precision="fp32"), psi.astype(np.complex64), line)
    assert np.abs(b - a).max() < 2e-5 * np.abs(a).max()
