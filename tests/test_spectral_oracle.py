"""Pins of the FFT-side diagnostics' oracle (oracle/spectral.py).

The reference's FFT libraries are absent, so these checks pin the
restatement by closed forms (a plane wave is one mode; Parseval; a uniform
field has no excitations; the finite-difference symbol's Taylor series) and
by literal loop transcriptions of spectrum and of condensed_particles'
k-space sum."""
import numpy as np
import pytest

from oracle import spectral as S

SHAPE = (5, 6, 8)  # (nz, ny, nx)


def _plane(shape, m):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.exp(2j * np.pi * (m[0] * x / nx + m[1] * y / ny + m[2] * z / nz))


def test_forward_is_the_unitary_dft():
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    a = S.forward(psi)
    # brute-force DFT, exp(-2 pi i k.n/N) / sqrt(V), x fastest
    nz, ny, nx = psi.shape
    want = np.zeros_like(a)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                for z in range(nz):
                    for y in range(ny):
                        for x in range(nx):
                            want[kz, ky, kx] += psi[z, y, x] * np.exp(
                                -2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    want /= np.sqrt(psi.size)
    assert np.max(np.abs(a - want)) < 1e-12
    # Parseval, and backward undoes forward
    assert np.sum(np.abs(a) ** 2) == pytest.approx(np.sum(np.abs(psi) ** 2), rel=1e-13)
    assert np.max(np.abs(S.backward(a) - psi)) < 1e-13


def test_plane_wave_is_one_mode():
    m = (3, -2, 1)
    a = S.forward(_plane(SHAPE, m))
    V = np.prod(SHAPE)
    idx = (m[2] % SHAPE[0], m[1] % SHAPE[1], m[0] % SHAPE[2])
    assert abs(a[idx]) ** 2 == pytest.approx(V, rel=1e-13)
    b = a.copy()
    b[idx] = 0
    assert np.max(np.abs(b)) < 1e-12


def test_wrap_matches_the_reference_rule():
    # n = 8: n1 = 7, i <= 4 counts as i, above as 8 - i
    assert list(S.wrap2(8)) == [0, 1, 4, 9, 16, 9, 4, 1]
    # n = 7 (odd): n1 = 6, i <= 4 counts as i: index 4 is 16, not (7-4)^2
    assert list(S.wrap2(7)) == [0, 1, 4, 9, 16, 4, 1]


def test_spectrum_matches_literal_transcription():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    nh, eta, F = S.spectrum(a)
    nh2, eta2, F2 = S.spectrum_literal(a)
    assert list(nh) == list(nh2)
    assert np.allclose(eta, eta2, rtol=1e-13, atol=0)
    assert np.allclose(F, F2, rtol=1e-13, atol=0)


def test_spectrum_of_a_plane_wave():
    m = (2, 1, 1)  # |k|^2 = 6 on this grid
    a = S.forward(_plane(SHAPE, m))
    V = float(np.prod(SHAPE))
    nh, eta, F = S.spectrum(a)
    # F(kk) = V once kk >= sqrt(6), 0 before (up to roundoff of the other modes)
    assert F[0] < 1e-20 and F[1] < 1e-20
    assert np.allclose(F[2:], V, rtol=1e-12)
    # the mode's shell: |log2(sqrt(6)/(2 pi))| = 1.36 -> m = 2
    assert int(np.floor(abs(np.log2(np.sqrt(6.0) / (2 * np.pi))))) + 1 == 2
    assert eta[1] == pytest.approx(V / nh[1], rel=1e-12)
    assert np.all(np.delete(eta, 1) < 1e-20)
    # every mode with k != 0 lies in some shell on this small grid
    assert nh.sum() == V - 1


def test_fd_symbol_taylor_series():
    # (-2 cos 2x + 32 cos x - 30)/d^2 * (-1/12) = k^2 - k^6 d^4 / 90 + O(k^8 d^6)
    d = 0.4
    for k in (0.05, 0.1, 0.2):
        got = -(1.0 / 12.0) * S.fd_symbol(k, d)
        assert got == pytest.approx(k * k - k ** 6 * d ** 4 / 90.0, rel=1e-8)
    assert S.fd_symbol(0.0, d) == 0.0


@pytest.mark.parametrize("rho0", [0.0, 0.7])
def test_condensed_sum_matches_literal_transcription(rho0):
    shape, d = (4, 5, 6), (0.4, 0.5, 0.3)
    assert S.condensed_sum(shape, *d, rho0) == pytest.approx(S.condensed_sum_literal(shape, *d, rho0), rel=1e-13)


def test_condensed_particles_of_a_uniform_field():
    # uniform psi = c, dx = dy = dz = 1: all particles condensed (n0 = M),
    # temp = 0, H/V = |c|^4 / 2
    shape = (6, 7, 8)
    c = 0.8 - 0.3j
    psi = np.full(shape, c)
    V = float(np.prod(shape))
    M = V * abs(c) ** 2
    xr, yr, zr = (shape[2] - 1) / 2.0, (shape[1] - 1) / 2.0, (shape[0] - 1) / 2.0
    r = S.condensed_particles(shape, 1.0, 1.0, 1.0, xr, yr, zr, psi, M)
    assert r["n0"] == pytest.approx(M, rel=1e-13)
    assert abs(r["temp"]) < 1e-12 * M
    assert r["energy_per_site"] == pytest.approx(abs(c) ** 4 / 2, rel=1e-12)
    assert r["density"] == pytest.approx(M / (8 * xr * yr * zr), rel=1e-15)


def test_filtered_closed_forms():
    m = (1, 1, 0)  # |k|^2 = 2
    psi = _plane(SHAPE, m)
    a = S.forward(psi)
    assert np.max(np.abs(S.filtered(a, 1e9) - psi)) < 1e-12
    # inside the cut-off the mode is scaled by 1 - m k^2 / kc^2
    f = S.filtered(a, 3.0, m=2.0)
    assert np.max(np.abs(f - (1 - 2.0 * 2 / 9.0) * psi)) < 1e-12
    # beyond it, removed
    assert np.max(np.abs(S.filtered(a, 1.2))) < 1e-12
