"""Initial conditions for the benchmark configurations (host side).

The reference builds its initial state on the host (ic.f90: ics, vortex_ring,
vortex_line, random_phase, ...) and so does this package; the device path
starts from an uploaded field.  The generators here produce the states the
benchmark configurations name, on the reference grid x_i = xl + i*dx
(ic.f90:32-69).  Seeded draws use SplitMix64 so a (config, seed) pair is the
same field on every machine.
"""
import numpy as np

from .config import RunConfig

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n uint64 draws of SplitMix64 from `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """n doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _mesh(cfg: RunConfig):
    """Broadcastable coordinate arrays x[1,1,nx], y[1,ny,1], z[nz,1,1]."""
    x, y, z = (a.astype(np.float64) for a in cfg.coords())
    return x[None, None, :], y[None, :, None], z[:, None, None]


def trap_potential(cfg: RunConfig):
    """Vtrap = 0.5*(omx^2 x^2 + omy^2 y^2 + omz^2 z^2) (ic.f90:713-731)."""
    x, y, z = _mesh(cfg)
    return 0.5 * ((cfg.omx ** 2) * x * x + (cfg.omy ** 2) * y * y + (cfg.omz ** 2) * z * z)


def vortex_ring(cfg: RunConfig, radius, x0=0.0, y0=0.0, z0=0.0):
    """Ring of the given radius in the z = z0 plane, centred at (x0, y0):
    amplitude tanh(d/sqrt 2) with d the distance to the core circle, phase
    atan2(z, rho - R) - atan2(z, rho + R)."""
    x, y, z = _mesh(cfg)
    rho = np.sqrt((x - x0) ** 2 + (y - y0) ** 2)
    zz = z - z0
    d = np.sqrt((rho - radius) ** 2 + zz ** 2)
    amp = np.tanh(d / np.sqrt(2.0))
    phase = np.arctan2(zz, rho - radius) - np.arctan2(zz, rho + radius)
    return (amp * np.exp(1j * phase)).astype(np.complex128)


def line_pair(cfg: RunConfig, sep=3.0, mu_ic=10.0):
    """Two straight vortex lines along z at (+-sep, 0) in a Thomas-Fermi
    background sqrt(max(0, mu_ic - V)): phase atan2(y, x-sep) - atan2(y, x+sep),
    core amplitude tanh(r/sqrt 2) for each line."""
    x, y, z = _mesh(cfg)
    V = trap_potential(cfg)
    tf = np.sqrt(np.maximum(0.0, mu_ic - V))
    rp = np.sqrt((x - sep) ** 2 + y ** 2)
    rm = np.sqrt((x + sep) ** 2 + y ** 2)
    amp = np.tanh(rp / np.sqrt(2.0)) * np.tanh(rm / np.sqrt(2.0))
    phase = np.arctan2(y, x - sep) - np.arctan2(y, x + sep)
    psi = tf * amp * np.exp(1j * phase)
    return np.broadcast_to(psi, cfg.shape).astype(np.complex128)


def random_phase(cfg: RunConfig, seed, kmax=0.5, workers=-1):
    """Random-phase field: every Fourier mode k = 2*pi*m/(n*dx) with |k| <= kmax
    gets unit amplitude and a uniform phase in [0, 2*pi) (SplitMix64 from
    `seed`, drawn in the order of the modes' flat index); inverse FFT; then
    rescaled so <|psi|^2> = 1."""
    import scipy.fft as sfft

    nz, ny, nx = cfg.shape
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=cfg.dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=cfg.dy)
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=cfg.dz)
    # modes inside the sphere only (a small set), then scatter
    mx = np.nonzero(np.abs(kx) <= kmax)[0]
    my = np.nonzero(np.abs(ky) <= kmax)[0]
    mz = np.nonzero(np.abs(kz) <= kmax)[0]
    K2 = kz[mz][:, None, None] ** 2 + ky[my][None, :, None] ** 2 + kx[mx][None, None, :] ** 2
    sel = np.nonzero(K2 <= kmax * kmax)
    iz, iy, ix = mz[sel[0]], my[sel[1]], mx[sel[2]]
    flat = (iz.astype(np.int64) * ny + iy) * nx + ix
    order = np.argsort(flat)
    flat = flat[order]
    ph = 2 * np.pi * uniform(seed, flat.size)
    spec = np.zeros(cfg.shape, dtype=np.complex128)
    spec.reshape(-1)[flat] = np.exp(1j * ph)
    psi = sfft.ifftn(spec, workers=workers, overwrite_x=True)
    psi *= 1.0 / np.sqrt(np.mean(np.abs(psi) ** 2))
    return np.ascontiguousarray(psi, dtype=np.complex128)


def trapped_ring(cfg: RunConfig, radius, mu_ic=5.0):
    """Centred vortex ring (amplitude/phase of vortex_ring) times
    sqrt(max(0, mu_ic - V))."""
    return vortex_ring(cfg, radius) * np.sqrt(np.maximum(0.0, mu_ic - trap_potential(cfg)))


def plane_wave(cfg: RunConfig, m=(1, 0, 0), amp=1.0):
    """amp * exp(i k.x) with k = 2*pi*m/(n*d) (a periodic-grid eigenmode)."""
    nz, ny, nx = cfg.shape
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    ph = 2 * np.pi * (m[0] * i / nx + m[1] * j / ny + m[2] * k / nz)
    return (amp * np.exp(1j * ph)).astype(np.complex128)
