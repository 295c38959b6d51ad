"""gpe_energy_noalloc: energy() of a caller's device field with no device
allocation (its partial sums in the library's static scratch array).

Called through the C ABI on device buffers the test allocates with the
library's own HIP runtime, twice in a row on the same GPU -- the second
call on a smaller grid, so it reuses the scratch array the first call
filled -- and each result compared
with gpe_sim_energy on a context of the default launch configuration, bit
for bit (the same kernels and the same fixed summation order)."""
import ctypes

import numpy as np
import pytest

from gpe_amd import _lib, ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu


def _cfg(n, prec, **kw):
    base = dict(nx=n, ny=n - 7, nz=n - 13, xr=0.2 * n, yr=0.2 * (n - 7), zr=0.2 * (n - 13), eqn=4, g=1.0, mu=1.0,
                omx=0.1, omy=0.12, omz=0.15, precision=prec)
    base.update(kw)
    return RunConfig(**base)


def _field(cfg):
    psi = ic.vortex_ring(cfg, 0.2 * cfg.nx * cfg.dx)
    return psi * (1 + 0.02 * np.random.default_rng(cfg.nx).standard_normal(cfg.shape))


def _hip():
    """The HIP runtime the library itself loaded (device buffers for the
    test come from it, so the library and the test share one runtime)."""
    _lib.lib()
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not loaded"
    H = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    H.hipFree.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return H


H2D, D2H = 1, 2


def _noalloc(cfg, psi):
    from gpe_amd.sim import make_params
    H = _hip()
    host = np.ascontiguousarray(psi.astype(np.complex128 if cfg.precision == "fp64" else np.complex64))
    dpsi, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dpsi), host.nbytes) == 0
    assert H.hipMalloc(ctypes.byref(dout), 8) == 0
    try:
        assert H.hipMemcpy(dpsi, host.ctypes.data_as(ctypes.c_void_p), host.nbytes, H2D) == 0
        L = _lib.lib()
        _lib.check(L.gpe_energy_noalloc(ctypes.byref(make_params(cfg)), dpsi,
                                        ctypes.cast(dout, ctypes.POINTER(ctypes.c_double)), None))
        assert H.hipDeviceSynchronize() == 0
        out = np.full(1, np.nan)
        assert H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dout, 8, D2H) == 0
    finally:
        H.hipFree(dpsi)
        H.hipFree(dout)
    return float(out[0])


def _ctx_energy(cfg, psi):
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi, tuned=False) as sim:
        return sim.energy()


@pytest.mark.parametrize("prec,order,bcs", [("fp64", 4, 2), ("fp64", 2, 1), ("fp32c", 4, 1), ("fp32", 2, 2)])
def test_energy_noalloc_equals_context_energy_twice(prec, order, bcs):
    print("[case] energy_noalloc", prec, order, bcs, flush=True)
    for n in (150, 70):  # the second call reuses the scratch on a smaller grid
        cfg = _cfg(n, prec, order=order, bcs=bcs)
        psi = _field(cfg)
        got = _noalloc(cfg, psi)
        want = _ctx_energy(cfg, psi)
        print("[case] n=%d noalloc %.17g context %.17g" % (n, got, want), flush=True)
        assert np.isfinite(got)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)


def test_energy_noalloc_rejects_bad_arguments():
    from gpe_amd.sim import make_params
    L = _lib.lib()
    p = make_params(_cfg(40, "fp64"))
    out = ctypes.c_double()
    assert L.gpe_energy_noalloc(ctypes.byref(p), None, ctypes.byref(out), None) == -1
    p.order = 3
    assert L.gpe_energy_noalloc(ctypes.byref(p), ctypes.c_void_p(1), ctypes.byref(out), None) == -1
    assert b"order" in L.gpe_last_error()
    # a grid needing more partial sums than the scratch holds is refused
    # before anything is launched (the field pointer is never read)
    p = make_params(_cfg(40, "fp64"))
    p.nx, p.ny, p.nz = 16384, 16384, 64
    assert L.gpe_energy_noalloc(ctypes.byref(p), ctypes.c_void_p(1), ctypes.byref(out), None) == -1
    assert b"workgroups" in L.gpe_last_error()
