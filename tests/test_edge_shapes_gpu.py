"""Edge shapes and the stationary state on the GPU.

* The smallest grids the library takes (5 points per axis), grids smaller
  than one tile in every direction, exact tile multiples and one point past
  them, z extents shorter than the prefetch and the priming window: three
  FP64 solver() calls against the oracle, bit for bit in psi, t and dt, for
  RK4 and RKF45, both orders and both boundary conditions.  Index maps at a
  wall stay in range for n >= 3 (the mirror of n+1 is n-3); every tile's
  rows and columns beyond the grid are clamped to n-1+R before mapping.
* psi = 1 with V = 0 and g = mu = 1 has a right-hand side of exactly zero
  (the stencil weights sum to zero, |psi|^2 - mu = 0): every scheme and
  every precision form must leave it unchanged bit for bit, the compensated
  form's low word included.
"""
import os
import sys

import numpy as np
import pytest

from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)

SHAPES = [(5, 5, 5), (5, 6, 7), (7, 5, 9), (64, 16, 8), (65, 17, 9), (128, 8, 5), (6, 40, 5)]


def _cfg(shape, **kw):
    nx, ny, nz = shape
    dx = 0.4
    base = dict(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2, eqn=4,
                g=1.0, mu=1.0, omx=0.3, omy=0.2, omz=0.4, tau=0.01)
    base.update(kw)
    return RunConfig(**base)


@pytest.mark.parametrize("bcs", [1, 2])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("scheme", ["rk4", "rk45"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "%dx%dx%d" % s)
def test_edge_shapes_fp64_bitwise(oracle, shape, scheme, order, bcs):
    from gpe_amd.sim import Simulation
    cfg = _cfg(shape, scheme=scheme, order=order, bcs=bcs)
    rng = np.random.default_rng(sum(shape) * 131 + order * 7 + bcs)
    psi = (1.0 + 0.2 * rng.standard_normal(cfg.shape)) * np.exp(1j * rng.uniform(0, 0.5, cfg.shape))
    print("[case] edge shape", shape, scheme, order, bcs, flush=True)
    sys.stdout.flush()
    st = oracle.make_state(cfg.dt)
    want = psi
    for _ in range(3):
        want = oracle.solver(cfg, st, want, NT)
    with Simulation(cfg, psi) as sim:
        sim.advance(3)
        got = sim.download()
        s = sim.state()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert s["t"] == st.t and s["dt"].real == st.dt_re


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4", "rk45"])
def test_uniform_state_is_stationary(precision, scheme):
    from gpe_amd.sim import Simulation
    cfg = _cfg((37, 21, 11), scheme=scheme, order=4, bcs=1, omx=0.0, omy=0.0, omz=0.0, precision=precision)
    print("[case] uniform state", precision, scheme, flush=True)
    sys.stdout.flush()
    psi = np.ones(cfg.shape, dtype=np.complex128)
    with Simulation(cfg, psi) as sim:
        sim.advance(5)
        got = sim.download(np.complex128)
        s = sim.state()
    assert np.array_equal(got, psi)
    assert s["p"] == 5
