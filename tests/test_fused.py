"""Fused stage pairs against the stage-by-stage sweeps.

The fused kernels compute every value with the same expressions in the same
order as the one-stage kernels, so the fields, times and step sizes must be
bitwise identical -- for both boundary conditions (the reflective walls
exercise the mirror-partner handling in x, y and z), both orders, both FP32
forms and every pair tile variant, with short z-chunks so chunk seams are
crossed.  FP64, the parity path, has no fused pairs."""
import os
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def cfg_of(shape, **kw):
    nx, ny, nz = shape
    dx = 0.35
    base = dict(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2, eqn=4,
                g=1.0, mu=1.0, omx=0.2, omy=0.15, omz=0.3, tau=0.005)
    base.update(kw)
    return RunConfig(**base)


def ic_of(cfg):
    psi = ic.vortex_ring(cfg, 0.25 * cfg.nx * cfg.dx)
    rng = np.random.default_rng(3)
    return psi * (1 + 0.05 * rng.standard_normal(cfg.shape))


def _run(cfg, psi, n, fused, tune=None):
    from gpe_amd import _lib
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi) as sim:
        sim.set_fusion(fused)
        if fused and not tune:
            tune = {p: (0, 0) for p in _lib.PAIR_KINDS}  # every pair fused
        if tune:
            for k, (v, z) in tune.items():
                sim.tune_set(k, v, z)
        sim.advance(n)
        return sim.download(np.complex128), sim.state()


SHAPES = [(70, 21, 19), (16, 16, 16), (40, 9, 25)]


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
@pytest.mark.parametrize("scheme", ["rk4", "rk45"])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_fused_equals_unfused(precision, scheme, order, bcs):
    for shape in SHAPES:
        cfg = cfg_of(shape, scheme=scheme, order=order, bcs=bcs, precision=precision,
                     tau=0.05 if scheme == "rk45" else 0.005)
        psi = ic_of(cfg)
        say("fused vs unfused", precision, scheme, order, bcs, shape)
        a, sa = _run(cfg, psi, 6, False)
        b, sb = _run(cfg, psi, 6, True)
        assert np.array_equal(a, b), (shape, float(np.abs(a - b).max()))
        assert sa["t"] == sb["t"] and sa["dt"] == sb["dt"] and sa["n_rejects"] == sb["n_rejects"]


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("scheme", ["rk4", "rk45"])
def test_pair_variants_and_chunks(variant, scheme):
    cfg = cfg_of((70, 37, 29), scheme=scheme, order=4, bcs=2, precision="fp32c",
                 tau=0.05 if scheme == "rk45" else 0.005)
    psi = ic_of(cfg)
    pairs = ["rk4_12", "rk4_34"] if scheme == "rk4" else ["ck_12", "ck_34", "ck_56"]
    say("pair variant", variant, scheme)
    a, _ = _run(cfg, psi, 5, False)
    b, _ = _run(cfg, psi, 5, True, tune={p: (variant, 8) for p in pairs})
    assert np.array_equal(a, b)


def test_fused_fp32c_vs_oracle_and_fp64_refuses(oracle):
    from gpe_amd import _lib
    from gpe_amd.sim import Simulation
    cfg = cfg_of((70, 21, 19), scheme="rk4", order=4, bcs=2)
    psi = ic_of(cfg)
    say("fused fp32c rk4 vs oracle; fp64 has no fusion")
    want, _ = oracle.run(cfg, oracle.make_state(cfg.dt), psi, 20, NT)
    with Simulation(cfg.with_(precision="fp32c"), psi) as sim:
        sim.set_fusion(True)
        for p in ("rk4_12", "rk4_34"):
            sim.tune_set(p, 0, 0)
        sim.advance(20)
        got = sim.download(np.complex128)
    assert float(np.linalg.norm(got - want) / np.linalg.norm(want)) < 1e-6
    with Simulation(cfg, psi) as sim:
        with pytest.raises(_lib.GpeError, match="one by one"):
            sim.set_fusion(True)
