"""FP64 bit-for-bit: the GPU field equals the oracle's in every bit.

The FP64 kernels evaluate the reference's expressions in its order with one
rounding per operation, the flang complex division, and the host libm's
hypot for abs() (gpe_amd/csrc/refmath.h); the host step control calls the
host libm's pow as the reference does.  Every operation of eqn 4 (and of
eqns 1 and 2 without the tanh sponge) therefore rounds as the oracle's, and
so as the reference binary's: the right-hand side and whole time steps,
including the rk45 error norm, accept/reject decisions and dt, are equal
bit for bit, not only within 1e-12.  (The sponge's tanh and eqn 3's sphere
use the device's tanh, which is held to 1e-14 in test_gpu_parity.py.)
"""
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


def _cfg(**kw):
    base = dict(nx=70, ny=21, nz=19, xr=7.0, yr=4.0, zr=3.6, eqn=4, g=1.0, mu=1.0, omx=0.3, omy=0.2, omz=0.4,
                tau=0.005)
    base.update(kw)
    return RunConfig(**base)


def _field(cfg):
    psi = ic.vortex_ring(cfg, 2.0)
    rng = np.random.default_rng(7)
    return psi * (1 + 0.05 * rng.standard_normal(cfg.shape))


def _bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("eqn,extra", [(4, {}), (2, {"urhs": -0.1}), (1, {"urhs": 0.2})])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_rhs_bitwise(gpu, oracle, eqn, extra, order, bcs):
    from gpe_amd.sim import Simulation
    cfg = _cfg(eqn=eqn, order=order, bcs=bcs, **extra)
    psi = _field(cfg)
    say("rhs bitwise", eqn, order, bcs)
    with Simulation(cfg, psi) as sim:
        got = sim.rhs()
    want = oracle.rhs(cfg, psi, NT)
    say("equal sites %.6f" % np.mean(got == want))
    assert _bits_equal(got, want)


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4", "rk45"])
@pytest.mark.parametrize("order", [2, 4])
def test_steps_bitwise(gpu, oracle, scheme, order):
    from gpe_amd.sim import Simulation
    cfg = _cfg(scheme=scheme, order=order, bcs=1 if order == 4 else 2,
               tau=0.002 if scheme == "euler" else (0.05 if scheme == "rk45" else 0.005))
    psi = _field(cfg)
    n = 25
    say("steps bitwise", scheme, order, n)
    st = oracle.make_state(cfg.dt)
    want, _ = oracle.run(cfg, st, psi, n, NT)
    with Simulation(cfg, psi) as sim:
        sim.advance(n)
        got = sim.download()
        gs = sim.state()
    say("equal sites %.6f, t %.17g vs %.17g" % (np.mean(got == want), gs["t"], st.t))
    assert _bits_equal(got, want)
    assert gs["t"] == st.t and gs["dt"].real == st.dt_re and gs["dt"].imag == st.dt_im


def test_rk45_batched_bitwise(gpu, oracle):
    """A batched rk45 instance runs its step control on the device, with a
    correctly rounded pow where the reference calls the host libm's; on this
    run the two agree, so the field is the oracle's bit for bit."""
    from gpe_amd.batch import Batch
    cfg = _cfg(scheme="rk45", order=4, bcs=2, tau=0.05)
    psi = _field(cfg)
    n = 25
    say("batched rk45 bitwise", n)
    st = oracle.make_state(cfg.dt)
    want, _ = oracle.run(cfg, st, psi, n, NT)
    with Batch() as b:
        i = b.add(cfg, psi, n)
        b.run_to_completion(chunk=8)
        got = b.download(i)
        q = b.query(i)
    say("equal sites %.6f, t %.17g vs %.17g" % (np.mean(got == want), q["t"], st.t))
    assert _bits_equal(got, want)
    assert q["t"] == st.t and q["dt"].real == st.dt_re
