"""RKF45 on a field with compact support (zero outside the Thomas-Fermi
radius of a tight trap), where the reference's error scale
|psi| + dt|f(psi)| + 1e-30 (solve.f90:176-178) is 1e-30 at the sites the
later stages reach first.  The reference shrinks dt until abs(err/scale) <= 1
(solve.f90:184) -- 12 rejections to dt = 2.5e-11 in double precision, 16 to
7e-16 in single -- and stays finite; its complex division and abs scale
their operands, so |scale|^2 never underflows.  The GPU forms must do the
same: FP64 and plain FP32 against the oracle, the compensated form finite
with a comparable step, alone and in a batch."""
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _cfg(precision="fp64"):
    n, dx = 48, 0.25
    xr = dx * (n - 1) / 2
    return RunConfig(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, scheme="rk45", order=4, bcs=2, tau=0.01, eqn=4, g=1.0,
                     mu=1.0, omx=0.6, omy=0.6, omz=0.85, eps=1e-6, precision=precision)


def _psi(cfg):
    psi = ic.line_pair(cfg, sep=3.0, mu_ic=10.0)
    assert (psi == 0).sum() > 10000
    return psi


def _run(cfg, psi, n):
    from gpe_amd.sim import Simulation
    out = []
    with Simulation(cfg, psi) as sim:
        for _ in range(n):
            sim.advance(1)
            out.append(sim.state())
        return sim.download(np.complex64 if cfg.precision == "fp32" else np.complex128), out


def _oracle(oracle, cfg, psi, n):
    st = oracle.make_state(cfg.dt)
    hist = []
    for _ in range(n):
        psi = oracle.solver(cfg, st, psi)
        hist.append((st.dt_re, st.t, st.n_rejects))
    return psi, hist


def test_fp64_matches_oracle(oracle):
    c = _cfg()
    psi = _psi(c)
    say("rk45 compact support fp64")
    got, sts = _run(c, psi, 3)
    want, hist = _oracle(oracle, c, psi, 3)
    assert [s["n_rejects"] for s in sts] == [h[2] for h in hist] and hist[0][2] >= 10
    for s, h in zip(sts, hist):
        assert s["dt"].real == pytest.approx(h[0], rel=1e-12)
        assert s["t"] == pytest.approx(h[1], rel=1e-12)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


def test_fp32_matches_single_precision_oracle(oracle):
    c = _cfg("fp32")
    psi = _psi(c).astype(np.complex64)
    say("rk45 compact support fp32")
    got, sts = _run(c, psi, 3)
    want, hist = _oracle(oracle, c, psi, 3)
    assert np.isfinite(got).all()
    assert sts[0]["n_rejects"] == hist[0][2] and hist[0][2] >= 10
    for s, h in zip(sts, hist):
        assert s["dt"].real == pytest.approx(h[0], rel=1e-4)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


def test_fp32c_finite_alone_and_batched():
    from gpe_amd.batch import Batch
    c = _cfg("fp32c")
    psi = _psi(c)
    say("rk45 compact support fp32c")
    got, sts = _run(c, psi, 3)
    assert np.isfinite(got).all()
    assert all(0 < s["dt"].real < 1e-8 for s in sts)
    with Batch() as b:
        i = b.add(c, psi, 3)
        st = b.run_to_completion(chunk=4)
        q = b.query(i)
        out = b.download(i)
    assert st[i] == 1 and q["steps"] == 3
    assert np.isfinite(out).all()
    assert q["t"] == pytest.approx(sts[-1]["t"], rel=1e-6)
