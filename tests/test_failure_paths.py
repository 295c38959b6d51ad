"""Failure handling of the time-stepping path.

A non-finite field makes the RKF45 error norm NaN.  rkqs (solve.f90:180-229)
has no test for it; the library checks the reduced norm on the host after
the last stage (every kernel runs to its end) and returns GPE_ENONFINITE,
the batch marks the instance "nan" and leaves the others running, and the
Fortran drop-in ends the run through the reference's emergency_stop
(error.f90:24-46): ERROR written, RUNNING removed, no end state.  Nothing
here makes a kernel fault or wait: the conditions are set from the host.
"""
import os
import subprocess

import numpy as np
import pytest

from gpe_amd import _lib, ic
from gpe_amd.config import RunConfig
from gpe_amd.state_io import write_run_in, write_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN_D = os.path.join(ROOT, "oracle", "_ref", "gpe_dropin_d")
GPE_ENONFINITE = -4


def _cfg(**kw):
    base = dict(nx=24, ny=20, nz=18, xr=4.6, yr=3.8, zr=3.4, eqn=4, g=1.0, mu=1.0, omx=0.3, omy=0.2, omz=0.4,
                bcs=2, order=4, scheme="rk45", tau=0.01)
    base.update(kw)
    return RunConfig(**base)


def _nan_field(cfg):
    psi = ic.vortex_ring(cfg, 1.5)
    psi[cfg.nz // 2, cfg.ny // 3, cfg.nx // 4] = complex(np.nan, 0.0)
    return psi


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_nan_state_returns_enonfinite(precision):
    from gpe_amd.sim import Simulation
    cfg = _cfg(precision=precision)
    print("[case] NaN state rk45", precision, flush=True)
    psi = _nan_field(cfg)
    if precision == "fp32":
        psi = psi.astype(np.complex64)
    with Simulation(cfg, psi) as sim:
        before = sim.state()
        with pytest.raises(_lib.GpeError) as e:
            sim.advance(1)
        assert e.value.code == GPE_ENONFINITE
        assert "NaN in the error norm" in str(e.value)
        # the failed call commits no step
        after = sim.state()
        assert after["t"] == before["t"] and after["p"] == before["p"]
    # a finite field of the same problem still steps (no state left behind)
    with Simulation(cfg, ic.vortex_ring(cfg, 1.5)) as sim:
        sim.advance(2)
        assert np.isfinite(sim.norm())


@pytest.mark.gpu
def test_nan_instance_stops_alone_in_a_batch():
    from gpe_amd.batch import Batch
    good = _cfg(precision="fp64")
    print("[case] NaN instance in a batch", flush=True)
    with Batch() as b:
        ok = b.add(good, ic.vortex_ring(good, 1.5), 5)
        bad = b.add(good, _nan_field(good), 5)
        done = b.run_to_completion(chunk=4)
        assert done[bad] == -1 and done[ok] == 1
        assert b.query(bad)["steps"] == 0 and b.query(ok)["steps"] == 5
        got = b.download(ok)
    with Batch() as b:
        i = b.add(good, ic.vortex_ring(good, 1.5), 5)
        b.run_to_completion(chunk=4)
        assert np.array_equal(b.download(i), got)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BIN_D), reason="drop-in harness not built (make -C oracle ref)")
@pytest.mark.parametrize("resident", [False, True])
def test_dropin_nan_state_emergency_stop(tmp_path, resident):
    cfg = _cfg()
    print("[case] drop-in NaN state, resident=%s" % resident, flush=True)
    write_run_in(str(tmp_path / "run.in"), cfg, 4.5 * cfg.tau)
    write_state(str(tmp_path / "end_state0000.dat"), _nan_field(cfg), dt=cfg.dt)
    open(tmp_path / "RUNNING", "w").close()
    env = dict(os.environ)
    env.pop("GPE_RESIDENT", None)
    if resident:
        env["GPE_RESIDENT"] = "1"
    r = subprocess.run([BIN_D, "run.in", "end_state0000.dat", "out.dat"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120, env=env)
    assert (tmp_path / "ERROR").exists(), r.stdout + r.stderr
    # list-directed output (error.f90:40) wraps the message at 80 columns
    err = (tmp_path / "ERROR").read_text().replace("\n ", "")
    assert "NaN in the error norm" in err
    assert not (tmp_path / "RUNNING").exists()
    assert not (tmp_path / "out.dat").exists()
