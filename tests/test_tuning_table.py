"""The stored launch-configuration table (gpe_tune_key / gpe_tune_load).

The table is written by the Python autotuner and read by the library, so C
and Fortran callers apply the same entries.  Launch configurations change
speed only: an RK4 field advanced under any stored configuration is bitwise
the field of the default one (the stage updates are pointwise)."""
import json
import os
import subprocess

import numpy as np
import pytest

from gpe_amd import _lib, ic, tuning
from gpe_amd.config import RunConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN_D = os.path.join(ROOT, "oracle", "_ref", "gpe_dropin_d")

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    base = dict(nx=24, ny=20, nz=18, xr=4.6, yr=3.8, zr=3.4, eqn=4, g=1.0, mu=1.0, omx=0.3, omy=0.2, omz=0.4, bcs=2,
                order=4, scheme="rk4", tau=0.01)
    base.update(kw)
    return RunConfig(**base)


ODD = {"rk4_1": [4, 8], "rk4_2": [2, 4], "rk4_3": [3, 0], "rk4_4": [0, 6]}


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
@pytest.mark.parametrize("scheme,bcs,order", [("rk4", 2, 4), ("rk45", 1, 2), ("euler", 2, 2)])
def test_key_matches_python_key(precision, scheme, bcs, order):
    from gpe_amd.sim import Simulation
    cfg = _cfg(precision=precision, scheme=scheme, bcs=bcs, order=order)
    print("[case] key", precision, scheme, bcs, order, flush=True)
    with Simulation(cfg, ic.vortex_ring(cfg, 1.5), tuned=False) as sim:
        assert sim.tune_key() == tuning.key(cfg)
        if precision != "fp64":
            sim.set_fusion(False)
            assert sim.tune_key() == tuning.key(cfg, fused=False)


def test_load_applies_entry_and_keeps_results(tmp_path):
    from gpe_amd.sim import Simulation
    cfg = _cfg()
    psi = ic.vortex_ring(cfg, 1.5) * (1 + 0.02 * np.random.default_rng(5).standard_normal(cfg.shape))
    path = str(tmp_path / "t.json")
    other = dict(ODD, rk4_1=[1, 0])
    # a key of another problem, an unknown kind name (skipped) and the entry
    json.dump({"8x8x8/o2/fp64/rk4/b1": other, tuning.key(cfg): dict(ODD, future_kind=[9, 9])}, open(path, "w"))
    print("[case] load", flush=True)
    with Simulation(cfg, psi, tuned=False) as ref:
        ref.advance(4)
        want = ref.download()
    with Simulation(cfg, psi, tuned=False) as sim:
        assert sim.tune_load(path)
        got_cfg = sim.tune_get()
        for k, (v, z) in ODD.items():
            assert got_cfg[k][:2] == (v, z)
        sim.advance(4)
        assert np.array_equal(sim.download(), want)


def test_load_skips_other_json_values(tmp_path):
    # other keys may hold any JSON value (escaped strings, nested objects,
    # exponents, true/false/null): the reader skips them to find the entry
    from gpe_amd.sim import Simulation
    cfg = _cfg()
    path = tmp_path / "mixed.json"
    path.write_text('{"note": "a \\"quoted\\" {not, an} [entry]", "meta": {"a": [1, 2.5e-3, -4E+2, true, false, '
                    'null], "b": {"c": []}, "d": {}}, "%s": {"rk4_1": [4, 8], "rk4_2": [2, 4]}, "tail": [[], {}]}'
                    % tuning.key(cfg))
    json.loads(path.read_text())  # the file is valid JSON
    print("[case] mixed table", flush=True)
    with Simulation(cfg, ic.vortex_ring(cfg, 1.5), tuned=False) as sim:
        assert sim.tune_load(str(path))
        got = sim.tune_get()
        assert got["rk4_1"][:2] == (4, 8) and got["rk4_2"][:2] == (2, 4)


def test_load_missing_and_malformed(tmp_path):
    from gpe_amd.sim import Simulation
    cfg = _cfg()
    print("[case] missing/malformed", flush=True)
    with Simulation(cfg, ic.vortex_ring(cfg, 1.5), tuned=False) as sim:
        before = sim.tune_get()
        assert not sim.tune_load(str(tmp_path / "absent.json"))
        json.dump({"8x8x8/o2/fp64/rk4/b1": ODD}, open(tmp_path / "other.json", "w"))
        assert not sim.tune_load(str(tmp_path / "other.json"))
        (tmp_path / "bad.json").write_text('{"%s": {"rk4_1": [1, }' % tuning.key(cfg))
        with pytest.raises(_lib.GpeError) as e:
            sim.tune_load(str(tmp_path / "bad.json"))
        assert e.value.code == _lib.GPE_EIO
        # an out-of-range variant is rejected and nothing of the entry is kept
        json.dump({tuning.key(cfg): dict(ODD, rk4_4=[99, 0])}, open(tmp_path / "range.json", "w"))
        with pytest.raises(_lib.GpeError) as e:
            sim.tune_load(str(tmp_path / "range.json"))
        assert "rk4_4" in str(e.value)
        assert sim.tune_get() == before
        # a value beyond int32 is rejected, not wrapped (4294967297 would wrap to 1)
        for bad in ([4294967297, 0], [1, -4294967296]):
            json.dump({tuning.key(cfg): dict(ODD, rk4_3=bad)}, open(tmp_path / "wrap.json", "w"))
            with pytest.raises(_lib.GpeError) as e:
                sim.tune_load(str(tmp_path / "wrap.json"))
            assert e.value.code == _lib.GPE_EIO and "rk4_3" in str(e.value)
            assert sim.tune_get() == before
        # FP64 Cash-Karp stages refuse the (4,4) and (8,2) tiles
        c45 = cfg.with_(scheme="rk45")
        with Simulation(c45, ic.vortex_ring(c45, 1.5), tuned=False) as s45:
            for v in (0, 1):
                with pytest.raises(_lib.GpeError):
                    s45.tune_set("ck_6", v, 0)
            assert s45.tune_get()["ck_6"][0] == 4
            s45.tune_set("ck_6", 3, 0)


@pytest.mark.skipif(not os.path.exists(BIN_D), reason="drop-in harness not built (make -C oracle ref)")
def test_fortran_dropin_reads_the_table(tmp_path):
    from test_dropin import _run_mode
    from gpe_amd.state_io import read_state, write_run_in, write_state
    cfg = _cfg()
    psi = ic.vortex_ring(cfg, 1.5) * (1 + 0.02 * np.random.default_rng(6).standard_normal(cfg.shape))
    good, rng_bad = tmp_path / "good.json", tmp_path / "range.json"
    json.dump({tuning.key(cfg): ODD}, open(good, "w"))
    json.dump({tuning.key(cfg): dict(ODD, rk4_2=[99, 0])}, open(rng_bad, "w"))
    print("[case] drop-in tuning table", flush=True)
    runs = {}
    for name, table in (("default", tmp_path / "absent.json"), ("tuned", good)):
        (tmp_path / name).mkdir()
        d, _, _ = _run_mode(tmp_path / name, BIN_D, cfg, psi, 5, None, False, GPE_TUNING_FILE=str(table))
        runs[name] = read_state(str(d / "out.dat"))[0]
    assert np.array_equal(runs["tuned"], runs["default"])
    # the entry is found and applied: an out-of-range variant in it stops the
    # run through emergency_stop (error.f90:24-46) with the library's message
    d = tmp_path / "bad"
    d.mkdir()
    write_run_in(str(d / "run.in"), cfg, 4.5 * cfg.tau)
    write_state(str(d / "end_state0000.dat"), psi, dt=cfg.dt)
    open(d / "RUNNING", "w").close()
    env = dict(os.environ, GPE_TUNING_FILE=str(rng_bad))
    subprocess.run([BIN_D, "run.in", "end_state0000.dat", "out.dat"], cwd=str(d), capture_output=True, text=True,
                   timeout=120, env=env)
    # list-directed output (error.f90:40) wraps the message at 80 columns
    err = (d / "ERROR").read_text().replace("\n ", "")
    assert "kind rk4_2: variant out of range" in err and "range.json" in err
    assert not (d / "out.dat").exists()
