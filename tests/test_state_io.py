"""The reference's file formats: end_state (io.f90:843-853) and the 3-D
dumps of idl_surface (io.f90:480-520), byte layout checked field by field;
and (GPU) batch checkpoints, whose restart continues bitwise as if the run
had not stopped."""
import os

import numpy as np
import pytest

from gpe_amd import ic, state_io
from gpe_amd.config import RunConfig


def _cfg(**kw):
    base = dict(nx=12, ny=10, nz=9, xr=2.2, yr=1.8, zr=1.6)
    base.update(kw)
    return RunConfig(**base)


def test_itos():
    assert state_io.itos(0) == "0000000"
    assert state_io.itos(123) == "0000123"
    assert state_io.itos(12345678) == "2345678"  # seven digits kept


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_dens_layout(tmp_path, dtype):
    c = _cfg(precision="fp64" if dtype == np.complex128 else "fp32")
    rng = np.random.default_rng(1)
    psi = (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)).astype(dtype)
    proc = str(tmp_path / "proc0000")
    path = state_io.write_dens(proc, psi, c, p=42, t=1.5, im_t=0.25)
    assert os.path.basename(path) == "dens0000042.dat"
    raw = open(path, "rb").read()
    rs = 8 if dtype == np.complex128 else 4
    rt = np.float64 if rs == 8 else np.float32
    assert np.frombuffer(raw[:rs], dtype=rt)[0] == rt(1.75)
    ints = np.frombuffer(raw[rs:rs + 36], dtype=np.int32)
    assert list(ints) == [12, 10, 9, 1, 1, 0, 9, 0, 8]
    n = 12 * 10 * 9
    body = np.frombuffer(raw[rs + 36:rs + 36 + 2 * rs * n], dtype=dtype).reshape(c.shape)
    assert np.array_equal(body, psi)
    x, y, z = c.coords()
    tail = np.frombuffer(raw[rs + 36 + 2 * rs * n:], dtype=rt)
    assert np.array_equal(tail, np.concatenate([x, y, z]).astype(rt))
    got = state_io.read_dens(path, single=(rs == 4))
    assert np.array_equal(got[0], psi) and got[1] == pytest.approx(1.75)
    assert open(tmp_path / "p_saved.dat").read() == "%10d\n" % 42


def test_end_state_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((9, 10, 12)) + 1j * rng.standard_normal((9, 10, 12))
    p = str(tmp_path / "end_state.dat")
    state_io.write_state(p, psi, p=7, t=0.5, im_t=0.0, dt=0.01 + 0j)
    raw = open(p, "rb").read()
    assert list(np.frombuffer(raw[:16], dtype=np.int32)) == [12, 10, 9, 7]
    back, pp, t, im_t, dt = state_io.read_state(p)
    assert np.array_equal(back, psi) and (pp, t, im_t, dt) == (7, 0.5, 0.0, 0.01 + 0j)


@pytest.mark.gpu
def test_batch_checkpoint_restart_bitwise(tmp_path):
    from gpe_amd.batch import Batch
    cases = [
        (_cfg(nx=24, ny=20, nz=18, scheme="rk4", order=2, bcs=2, omx=0.1, omy=0.1, omz=0.2, tau=0.005), 14),
        (_cfg(nx=33, ny=21, nz=16, scheme="rk45", order=4, bcs=1, tau=0.05, precision="fp32c"), 9),
        (_cfg(nx=20, ny=20, nz=20, scheme="rk45", order=2, bcs=2, omx=0.2, omy=0.2, omz=0.3, tau=0.5), 7),
    ]
    psis = [ic.vortex_ring(c, 0.3 * c.nx * c.dx) for c, _ in cases]
    print("[case] batch checkpoint/restart", flush=True)
    with Batch() as b:
        ids = [b.add(c, p, n) for (c, n), p in zip(cases, psis)]
        b.run_to_completion(chunk=3)
        want = [(b.download(i), b.query(i)) for i in ids]
    ck = str(tmp_path / "ck")
    with Batch() as b:
        ids = [b.add(c, p, n) for (c, n), p in zip(cases, psis)]
        b.run(4)
        man = b.checkpoint(ck)
        assert len(man) == 3
    with Batch() as b:
        new = b.restore(ck)
        b.run_to_completion(chunk=3)
        got = [(b.download(new[i]), b.query(new[i])) for i in ids]
    for (w, qw), (g, qg) in zip(want, got):
        assert np.array_equal(w, g)
        assert qw["t"] == qg["t"] and qw["dt"] == qg["dt"]


def test_filtered_file_is_the_dens_layout_with_t(tmp_path):
    c = _cfg()
    psi = np.arange(c.nx * c.ny * c.nz, dtype=np.float64).reshape(c.shape) * (1 - 0.5j)
    path = state_io.write_filtered(str(tmp_path / "proc0000"), psi, c, p=7, t=2.5)
    assert os.path.basename(path) == "filtered0000007.dat"
    got = state_io.read_dens(path)
    assert np.array_equal(got[0], psi) and got[1] == 2.5  # t, not t + im_t (io.f90:588)
    assert not (tmp_path / "p_saved.dat").exists()


def test_mass_and_spectrum_records(tmp_path):
    rec = dict(zip(state_io.MASS_FIELDS, [1.5, 0.0, 0.98, 2.8e-5, 1.2e6, 3.7e3, 0.64, -1.0e-3, 0.66]))
    state_io.append_mass(str(tmp_path / "mass.dat"), rec)
    state_io.append_mass(str(tmp_path / "mass.dat"), rec)
    lines = open(tmp_path / "mass.dat").read().splitlines()
    # '(9e17.9)': nine 17-character fields
    assert len(lines) == 2 and len(lines[0]) == 9 * 17
    assert [float(lines[0][17 * q:17 * q + 17]) for q in range(9)] == pytest.approx(list(rec.values()), rel=1e-8)
    assert lines[0][:17] == "  0.150000000E+01"
    ps, pi, pe = state_io.write_spectrum(str(tmp_path), 12, 0.5, [3, 10, 0, 0, 0, 0, 0],
                                         [1.25, 0.5, 0, 0, 0, 0, 0], [1.0, 2.0, 4.0])
    assert ps.endswith(os.path.join("spectrum", "spectrum0000012.dat"))
    sl = open(ps).read().splitlines()
    # '(2i9,e17.9)'
    assert len(sl) == 7 and sl[0] == "        1        3  0.125000000E+01"
    il = open(pi).read().splitlines()
    assert il == ["        1  0.100000000E+01", "        2  0.200000000E+01", "        3  0.400000000E+01"]
    # '(4e17.9)': t, eta(1:3)
    assert open(pe).read() == "  0.500000000E+00  0.125000000E+01  0.500000000E+00  0.000000000E+00\n"
