"""Malformed input to the host-side readers (CPU).

* The launch-configuration table reader (gpe_amd/csrc/tune_table.cpp) is
  built for the host with AddressSanitizer and UndefinedBehaviorSanitizer
  (g++, no device code) and fed every prefix of a valid table, a thousand
  seeded byte corruptions of it, deep nesting, huge and negative numbers,
  unterminated strings, a trailing escape and non-ASCII bytes.  Any memory
  error aborts the harness; every input must give GPE_OK or GPE_EIO, and the
  valid table its entry.
* The end_state / dens readers (gpe_amd/state_io.py) get truncated files,
  corrupted headers and size mismatches: each must raise ValueError before
  reading the field (no over-allocation, no misread).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpe_amd import state_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpe_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <vector>
#include "host_common.h"
int main(int argc, char **argv) {
  for (int i = 2; i < argc; ++i) {
    std::vector<gpe::TuneItem> out;
    bool found = false;
    int rc = gpe::read_tune_table(argv[i], argv[1], out, found);
    std::printf("%d %d %zu", rc, (int)found, out.size());
    for (const auto &it : out) {
      std::printf(" ");
      for (unsigned char ch : it.name) {
        if (ch > 32 && ch < 127) std::printf("%c", ch);
        else std::printf("\\x%02x", ch);
      }
      std::printf(":%d:%d", it.variant, it.zchunk);
    }
    std::printf("\n");
  }
  return 0;
}
"""

VALID = ('{"64x64x64/o2/fp64/rk4/b2": {"rk4_1": [3, 0], "rk4_2": [1, 64]},\n'
         ' "k": {"a\\"b": [1, 2], "ck_6": [4, 512], "energy": [0, -1]},\n'
         ' "x": [1, 2.5e3, true, null, {"y": "z\\\\"}]}')


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("asan")
    (d / "h.cpp").write_text(HARNESS)
    exe = str(d / "h")
    subprocess.check_call(["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
                           "-I/opt/rocm/include", "-I" + CSRC, str(d / "h.cpp"), os.path.join(CSRC, "tune_table.cpp"),
                           os.path.join(CSRC, "errors.cpp"), "-o", exe])
    return exe, d


def _run(harness, key, blobs):
    exe, d = harness
    paths = []
    for i, b in enumerate(blobs):
        p = d / ("in%05d.json" % i)
        p.write_bytes(b)
        paths.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, key] + paths, capture_output=True, text=True, errors="replace", env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = r.stdout.splitlines()
    assert len(lines) == len(blobs)
    return [ln.split() for ln in lines]


def test_valid_table(harness):
    (rc, found, n, *items), = _run(harness, "k", [VALID.encode()])
    assert (rc, found, n) == ("0", "1", "3")
    assert items == ['a"b:1:2', "ck_6:4:512", "energy:0:-1"]


def test_prefixes_and_corruptions(harness):
    v = VALID.encode()
    blobs = [v[:i] for i in range(len(v))]
    rng = np.random.default_rng(20261202)
    for _ in range(1000):
        b = bytearray(v)
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(len(b)))] = int(rng.integers(256))
        blobs.append(bytes(b))
    for rc, *_ in _run(harness, "k", blobs):
        assert rc in ("0", "-8")


def test_hostile_values(harness):
    blobs = [
        b"", b"   ", b"{", b"}", b"[]", b"{}", b'{"k"', b'{"k":', b'{"k": {', b'{"k": {"a": [',
        b'{"z": ' + b"[" * 200000 + b"]" * 200000 + b', "k": {"a": [1, 2]}}',  # nesting deeper than the limit
        b'{"z": ' + b"[" * 60 + b"]" * 60 + b', "k": {"a": [1, 2]}}',  # within it
        b'{"k": {"a": [99999999999999999999999, 0]}}', b'{"k": {"a": [4294967297, 0]}}',
        b'{"k": {"a": [-2147483649, 0]}}', b'{"k": {"a": [1, 2]', b'{"k": {"a\\', b'{"k": {"\xff\xfe": [1, 2]}}',
        b'{"k": {"a": [1 2]}}', b'{"k": {"a": ["1", 2]}}', b'{"q": "unterminated',
        b"\x00" * 64, bytes(range(256)),
    ]
    res = _run(harness, "k", blobs)
    assert all(r[0] in ("0", "-8") for r in res)
    assert res[10][0] == "-8"                  # too deep: refused, no stack overflow
    assert res[11][:3] == ["0", "1", "1"]      # 60 levels: skipped, entry found
    assert res[12][0] == res[13][0] == res[14][0] == "-8"  # out of int32 range
    assert res[17][:2] == ["0", "1"]           # any bytes in a name


def _state_file(tmp_path, nx=6, ny=5, nz=4):
    psi = (np.arange(nx * ny * nz) * (1 + 0.5j)).reshape(nz, ny, nx)
    p = tmp_path / "end_state.dat"
    state_io.write_state(str(p), psi, p=3, t=0.5, dt=0.01)
    return p, psi


def test_state_reader_malformed(tmp_path):
    p, psi = _state_file(tmp_path)
    good = p.read_bytes()
    got = state_io.read_state(str(p))
    assert np.array_equal(got[0], psi) and got[1] == 3
    bad = tmp_path / "bad.dat"
    cases = [good[:n] for n in (0, 3, 15, 16, 40, len(good) - 1)] + [good + b"x"]
    for nx, ny, nz in ((0, 5, 4), (-6, 5, 4), (6, -5, -4), (2 ** 31 - 1, 2 ** 31 - 1, 2 ** 31 - 1), (7, 5, 4)):
        cases.append(np.array([nx, ny, nz, 3], dtype=np.int32).tobytes() + good[16:])
    for b in cases:
        bad.write_bytes(b)
        with pytest.raises(ValueError):
            state_io.read_state(str(bad))


def test_dens_reader_malformed(tmp_path):
    from gpe_amd.config import RunConfig
    cfg = RunConfig(nx=6, ny=5, nz=4, xr=1.0, yr=1.0, zr=1.0)
    psi = (np.arange(cfg.nx * cfg.ny * cfg.nz) * (1 - 0.25j)).reshape(cfg.shape)
    d = tmp_path / "proc0000"
    d.mkdir()
    state_io.write_dens(str(d), psi, cfg, p=7)
    (f,) = list(d.iterdir())
    good = f.read_bytes()
    assert np.array_equal(state_io.read_dens(str(f))[0], psi)
    bad = tmp_path / "bad.dat"
    head = 8
    cases = [good[:n] for n in (0, 7, 20, head + 36, len(good) - 1)] + [good + b"\0"]
    hdr = np.frombuffer(good[head:head + 36], dtype=np.int32).copy()
    for i, v in ((0, 0), (5, 3), (6, 9), (7, 2), (8, 99), (1, -1)):
        h = hdr.copy()
        h[i] = v
        cases.append(good[:head] + h.tobytes() + good[head + 36:])
    for b in cases:
        bad.write_bytes(b)
        with pytest.raises(ValueError):
            state_io.read_dens(str(bad))
