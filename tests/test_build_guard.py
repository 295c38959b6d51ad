"""The library build (gpe_amd/build.py) never links objects that may come
from two versions of the shared headers (CPU; nothing is compiled here).

The host code and the kernels pass argument structs by layout, so a library
linked from objects compiled against different header versions would load
and run without complaint.  The build holds a lock on its object folder and
refuses to link when a file under csrc/ changed while its units compiled.
"""
import os

import pytest

from gpe_amd import build as b


def _fake(monkeypatch, tmp_path, digests):
    seq = iter(digests)
    compiled, linked = [], []
    monkeypatch.setattr(b, "OBJ", str(tmp_path / "obj"))
    monkeypatch.setattr(b, "_sources_digest", lambda: next(seq))
    monkeypatch.setattr(b, "_compile", lambda unit: compiled.append(unit[1]) or unit[1] + ".o")
    monkeypatch.setattr(b, "_link", lambda objs, verbose, lib_path: linked.append(list(objs)) or "lib")
    return compiled, linked


def test_source_change_during_build_refuses_to_link(monkeypatch, tmp_path):
    compiled, linked = _fake(monkeypatch, tmp_path, ["a", "b"])
    with pytest.raises(RuntimeError, match="changed during the build"):
        b.build(verbose=False)
    assert len(compiled) == len(b.UNITS) and linked == []
    assert os.path.exists(os.path.join(b.OBJ, ".lock"))


def test_stable_sources_link_every_unit(monkeypatch, tmp_path):
    compiled, linked = _fake(monkeypatch, tmp_path, ["a", "a"])
    assert b.build(verbose=False) == "lib"
    assert linked == [[u[1] + ".o" for u in b.UNITS]]


def test_digest_covers_every_source_file():
    names = sorted(os.listdir(b.CSRC))
    assert any(n.endswith(".hpp") for n in names) and any(n.endswith(".h") for n in names)
    d = b._sources_digest()
    assert len(d) == 64 and d == b._sources_digest()
