"""The C-ABI library loads and exports every entry point include/gpe_hip.h
declares (no GPU needed: nothing here launches work)."""
import ctypes
import os
import re

from gpe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(gpe_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"gpe_ctx_create", "gpe_solver_step", "gpe_sim_advance", "gpe_last_error"} <= names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, "not exported: %s" % missing


def test_python_binding_covers_header():
    assert declared_symbols() == set(_lib.EXPORTS)


def test_struct_layouts_match_header():
    # gpe_params: 10 int32 then 17 doubles; gpe_step_state: 5 doubles, 2 int32
    assert ctypes.sizeof(_lib.GpeParams) == 10 * 4 + 17 * 8
    assert ctypes.sizeof(_lib.GpeStepState) == 5 * 8 + 2 * 4
    # gpe_line_param: 7 doubles, 2 int32; gpe_zero_point: 2 int32, int64, 3 doubles
    assert ctypes.sizeof(_lib.GpeLineParam) == 7 * 8 + 2 * 4
    assert _lib.ZERO_DTYPE.itemsize == 2 * 4 + 8 + 3 * 8
    assert ctypes.sizeof(_lib.GpeCondensed) == 9 * 8


def test_invalid_params_rejected_without_gpu_work():
    # parameter validation happens before any device call
    from gpe_amd.sim import make_params
    from gpe_amd.config import RunConfig
    L = _lib.lib()
    p = make_params(RunConfig(nx=4, ny=8, nz=8))
    ctx = ctypes.c_void_p()
    rc = L.gpe_ctx_create(ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -1 and not ctx
    assert b"at least 5" in L.gpe_last_error()
    p = make_params(RunConfig(nx=8, ny=8, nz=8))
    p.scheme = 3
    assert L.gpe_ctx_create(ctypes.byref(p), ctypes.byref(ctx)) == -1
    assert b"Unrecognised time stepping scheme" in L.gpe_last_error()


def test_new_entry_points_reject_bad_arguments_without_gpu_work():
    L = _lib.lib()
    n = ctypes.c_int64()
    i64 = ctypes.c_int64
    # NULL context / outputs
    assert L.gpe_sim_zeros(None, 0, None, 0, ctypes.byref(n)) == -1
    assert L.gpe_sim_random_phase(None, ctypes.c_uint64(1), 0.5) == -1
    assert L.gpe_sim_imprint_vortex_line(None, ctypes.byref(_lib.GpeLineParam()), 1.0) == -1
    assert L.gpe_sim_condensed_particles(None, ctypes.byref(_lib.GpeCondensed())) == -1
    assert L.gpe_sim_spectrum(None, None, None, None, 0) == -1
    assert L.gpe_sim_filtered(None, 1.0, 1.0, None, 0, None) == -1
    assert L.gpe_tune_load(None, None, None) == -1
    assert L.gpe_transfer_bytes(None, ctypes.byref(i64()), ctypes.byref(i64())) == -1
    st = _lib.GpeStepState()
    assert L.gpe_solver_step_ex(None, None, 2, None, ctypes.byref(st), 3) == -1
    # the mode list is host-only: bad grids and kmax are refused, a good one counts
    for args in [(0, 8, 8, 0.5, 0.5, 0.5, 0.5), (8, 8, 8, -0.5, 0.5, 0.5, 0.5), (8, 8, 8, 0.5, 0.5, 0.5, -1.0)]:
        assert L.gpe_random_phase_modes(*args, ctypes.c_uint64(1), None, None, 0, ctypes.byref(n)) == -1
    assert b"random phase" in L.gpe_last_error()
    assert L.gpe_random_phase_modes(8, 8, 8, 0.5, 0.5, 0.5, 0.0, ctypes.c_uint64(1), None, None, 0,
                                    ctypes.byref(n)) == 0
    assert n.value == 1  # kmax = 0: the zero mode only
    assert L.gpe_random_phase_modes(8, 8, 8, 0.5, 0.5, 0.5, 0.5, ctypes.c_uint64(1), None, None, 4,
                                    ctypes.byref(n)) == -1  # cap without buffers
