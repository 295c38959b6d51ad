"""ctypes binding of libgpehip.so (the C ABI declared in include/gpe_hip.h).

There is no fallback: if the library is not built, importing a function
raises.  Build it with `python -m gpe_amd.build`.
"""
import ctypes
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPE_LIBRARY") or os.path.join(HERE, "libgpehip.so")

GPE_OK = 0
GPE_EFULL = -7
GPE_EIO = -8
GPE_EULER, GPE_RK2, GPE_RK4, GPE_RK45 = 1, 2, 4, 45
GPE_FP64, GPE_FP32, GPE_FP32C = 0, 1, 2
GPE_HOST_C128, GPE_HOST_C64 = 0, 1
GPE_PROF_KINDS = 24

SCHEME_CODES = {"euler": GPE_EULER, "rk2": GPE_RK2, "rk4": GPE_RK4, "rk45": GPE_RK45}
PRECISION_CODES = {"fp64": GPE_FP64, "fp32": GPE_FP32, "fp32c": GPE_FP32C}

# Stage kinds, in the order of gpe::Kind (csrc/stage_args.h).
KIND_NAMES = ["rhs", "euler", "rk2_1", "rk2_2", "rk4_1", "rk4_2", "rk4_3", "rk4_4",
              "ck_1", "ck_2", "ck_3", "ck_4", "ck_5", "ck_6", "energy", "-",
              # fused stage pairs (ids 16..20)
              "rk4_12", "rk4_34", "ck_12", "ck_34", "ck_56"]
STAGE_KINDS = KIND_NAMES[:15]
PAIR_KINDS = KIND_NAMES[16:]

class GpeLineParam(ctypes.Structure):
    """struct gpe_line_param (include/gpe_hip.h)."""
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
                ("amp1", ctypes.c_double), ("amp2", ctypes.c_double), ("ll", ctypes.c_double),
                ("sgn", ctypes.c_double), ("dir", ctypes.c_int32), ("imprint_phase", ctypes.c_int32)]


class GpeCondensed(ctypes.Structure):
    """struct gpe_condensed (include/gpe_hip.h): one mass.dat record."""
    _fields_ = [(n, ctypes.c_double) for n in ("time", "im_t", "density", "n0_per_site", "mass", "n0", "temp",
                                                "temp2", "energy_per_site")]


# struct gpe_zero_point (include/gpe_hip.h)
ZERO_DTYPE = np.dtype([("section", np.int32), ("reserved", np.int32), ("cell", np.int64), ("x", np.float64),
                       ("y", np.float64), ("z", np.float64)])
ZERO_FILES = {"zeros": 0, "re_zeros": 1, "im_zeros": 2}

# Exported entry points (must match include/gpe_hip.h).
EXPORTS = [
    "gpe_last_error", "gpe_ctx_create", "gpe_ctx_destroy", "gpe_solver_step", "gpe_solver_step_ex",
    "gpe_transfer_bytes", "gpe_sim_imprint_vortex_line", "gpe_sim_random_phase", "gpe_random_phase_prepare", "gpe_random_phase_modes", "gpe_batch_add_random_phase",
    "gpe_sim_upload", "gpe_sim_download", "gpe_sim_set_state", "gpe_sim_get_state",
    "gpe_sim_advance", "gpe_sim_norm_history", "gpe_sim_norm", "gpe_sim_energy",
    "gpe_sim_mass", "gpe_sim_hamiltonian", "gpe_sim_linelength", "gpe_sim_zeros", "gpe_sim_condensed_particles", "gpe_sim_spectrum", "gpe_sim_filtered", "gpe_sim_renormalise", "gpe_sim_rhs", "gpe_prof_enable", "gpe_prof_query",
    "gpe_tune_set", "gpe_set_fusion", "gpe_tune_get", "gpe_tune_key", "gpe_tune_load", "gpe_autotune", "gpe_probe_bandwidth",
    "gpe_batch_create", "gpe_batch_destroy", "gpe_batch_footprint", "gpe_batch_add",
    "gpe_batch_remove", "gpe_batch_run", "gpe_batch_sync", "gpe_batch_poll", "gpe_batch_query",
    "gpe_batch_download", "gpe_batch_stats", "gpe_batch_use_graph", "gpe_batch_tune_set",
    "gpe_batch_tune_get", "gpe_batch_autotune", "gpe_energy_noalloc",
    "gpe_serve_create", "gpe_serve_destroy", "gpe_serve_submit", "gpe_serve_step", "gpe_serve_poll",
    "gpe_serve_result",
]


class GpeParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("eqn", ctypes.c_int32), ("bcs", ctypes.c_int32),
        ("order", ctypes.c_int32), ("precision", ctypes.c_int32), ("real_time", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("g", ctypes.c_double), ("mu", ctypes.c_double),
        ("omx", ctypes.c_double), ("omy", ctypes.c_double), ("omz", ctypes.c_double),
        ("urhs", ctypes.c_double), ("diss_amp", ctypes.c_double),
        ("eps", ctypes.c_double), ("safety", ctypes.c_double),
        ("dt_increase", ctypes.c_double), ("dt_decrease", ctypes.c_double),
    ]


class GpeStepState(ctypes.Structure):
    _fields_ = [
        ("dt_re", ctypes.c_double), ("dt_im", ctypes.c_double),
        ("t", ctypes.c_double), ("im_t", ctypes.c_double),
        ("errmax", ctypes.c_double), ("n_rejects", ctypes.c_int32), ("p", ctypes.c_int32),
    ]


class GpeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gpe error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load libgpehip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libgpehip.so is not built (run `python -m gpe_amd.build`); "
                           "there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    P, S = ctypes.POINTER(GpeParams), ctypes.POINTER(GpeStepState)
    sig = {
        "gpe_last_error": ([], ctypes.c_char_p),
        "gpe_ctx_create": ([P, ctypes.POINTER(vp)], ctypes.c_int),
        "gpe_ctx_destroy": ([vp], None),
        "gpe_solver_step": ([vp, vp, i32, vp, S], ctypes.c_int),
        "gpe_solver_step_ex": ([vp, vp, i32, vp, S, i32], ctypes.c_int),
        "gpe_transfer_bytes": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], ctypes.c_int),
        "gpe_sim_random_phase": ([vp, ctypes.c_uint64, ctypes.c_double], ctypes.c_int),
        "gpe_random_phase_prepare": ([i32, i32, i32], ctypes.c_int),
        "gpe_sim_imprint_vortex_line": ([vp, ctypes.POINTER(GpeLineParam), ctypes.c_double], ctypes.c_int),
        "gpe_random_phase_modes": ([i32, i32, i32] + [ctypes.c_double] * 4 + [ctypes.c_uint64, vp, vp, i64,
                                                                                ctypes.POINTER(i64)], ctypes.c_int),
        "gpe_batch_add_random_phase": ([vp, P, ctypes.c_uint64, ctypes.c_double, S, i64,
                                        ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_sim_upload": ([vp, vp, i32], ctypes.c_int),
        "gpe_sim_download": ([vp, vp, i32], ctypes.c_int),
        "gpe_sim_set_state": ([vp, S], ctypes.c_int),
        "gpe_sim_get_state": ([vp, S], ctypes.c_int),
        "gpe_sim_advance": ([vp, i64, i32], ctypes.c_int),
        "gpe_sim_norm_history": ([vp, dp, i64], ctypes.c_int),
        "gpe_sim_norm": ([vp, dp], ctypes.c_int),
        "gpe_sim_energy": ([vp, dp], ctypes.c_int),
        "gpe_sim_mass": ([vp, dp], ctypes.c_int),
        "gpe_sim_linelength": ([vp, dp], ctypes.c_int),
        "gpe_sim_hamiltonian": ([vp, dp], ctypes.c_int),
        "gpe_sim_zeros": ([vp, i32, vp, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "gpe_sim_condensed_particles": ([vp, ctypes.POINTER(GpeCondensed)], ctypes.c_int),
        "gpe_sim_spectrum": ([vp, ctypes.POINTER(i32), dp, dp, i32], ctypes.c_int),
        "gpe_sim_filtered": ([vp, ctypes.c_double, ctypes.c_double, vp, i32, dp], ctypes.c_int),
        "gpe_sim_renormalise": ([vp, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "gpe_sim_rhs": ([vp, vp, i32], ctypes.c_int),
        "gpe_prof_enable": ([vp, i32], ctypes.c_int),
        "gpe_prof_query": ([vp, i32, ctypes.POINTER(i64), dp, dp, ctypes.c_char_p], ctypes.c_int),
        "gpe_tune_set": ([vp, i32, i32, i32], ctypes.c_int),
        "gpe_set_fusion": ([vp, i32], ctypes.c_int),
        "gpe_tune_get": ([vp, i32] + [ctypes.POINTER(i32)] * 4, ctypes.c_int),
        "gpe_tune_key": ([vp, ctypes.c_char_p, i64], ctypes.c_int),
        "gpe_tune_load": ([vp, ctypes.c_char_p, ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_autotune": ([vp, i32], ctypes.c_int),
        "gpe_probe_bandwidth": ([i64, i32, dp, dp, dp], ctypes.c_int),
        "gpe_batch_create": ([i64, ctypes.POINTER(vp)], ctypes.c_int),
        "gpe_batch_destroy": ([vp], None),
        "gpe_batch_footprint": ([P, ctypes.POINTER(i64)], ctypes.c_int),
        "gpe_batch_add": ([vp, P, vp, i32, S, i64, ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_batch_remove": ([vp, i32], ctypes.c_int),
        "gpe_batch_run": ([vp, i32], ctypes.c_int),
        "gpe_batch_sync": ([vp], ctypes.c_int),
        "gpe_batch_poll": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32), i32, ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_batch_query": ([vp, i32, S, ctypes.POINTER(i64), ctypes.POINTER(i32), dp, ctypes.POINTER(i32)],
                            ctypes.c_int),
        "gpe_batch_download": ([vp, i32, vp, i32], ctypes.c_int),
        "gpe_batch_stats": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i64),
                             ctypes.POINTER(i64)], ctypes.c_int),
        "gpe_batch_use_graph": ([vp, i32], ctypes.c_int),
        "gpe_batch_tune_set": ([vp, i32, i32, i32, i32, i32, i32], ctypes.c_int),
        "gpe_batch_tune_get": ([vp, i32, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_batch_autotune": ([vp, i32], ctypes.c_int),
        "gpe_energy_noalloc": ([P, vp, dp, vp], ctypes.c_int),
        "gpe_serve_create": ([i64, i64, ctypes.POINTER(vp)], ctypes.c_int),
        "gpe_serve_destroy": ([vp], None),
        "gpe_serve_submit": ([vp, P, vp, i32, ctypes.c_uint64, ctypes.c_double, S, i64, ctypes.POINTER(i64)],
                             ctypes.c_int),
        "gpe_serve_step": ([vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_serve_poll": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i32), i32, ctypes.POINTER(i32)], ctypes.c_int),
        "gpe_serve_result": ([vp, i64, vp, i32, S], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def probe_bandwidth(nbytes=1 << 31, reps=20):
    """Measured HBM read / write / copy GB/s of plain float4 kernels."""
    r, w, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().gpe_probe_bandwidth(int(nbytes), int(reps), ctypes.byref(r), ctypes.byref(w), ctypes.byref(c)))
    return {"read": r.value, "write": w.value, "copy": c.value}


def check(rc):
    if rc != GPE_OK:
        msg = lib().gpe_last_error()
        raise GpeError(rc, msg.decode() if msg else "")
    return rc
