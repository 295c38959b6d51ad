"""ctypes wrapper of the C oracle (oracle/gpe_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, to compare against.  The product (gpe_amd/) never
imports it.

Fields are numpy arrays of shape (nz, ny, nx), C-contiguous, so x is fastest
as in the reference's psi(0:nx1, js:je, ks:ke).  complex128 selects the FP64
restatement (orc64_*), complex64 the FP32 one (orc32_*).
"""
import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("scheme", ctypes.c_int), ("eqn", ctypes.c_int), ("bcs", ctypes.c_int),
        ("order", ctypes.c_int), ("real_time", ctypes.c_int),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("g", ctypes.c_double), ("mu", ctypes.c_double),
        ("omx", ctypes.c_double), ("omy", ctypes.c_double), ("omz", ctypes.c_double),
        ("urhs", ctypes.c_double), ("diss_amp", ctypes.c_double),
        ("eps", ctypes.c_double), ("safety", ctypes.c_double),
        ("dt_increase", ctypes.c_double), ("dt_decrease", ctypes.c_double),
        ("nthreads", ctypes.c_int),
    ]


class OrcState(ctypes.Structure):
    _fields_ = [
        ("dt_re", ctypes.c_double), ("dt_im", ctypes.c_double),
        ("t", ctypes.c_double), ("im_t", ctypes.c_double),
        ("errmax", ctypes.c_double), ("n_rejects", ctypes.c_int), ("p", ctypes.c_int),
    ]


SCHEMES = {"euler": 1, "rk2": 2, "rk4": 4, "rk45": 45}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
        P, D, I = ctypes.POINTER(OrcParams), ctypes.c_double, ctypes.c_int
        vp = ctypes.c_void_p
        for pfx in ("orc64", "orc32"):
            f = getattr(_LIB, pfx + "_laplacian"); f.argtypes = [P, vp, vp]; f.restype = None
            f = getattr(_LIB, pfx + "_deriv_x"); f.argtypes = [P, vp, vp]; f.restype = None
            f = getattr(_LIB, pfx + "_rhs"); f.argtypes = [P, vp, vp]; f.restype = None
            f = getattr(_LIB, pfx + "_grid"); f.argtypes = [P, vp, vp, vp]; f.restype = None
            f = getattr(_LIB, pfx + "_solver")
            f.argtypes = [P, ctypes.POINTER(OrcState), vp, vp]; f.restype = I
            f = getattr(_LIB, pfx + "_run")
            f.argtypes = [P, ctypes.POINTER(OrcState), vp, I, vp]; f.restype = I
            for nm in ("_norm", "_energy", "_mass", "_linelength"):
                f = getattr(_LIB, pfx + nm); f.argtypes = [P, vp]; f.restype = D
            f = getattr(_LIB, pfx + "_renormalise"); f.argtypes = [P, vp, D, D]; f.restype = None
            f = getattr(_LIB, pfx + "_zeros"); f.argtypes = [P, vp, I, vp, ctypes.c_long]; f.restype = ctypes.c_long
            f = getattr(_LIB, pfx + "_imprint_vortex_line"); f.argtypes = [P, vp, ctypes.POINTER(OrcLine), D]
            f.restype = I
            f = getattr(_LIB, pfx + "_set_control_pow"); f.argtypes = [I]; f.restype = None
            f = getattr(_LIB, pfx + "_control_pow"); f.argtypes = [D, D]; f.restype = D
            f = getattr(_LIB, pfx + "_control_pow_ambiguous"); f.argtypes = []; f.restype = ctypes.c_long
    return _LIB


@contextlib.contextmanager
def correctly_rounded_control_pow():
    """Within the block the FP64 oracle's rk45 step control rounds
    errmax**dt_decrease and errmax**dt_increase correctly (quad precision)
    instead of calling the C library's pow as the reference does: the
    rounding of a device step control (batched instances)."""
    L = lib()
    L.orc64_set_control_pow(1)
    try:
        yield
    finally:
        L.orc64_set_control_pow(0)


def control_pow(x, y):
    """The FP64 oracle's step-control pow in its current mode."""
    return lib().orc64_control_pow(float(x), float(y))


def control_pow_ambiguous():
    """Correctly rounded control pows whose quad result lay within 2^-100
    of a double rounding boundary (then the rounding is not certain)."""
    return lib().orc64_control_pow_ambiguous()


def params(cfg, shape, nthreads=1):
    """cfg: any object with the run attributes (scheme, eqn, bcs, order,
    real_time, xl.., dx.., g, mu, omx.., urhs, diss_amp, eps, safety,
    dt_increase, dt_decrease).  shape: (nz, ny, nx)."""
    nz, ny, nx = shape
    p = OrcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    sch = cfg.scheme
    p.scheme = SCHEMES[sch] if isinstance(sch, str) else int(sch)
    p.eqn, p.bcs, p.order = int(cfg.eqn), int(cfg.bcs), int(cfg.order)
    p.real_time = 1 if cfg.real_time else 0
    p.xl, p.yl, p.zl = cfg.xl, cfg.yl, cfg.zl
    p.dx, p.dy, p.dz = cfg.dx, cfg.dy, cfg.dz
    p.g, p.mu = cfg.g, cfg.mu
    p.omx, p.omy, p.omz = cfg.omx, cfg.omy, cfg.omz
    p.urhs, p.diss_amp = cfg.urhs, cfg.diss_amp
    p.eps, p.safety = cfg.eps, cfg.safety
    p.dt_increase, p.dt_decrease = cfg.dt_increase, cfg.dt_decrease
    p.nthreads = int(nthreads)
    return p


def _pfx(a):
    if a.dtype == np.complex128:
        return "orc64"
    if a.dtype == np.complex64:
        return "orc32"
    raise TypeError("oracle fields are complex128 or complex64, got %s" % a.dtype)


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def laplacian(cfg, psi, nthreads=1):
    out = np.empty_like(psi)
    getattr(lib(), _pfx(psi) + "_laplacian")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi), _ptr(out))
    return out


def deriv_x(cfg, psi, nthreads=1):
    out = np.empty_like(psi)
    getattr(lib(), _pfx(psi) + "_deriv_x")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi), _ptr(out))
    return out


def rhs(cfg, psi, nthreads=1):
    out = np.empty_like(psi)
    getattr(lib(), _pfx(psi) + "_rhs")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi), _ptr(out))
    return out


def grid(cfg, shape, dtype=np.float64):
    nz, ny, nx = shape
    x, y, z = (np.empty(n, dtype=dtype) for n in (nx, ny, nz))
    pfx = "orc64" if dtype == np.float64 else "orc32"
    getattr(lib(), pfx + "_grid")(ctypes.byref(params(cfg, shape)), _ptr(x), _ptr(y), _ptr(z))
    return x, y, z


def make_state(dt, t=0.0, im_t=0.0, p=0):
    s = OrcState()
    dt = complex(dt)
    s.dt_re, s.dt_im, s.t, s.im_t, s.p = dt.real, dt.imag, t, im_t, p
    return s


def solver(cfg, state, psi, nthreads=1):
    """One call of the reference's solver(); returns the new field and
    updates state (dt, t, im_t, errmax, n_rejects) in place."""
    out = np.empty_like(psi)
    rc = getattr(lib(), _pfx(psi) + "_solver")(
        ctypes.byref(params(cfg, psi.shape, nthreads)), ctypes.byref(state), _ptr(psi), _ptr(out))
    if rc != 0:
        raise RuntimeError("oracle solver failed (rc=%d)" % rc)
    return out


def run(cfg, state, psi, nsteps, nthreads=1, with_norms=False):
    """nsteps solver() calls in place on a copy of psi; returns (psi, norms)."""
    work = np.ascontiguousarray(psi).copy()
    norms = np.empty(nsteps, dtype=np.float64) if with_norms else None
    rc = getattr(lib(), _pfx(work) + "_run")(
        ctypes.byref(params(cfg, work.shape, nthreads)), ctypes.byref(state), _ptr(work), int(nsteps),
        _ptr(norms) if with_norms else None)
    if rc != 0:
        raise RuntimeError("oracle run failed (rc=%d)" % rc)
    return work, norms


def norm(cfg, psi, nthreads=1):
    return getattr(lib(), _pfx(psi) + "_norm")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi))


def energy(cfg, psi, nthreads=1):
    return getattr(lib(), _pfx(psi) + "_energy")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi))


def mass(cfg, psi, nthreads=1):
    return getattr(lib(), _pfx(psi) + "_mass")(ctypes.byref(params(cfg, psi.shape, nthreads)), _ptr(psi))


def renormalise(cfg, psi, nn, nrm):
    out = np.ascontiguousarray(psi).copy()
    getattr(lib(), _pfx(out) + "_renormalise")(ctypes.byref(params(cfg, out.shape)), _ptr(out), float(nn), float(nrm))
    return out


class OrcLine(ctypes.Structure):
    """struct orc_line (type line_param, parameters.in:33-43)."""
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
                ("amp1", ctypes.c_double), ("amp2", ctypes.c_double), ("ll", ctypes.c_double),
                ("sgn", ctypes.c_double), ("dir", ctypes.c_int), ("imprint_phase", ctypes.c_int)]


def imprint_vortex_line(cfg, psi, line, scal=1.0):
    """imprint_vortex_line (variables.f90:1166-1180) of a copy of psi; line
    is a dict with x0, y0, z0, amp1, amp2, ll, sgn, dir ('x'|'y'|'z'),
    imprint_phase."""
    out = np.array(psi, copy=True, order="C")
    vl = OrcLine(float(line.get("x0", 0.0)), float(line.get("y0", 0.0)), float(line.get("z0", 0.0)),
                 float(line.get("amp1", 0.0)), float(line.get("amp2", 0.0)), float(line.get("ll", 0.0)),
                 float(line.get("sgn", 1.0)), "xyz".index(line.get("dir", "z")), int(bool(line.get("imprint_phase", True))))
    rc = getattr(lib(), _pfx(out) + "_imprint_vortex_line")(ctypes.byref(params(cfg, out.shape)), _ptr(out),
                                                            ctypes.byref(vl), float(scal))
    if rc != 0:
        raise ValueError("vortex line disturbance wavelength is zero, but amplitude is non-zero")
    return out


ZERO_DTYPE = np.dtype([("section", np.int32), ("reserved", np.int32), ("cell", np.int64), ("x", np.float64),
                       ("y", np.float64), ("z", np.float64)])


def zeros(cfg, psi, which=0):
    """Vortex-core points of the field (io.f90:886-1393): which = 0 for the
    zeros file (get_zeros + get_extra_zeros, sections 0..5), 1 for re_zeros,
    2 for im_zeros (sections 0..2).  A structured array in file order."""
    psi = np.ascontiguousarray(psi)
    f = getattr(lib(), _pfx(psi) + "_zeros")
    prm = params(cfg, psi.shape)
    n = f(ctypes.byref(prm), _ptr(psi), int(which), None, 0)
    out = np.zeros(max(n, 1), dtype=ZERO_DTYPE)
    n2 = f(ctypes.byref(prm), _ptr(psi), int(which), out.ctypes.data_as(ctypes.c_void_p), n)
    assert n2 == n
    return out[:n]


def linelength(cfg, psi):
    """linelength (variables.f90:896-1162) of the field, in the field's
    precision."""
    psi = np.ascontiguousarray(psi)
    return getattr(lib(), _pfx(psi) + "_linelength")(ctypes.byref(params(cfg, psi.shape)), _ptr(psi))
