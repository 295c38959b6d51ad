"""Python host interface of the GP time-stepping path.

`Simulation` owns one device context (one field on the GPU) and exposes the
reference's operations on it: solver() calls (gpe_sim_advance / the drop-in
solver_step), get_norm, energy, mass, renormalise and get_rhs.  Fields are
numpy arrays of shape (nz, ny, nx) -- x fastest, the Fortran order of the
reference's psi(0:nx1, js:je, ks:ke).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .config import RunConfig


def make_params(cfg: RunConfig) -> _lib.GpeParams:
    p = _lib.GpeParams()
    p.nx, p.ny, p.nz = cfg.nx, cfg.ny, cfg.nz
    p.scheme = _lib.SCHEME_CODES[cfg.scheme]
    p.eqn, p.bcs, p.order = cfg.eqn, cfg.bcs, cfg.order
    p.precision = _lib.PRECISION_CODES[cfg.precision]
    p.real_time = 1 if cfg.real_time else 0
    p.xl, p.yl, p.zl = cfg.xl, cfg.yl, cfg.zl
    p.dx, p.dy, p.dz = cfg.dx, cfg.dy, cfg.dz
    p.g, p.mu = cfg.g, cfg.mu
    p.omx, p.omy, p.omz = cfg.omx, cfg.omy, cfg.omz
    p.urhs, p.diss_amp = cfg.urhs, cfg.diss_amp
    p.eps, p.safety = cfg.eps, cfg.safety
    p.dt_increase, p.dt_decrease = cfg.dt_increase, cfg.dt_decrease
    return p


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Simulation:
    """One GP simulation resident on the GPU."""

    def __init__(self, cfg: RunConfig, psi=None, dt=None, t=0.0, im_t=0.0, p=0, tuned=True):
        """tuned: apply the stored launch configuration of this problem
        (gpe_amd/tuning.json or $GPE_TUNING_FILE) when there is one; the
        launch configuration changes speed only, never the results."""
        self.cfg = cfg
        self._L = _lib.lib()
        self._ctx = ctypes.c_void_p()
        _lib.check(self._L.gpe_ctx_create(ctypes.byref(make_params(cfg)), ctypes.byref(self._ctx)))
        self.set_state(cfg.dt if dt is None else dt, t=t, im_t=im_t, p=p)
        if tuned:
            from . import tuning
            tuning.apply(self, tuning.DEFAULT_PATH)
        if psi is not None:
            self.upload(psi)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self._ctx:
            self._L.gpe_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --------------------------------------------------------------
    @property
    def host_dtype(self):
        return np.complex64 if self.cfg.precision == "fp32" else np.complex128

    def set_state(self, dt, t=0.0, im_t=0.0, p=0):
        s = _lib.GpeStepState()
        dt = complex(dt)
        s.dt_re, s.dt_im, s.t, s.im_t, s.p = dt.real, dt.imag, t, im_t, p
        _lib.check(self._L.gpe_sim_set_state(self._ctx, ctypes.byref(s)))

    def state(self):
        s = _lib.GpeStepState()
        _lib.check(self._L.gpe_sim_get_state(self._ctx, ctypes.byref(s)))
        return {"dt": complex(s.dt_re, s.dt_im), "t": s.t, "im_t": s.im_t, "errmax": s.errmax,
                "n_rejects": s.n_rejects, "p": s.p}

    # -- data ---------------------------------------------------------------
    def upload(self, psi):
        psi = np.ascontiguousarray(psi)
        if psi.shape != self.cfg.shape:
            raise ValueError("field shape %s != %s" % (psi.shape, self.cfg.shape))
        if psi.dtype == np.complex128:
            ht = _lib.GPE_HOST_C128
        elif psi.dtype == np.complex64:
            ht = _lib.GPE_HOST_C64
        else:
            raise TypeError("complex128 or complex64 field expected")
        _lib.check(self._L.gpe_sim_upload(self._ctx, _ptr(psi), ht))

    def download(self, dtype=None):
        dtype = np.dtype(dtype or (np.complex64 if self.cfg.precision == "fp32" else np.complex128))
        out = np.empty(self.cfg.shape, dtype=dtype)
        ht = _lib.GPE_HOST_C128 if dtype == np.complex128 else _lib.GPE_HOST_C64
        _lib.check(self._L.gpe_sim_download(self._ctx, _ptr(out), ht))
        return out

    # -- stepping -----------------------------------------------------------
    def advance(self, nsteps, record_norm=False):
        """nsteps solver() calls; returns the per-step get_norm if asked."""
        _lib.check(self._L.gpe_sim_advance(self._ctx, int(nsteps), 1 if record_norm else 0))
        if record_norm:
            h = np.empty(int(nsteps), dtype=np.float64)
            _lib.check(self._L.gpe_sim_norm_history(
                self._ctx, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(nsteps)))
            return h
        return None

    def solver_step(self, var_in, halo=2):
        """The drop-in call: var_in has `halo` ghost layers in y and z, shape
        (nz+2h, ny+2h, nx); returns var_out (nz, ny, nx) and updates the
        state as solver() does."""
        var_in = np.ascontiguousarray(var_in)
        nz, ny, nx = self.cfg.shape
        if var_in.shape != (nz + 2 * halo, ny + 2 * halo, nx):
            raise ValueError("var_in shape %s" % (var_in.shape,))
        out = np.empty(self.cfg.shape, dtype=var_in.dtype)
        s = _lib.GpeStepState()
        _lib.check(self._L.gpe_sim_get_state(self._ctx, ctypes.byref(s)))
        _lib.check(self._L.gpe_solver_step(self._ctx, _ptr(var_in), int(halo), _ptr(out), ctypes.byref(s)))
        _lib.check(self._L.gpe_sim_set_state(self._ctx, ctypes.byref(s)))
        return out

    # -- reductions ---------------------------------------------------------
    def _scalar(self, fn):
        v = ctypes.c_double()
        _lib.check(fn(self._ctx, ctypes.byref(v)))
        return v.value

    def norm(self):
        return self._scalar(self._L.gpe_sim_norm)

    def energy(self):
        return self._scalar(self._L.gpe_sim_energy)

    def mass(self):
        return self._scalar(self._L.gpe_sim_mass)

    def hamiltonian(self):
        """The eqn-4 Hamiltonian the dynamics conserve (kinetic energy of
        the finite-difference Laplacian + trap + quartic - mu N)."""
        return self._scalar(self._L.gpe_sim_hamiltonian)

    def save_end_state(self, path):
        """The field and step state in the reference's end_state format
        (io.f90:843-853)."""
        from .state_io import write_state
        st = self.state()
        write_state(path, self.download(), p=st["p"], t=st["t"], im_t=st["im_t"], dt=st["dt"])

    def save_dens(self, proc_dir):
        """3-D dump in the format of idl_surface (io.f90:480-520)."""
        from .state_io import write_dens
        st = self.state()
        return write_dens(proc_dir, self.download(), self.cfg, st["p"], st["t"], st["im_t"])

    def linelength(self):
        """Total vortex-line length (linelength, variables.f90:896-1162)."""
        return self._scalar(self._L.gpe_sim_linelength)

    def imprint_vortex_line(self, line, scal=1.0):
        """imprint_vortex_line (variables.f90:1166-1180) on the resident
        field; line is a dict with x0, y0, z0, amp1, amp2, ll, sgn,
        dir ('x'|'y'|'z') and imprint_phase (type line_param)."""
        vl = _lib.GpeLineParam(float(line.get("x0", 0.0)), float(line.get("y0", 0.0)), float(line.get("z0", 0.0)),
                               float(line.get("amp1", 0.0)), float(line.get("amp2", 0.0)),
                               float(line.get("ll", 0.0)), float(line.get("sgn", 1.0)),
                               "xyz".index(line.get("dir", "z")), int(bool(line.get("imprint_phase", True))))
        _lib.check(self._L.gpe_sim_imprint_vortex_line(self._ctx, ctypes.byref(vl), float(scal)))

    def random_phase(self, seed, kmax=0.5):
        """Replace the resident field by the random-phase state built on the
        device (the state ic.random_phase builds on the host: unit modes
        with |k| <= kmax and SplitMix64 phases, inverse FFT, <|psi|^2> = 1)."""
        _lib.check(self._L.gpe_sim_random_phase(self._ctx, ctypes.c_uint64(int(seed)), float(kmax)))

    def zeros(self, which="zeros"):
        """Vortex-core points of the resident field as the reference's
        zeros files hold them (io.f90:886-1393): which = "zeros"
        (get_zeros + get_extra_zeros, six sections), "re_zeros" or
        "im_zeros" (get_re_im_zeros, three sections).  A structured array
        (section, cell, x, y, z) in file order."""
        w = _lib.ZERO_FILES[which]
        n = ctypes.c_int64()
        _lib.check(self._L.gpe_sim_zeros(self._ctx, w, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=_lib.ZERO_DTYPE)
        m = ctypes.c_int64()
        _lib.check(self._L.gpe_sim_zeros(self._ctx, w, out.ctypes.data_as(ctypes.c_void_p), n.value,
                                         ctypes.byref(m)))
        if m.value != n.value:
            raise RuntimeError("zero count changed between calls: %d, %d" % (n.value, m.value))
        return out[:n.value]

    def save_zeros(self, proc_dir):
        """zeros<p>.dat, re_zeros<p>.dat and im_zeros<p>.dat of the resident
        field in the reference's text format (io.f90:886-1393); returns the
        three paths."""
        from .state_io import write_zeros
        p = self.state()["p"]
        return [write_zeros(proc_dir, w, p, self.zeros(w), self.cfg.nz) for w in ("zeros", "re_zeros", "im_zeros")]

    # -- FFT-side diagnostics (unitary forward DFT of psi on the device) ---
    def condensed_particles(self):
        """condensed_particles (io.f90:304-395): the mass.dat record as a
        dict (time, im_t, density, n0_per_site, mass, n0, temp, temp2,
        energy_per_site)."""
        r = _lib.GpeCondensed()
        _lib.check(self._L.gpe_sim_condensed_particles(self._ctx, ctypes.byref(r)))
        return {n: getattr(r, n) for n, _ in r._fields_}

    def spectrum(self):
        """spectrum (io.f90:623-716): (nharm[7], eta[7], idf[nx/2]) - modes
        and mean occupation per octave shell, and the integral distribution
        F(kk), kk = 1..nx/2."""
        nb = self.cfg.nx // 2
        nh = np.zeros(7, np.int32)
        eta = np.zeros(7)
        idf = np.zeros(max(nb, 1))
        _lib.check(self._L.gpe_sim_spectrum(self._ctx, nh.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            eta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            idf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nb))
        return nh, eta, idf[:nb]

    def filtered(self, filter_kc, m=1.0, with_length=False, single=False):
        """filtered_surface (io.f90:524-619): the field low-pass filtered by
        max(1 - m k^2/filter_kc^2, 0) in k-space, as an (nz, ny, nx) array
        (complex64 if single), and with_length its vortex-line length."""
        out = np.empty(self.cfg.shape, np.complex64 if single else np.complex128)
        ln = ctypes.c_double()
        _lib.check(self._L.gpe_sim_filtered(self._ctx, float(filter_kc), float(m), _ptr(out),
                                            _lib.GPE_HOST_C64 if single else _lib.GPE_HOST_C128,
                                            ctypes.byref(ln) if with_length else None))
        return (out, ln.value) if with_length else out

    def renormalise(self, nn, norm):
        _lib.check(self._L.gpe_sim_renormalise(self._ctx, float(nn), float(norm)))

    def rhs(self):
        dtype = np.complex64 if self.cfg.precision == "fp32" else np.complex128
        out = np.empty(self.cfg.shape, dtype=dtype)
        ht = _lib.GPE_HOST_C128 if dtype == np.complex128 else _lib.GPE_HOST_C64
        _lib.check(self._L.gpe_sim_rhs(self._ctx, _ptr(out), ht))
        return out

    # -- kernel timing ------------------------------------------------------
    def prof_enable(self, on=True):
        _lib.check(self._L.gpe_prof_enable(self._ctx, 1 if on else 0))

    def prof_query(self):
        """{kind name: (launches, total_ms, algorithmic bytes per launch)}"""
        out = {}
        for k, nm in enumerate(_lib.KIND_NAMES):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            name = ctypes.create_string_buffer(64)
            _lib.check(self._L.gpe_prof_query(self._ctx, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by), name))
            if n.value:
                out[nm] = (n.value, ms.value, by.value)
        return out

    # -- launch configuration -----------------------------------------------
    def tune_set(self, kind, variant, zchunk=0):
        k = _lib.KIND_NAMES.index(kind) if isinstance(kind, str) else int(kind)
        _lib.check(self._L.gpe_tune_set(self._ctx, k, int(variant), int(zchunk)))

    def set_fusion(self, enable=True):
        """Allow fused RK4 / Cash-Karp stage pairs (bitwise the same fields);
        a pair runs fused once its variant is set >= 0 (tune_set or
        autotune)."""
        _lib.check(self._L.gpe_set_fusion(self._ctx, 1 if enable else 0))

    def tune_get(self):
        """{kind name: (variant, zchunk, waves, rows per wave)} for the stage
        kinds and the fused stage pairs"""
        out = {}
        for nm in _lib.STAGE_KINDS + _lib.PAIR_KINDS:
            k = _lib.KIND_NAMES.index(nm)
            v, z, w, r = (ctypes.c_int32() for _ in range(4))
            _lib.check(self._L.gpe_tune_get(self._ctx, k, ctypes.byref(v), ctypes.byref(z), ctypes.byref(w),
                                            ctypes.byref(r)))
            out[nm] = (v.value, z.value, w.value, r.value)
        return out

    def tune_key(self):
        """This problem's key in the launch-configuration table."""
        buf = ctypes.create_string_buffer(128)
        _lib.check(self._L.gpe_tune_key(self._ctx, buf, len(buf)))
        return buf.value.decode()

    def tune_load(self, path=None):
        """Apply the stored launch configuration of this problem from the
        table at path (default: $GPE_TUNING_FILE, else tuning.json beside the
        library); True if the table held one."""
        got = ctypes.c_int32()
        _lib.check(self._L.gpe_tune_load(self._ctx, os.fsencode(path) if path else None, ctypes.byref(got)))
        return bool(got.value)

    def autotune(self, reps=3):
        _lib.check(self._L.gpe_autotune(self._ctx, int(reps)))
        return self.tune_get()


def random_phase_prepare(cfg: RunConfig):
    """Create the cached inverse-FFT plan the device random-phase state of
    this grid uses (safe from a worker thread; no field is touched)."""
    _lib.check(_lib.lib().gpe_random_phase_prepare(int(cfg.nx), int(cfg.ny), int(cfg.nz)))


def random_phase_modes(cfg: RunConfig, seed, kmax=0.5):
    """Flat indices and unit phasors of the random-phase state's modes, as
    the device generator forms them (host only, no GPU work)."""
    L = _lib.lib()
    args = (int(cfg.nx), int(cfg.ny), int(cfg.nz), float(cfg.dx), float(cfg.dy), float(cfg.dz), float(kmax),
            ctypes.c_uint64(int(seed)))
    n = ctypes.c_int64()
    _lib.check(L.gpe_random_phase_modes(*args, None, None, 0, ctypes.byref(n)))
    idx = np.zeros(max(n.value, 1), dtype=np.int64)
    ph = np.zeros(max(n.value, 1), dtype=np.complex128)
    _lib.check(L.gpe_random_phase_modes(*args, idx.ctypes.data_as(ctypes.c_void_p), ph.ctypes.data_as(ctypes.c_void_p),
                                        n.value, ctypes.byref(n)))
    return idx[:n.value], ph[:n.value]
