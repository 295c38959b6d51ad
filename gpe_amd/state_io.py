"""The reference's restart/state file format.

end_state (io.f90:843-853) writes, as Fortran stream I/O on one rank:
  int32 nx, ny, nz, p;  real(pr) t, im_t;  complex(pr) dt;
  complex(pr) psi(0:nx1, 0:ny1, 0:nz1)   (x fastest)
and state_restart (ic.f90:117-191) reads the same layout from
end_state####.dat.  pr is double or single precision (parameters.in:22-23).
Fields here are numpy arrays of shape (nz, ny, nx).
"""
import numpy as np


def write_state(path, psi, p=0, t=0.0, im_t=0.0, dt=0.0):
    psi = np.ascontiguousarray(psi)
    if psi.dtype == np.complex128:
        rt, ct = np.float64, np.complex128
    elif psi.dtype == np.complex64:
        rt, ct = np.float32, np.complex64
    else:
        raise TypeError("complex128 or complex64 field expected")
    nz, ny, nx = psi.shape
    with open(path, "wb") as fh:
        fh.write(np.array([nx, ny, nz, p], dtype=np.int32).tobytes())
        fh.write(np.array([t, im_t], dtype=rt).tobytes())
        fh.write(np.array([dt], dtype=ct).tobytes())
        fh.write(psi.tobytes())


def read_state(path, single=False):
    """Returns (psi, p, t, im_t, dt)."""
    rt, ct = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    with open(path, "rb") as fh:
        nx, ny, nz, p = (int(v) for v in np.frombuffer(fh.read(16), dtype=np.int32))
        t, im_t = (float(v) for v in np.frombuffer(fh.read(2 * np.dtype(rt).itemsize), dtype=rt))
        dt = complex(np.frombuffer(fh.read(np.dtype(ct).itemsize), dtype=ct)[0])
        psi = np.frombuffer(fh.read(nx * ny * nz * np.dtype(ct).itemsize), dtype=ct).reshape(nz, ny, nx).copy()
    return psi, p, t, im_t, dt


def write_run_in(path, cfg, end_time):
    """run_params / eqn_params namelists (parameters.f90:40-56) for cfg."""
    b = lambda v: ".true." if v else ".false."  # noqa: E731
    with open(path, "w") as fh:
        fh.write("&run_params\n")
        fh.write("  tau = %r\n  end_time = %r\n" % (float(cfg.tau), float(end_time)))
        fh.write("  xr = %r\n  yr = %r\n  zr = %r\n" % (float(cfg.xr), float(cfg.yr), float(cfg.zr)))
        fh.write("  scheme = '%s'\n  eqn_to_solve = %d\n  bcs = %d\n  order = %d\n" %
                 (cfg.scheme, cfg.eqn, cfg.bcs, cfg.order))
        fh.write("  real_time = %s\n/\n" % b(cfg.real_time))
        fh.write("&eqn_params\n")
        fh.write("  Urhs = %r\n  diss_amp = %r\n  g = %r\n  mu = %r\n" %
                 (float(cfg.urhs), float(cfg.diss_amp), float(cfg.g), float(cfg.mu)))
        fh.write("  omx = %r\n  omy = %r\n  omz = %r\n/\n" % (float(cfg.omx), float(cfg.omy), float(cfg.omz)))
