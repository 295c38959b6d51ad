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


def itos(n):
    """The reference's 7-digit file counter (io.f90:29-47): the last seven
    decimal digits of n, zero padded."""
    return "%07d" % (int(n) % 10 ** 7)


def write_dens(proc_dir, psi, cfg, p, t=0.0, im_t=0.0):
    """3-D dump of idl_surface (io.f90:480-520) for one rank, as Fortran
    stream I/O: real(pr) t+im_t; int32 nx, ny, nz; int32 nyprocs, nzprocs;
    int32 js, je, ks, ke; complex(pr) psi(0:nx1, js:je, ks:ke); real(pr)
    x(0:nx1), y(0:ny1), z(0:nz1).  Appends p to p_saved.dat ('(i10)') in the
    parent directory, as rank 0 does.  Returns the file path."""
    import os
    psi = np.ascontiguousarray(psi)
    rt = np.float64 if psi.dtype == np.complex128 else np.float32
    nz, ny, nx = psi.shape
    x, y, z = (np.asarray(a, dtype=rt) for a in cfg.coords())
    os.makedirs(proc_dir, exist_ok=True)
    path = os.path.join(proc_dir, "dens" + itos(p) + ".dat")
    with open(path, "wb") as fh:
        fh.write(np.array([t + im_t], dtype=rt).tobytes())
        fh.write(np.array([nx, ny, nz, 1, 1, 0, ny - 1, 0, nz - 1], dtype=np.int32).tobytes())
        fh.write(psi.tobytes())
        fh.write(x.tobytes())
        fh.write(y.tobytes())
        fh.write(z.tobytes())
    with open(os.path.join(os.path.dirname(os.path.abspath(proc_dir.rstrip("/"))), "p_saved.dat"), "a") as fh:
        fh.write("%10d\n" % p)
    return path


def read_dens(path, single=False):
    """Returns (psi, t_plus_im_t, (nyprocs, nzprocs), (js, je, ks, ke), x, y, z)."""
    rt, ct = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    with open(path, "rb") as fh:
        tt = float(np.frombuffer(fh.read(np.dtype(rt).itemsize), dtype=rt)[0])
        nx, ny, nz, nyp, nzp, js, je, ks, ke = (int(v) for v in np.frombuffer(fh.read(36), dtype=np.int32))
        n = nx * (je - js + 1) * (ke - ks + 1)
        psi = np.frombuffer(fh.read(n * np.dtype(ct).itemsize), dtype=ct).reshape(ke - ks + 1, je - js + 1, nx)
        x = np.frombuffer(fh.read(nx * np.dtype(rt).itemsize), dtype=rt)
        y = np.frombuffer(fh.read(ny * np.dtype(rt).itemsize), dtype=rt)
        z = np.frombuffer(fh.read(nz * np.dtype(rt).itemsize), dtype=rt)
    return psi.copy(), tt, (nyp, nzp), (js, je, ks, ke), x.copy(), y.copy(), z.copy()
