"""The reference's restart/state file format.

end_state (io.f90:843-853) writes, as Fortran stream I/O on one rank:
  int32 nx, ny, nz, p;  real(pr) t, im_t;  complex(pr) dt;
  complex(pr) psi(0:nx1, 0:ny1, 0:nz1)   (x fastest)
and state_restart (ic.f90:117-191) reads the same layout from
end_state####.dat.  pr is double or single precision (parameters.in:22-23).
Fields here are numpy arrays of shape (nz, ny, nx).
"""
import os

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


def _read_exact(fh, n, what, path):
    b = fh.read(n)
    if len(b) != n:
        raise ValueError("%s: truncated %s (%d of %d bytes)" % (path, what, len(b), n))
    return b


def read_state(path, single=False):
    """Returns (psi, p, t, im_t, dt).  A file whose header does not describe
    its own size raises ValueError before the field is read."""
    rt, ct = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    rs, cs = np.dtype(rt).itemsize, np.dtype(ct).itemsize
    size = os.path.getsize(path)
    with open(path, "rb") as fh:
        nx, ny, nz, p = (int(v) for v in np.frombuffer(_read_exact(fh, 16, "header", path), dtype=np.int32))
        if min(nx, ny, nz) < 1:
            raise ValueError("%s: grid %dx%dx%d in the header" % (path, nx, ny, nz))
        want = 16 + 2 * rs + cs + nx * ny * nz * cs
        if size != want:
            raise ValueError("%s: %d bytes, the header's %dx%dx%d grid needs %d" % (path, size, nx, ny, nz, want))
        t, im_t = (float(v) for v in np.frombuffer(_read_exact(fh, 2 * rs, "times", path), dtype=rt))
        dt = complex(np.frombuffer(_read_exact(fh, cs, "dt", path), dtype=ct)[0])
        psi = np.frombuffer(_read_exact(fh, nx * ny * nz * cs, "field", path), dtype=ct).reshape(nz, ny, nx).copy()
    return psi, p, t, im_t, dt


def _nml_value(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, str):
        return "'%s'" % v
    return repr(int(v) if isinstance(v, (int, np.integer)) else float(v))


def write_run_in(path, cfg, end_time, io=None, run_extra=None, eqn_extra=None):
    """run_params / eqn_params namelists (parameters.f90:40-56) for cfg,
    and io_params (parameters.f90:74-76) from the dict `io` (save_rate,
    save_rate2, save_rate3, p_save) when given.  run_extra / eqn_extra add
    other entries of those namelists (e.g. renorm, nn)."""
    b = lambda v: ".true." if v else ".false."  # noqa: E731
    with open(path, "w") as fh:
        fh.write("&run_params\n")
        for k, v in (run_extra or {}).items():
            fh.write("  %s = %s\n" % (k, _nml_value(v)))
        fh.write("  tau = %r\n  end_time = %r\n" % (float(cfg.tau), float(end_time)))
        fh.write("  xr = %r\n  yr = %r\n  zr = %r\n" % (float(cfg.xr), float(cfg.yr), float(cfg.zr)))
        fh.write("  scheme = '%s'\n  eqn_to_solve = %d\n  bcs = %d\n  order = %d\n" %
                 (cfg.scheme, cfg.eqn, cfg.bcs, cfg.order))
        fh.write("  real_time = %s\n/\n" % b(cfg.real_time))
        fh.write("&eqn_params\n")
        for k, v in (eqn_extra or {}).items():
            fh.write("  %s = %s\n" % (k, _nml_value(v)))
        fh.write("  Urhs = %r\n  diss_amp = %r\n  g = %r\n  mu = %r\n" %
                 (float(cfg.urhs), float(cfg.diss_amp), float(cfg.g), float(cfg.mu)))
        fh.write("  omx = %r\n  omy = %r\n  omz = %r\n/\n" % (float(cfg.omx), float(cfg.omy), float(cfg.omz)))
        if io:
            fh.write("&io_params\n")
            for k, v in io.items():
                fh.write("  %s = %r\n" % (k, int(v) if k == "save_rate" else float(v)))
            fh.write("/\n")


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
    os.makedirs(proc_dir, exist_ok=True)
    path = os.path.join(proc_dir, "dens" + itos(p) + ".dat")
    _write_dump(path, t + im_t, psi, cfg)
    with open(os.path.join(os.path.dirname(os.path.abspath(proc_dir.rstrip("/"))), "p_saved.dat"), "a") as fh:
        fh.write("%10d\n" % p)
    return path


def _write_dump(path, tval, psi, cfg):
    psi = np.ascontiguousarray(psi)
    rt = np.float64 if psi.dtype == np.complex128 else np.float32
    nz, ny, nx = psi.shape
    x, y, z = (np.asarray(a, dtype=rt) for a in cfg.coords())
    with open(path, "wb") as fh:
        fh.write(np.array([tval], dtype=rt).tobytes())
        fh.write(np.array([nx, ny, nz, 1, 1, 0, ny - 1, 0, nz - 1], dtype=np.int32).tobytes())
        fh.write(psi.tobytes())
        fh.write(x.tobytes())
        fh.write(y.tobytes())
        fh.write(z.tobytes())


def write_filtered(proc_dir, psi, cfg, p, t=0.0):
    """filtered<p>.dat of filtered_surface(a, 0) (io.f90:586-596): the dens
    layout (read_dens reads it) with t in place of t+im_t.  Returns the
    path.  end_state_filtered.dat (flag 1) is the end_state layout:
    write_state."""
    os.makedirs(proc_dir, exist_ok=True)
    path = os.path.join(proc_dir, "filtered" + itos(p) + ".dat")
    _write_dump(path, t, psi, cfg)
    return path


MASS_FIELDS = ("time", "im_t", "density", "n0_per_site", "mass", "n0", "temp", "temp2", "energy_per_site")


def append_mass(path, rec):
    """One condensed_particles record appended to mass.dat, '(9e17.9)'
    (io.f90:376-381); rec is Simulation.condensed_particles()'s dict."""
    with open(path, "a") as fh:
        fh.write("".join(fortran_e(rec[k]) for k in MASS_FIELDS) + "\n")


def write_spectrum(root, p, t, nharm, eta, idf):
    """The files of spectrum (io.f90:691-708) under root: spectrum/
    spectrum<p>.dat ('(2i9,e17.9)' per shell), idf/idf<p>.dat ('(i9,e17.9)'
    per kk) and a '(4e17.9)' line t, eta(1:3) appended to eta_time.dat.
    Returns the three paths."""
    for d in ("spectrum", "idf"):
        os.makedirs(os.path.join(root, d), exist_ok=True)
    ps = os.path.join(root, "spectrum", "spectrum" + itos(p) + ".dat")
    with open(ps, "w") as fh:
        for m in range(len(eta)):
            fh.write("%9d%9d%s\n" % (m + 1, int(nharm[m]), fortran_e(eta[m])))
    pi = os.path.join(root, "idf", "idf" + itos(p) + ".dat")
    with open(pi, "w") as fh:
        for kk, v in enumerate(idf, start=1):
            fh.write("%9d%s\n" % (kk, fortran_e(v)))
    pe = os.path.join(root, "eta_time.dat")
    with open(pe, "a") as fh:
        fh.write(fortran_e(t) + "".join(fortran_e(v) for v in eta[:3]) + "\n")
    return ps, pi, pe


def read_dens(path, single=False):
    """Returns (psi, t_plus_im_t, (nyprocs, nzprocs), (js, je, ks, ke), x, y, z)."""
    rt, ct = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    rs, cs = np.dtype(rt).itemsize, np.dtype(ct).itemsize
    size = os.path.getsize(path)
    with open(path, "rb") as fh:
        tt = float(np.frombuffer(_read_exact(fh, rs, "time", path), dtype=rt)[0])
        nx, ny, nz, nyp, nzp, js, je, ks, ke = (int(v) for v in np.frombuffer(_read_exact(fh, 36, "header", path),
                                                                                dtype=np.int32))
        if min(nx, ny, nz) < 1 or not (0 <= js <= je < ny and 0 <= ks <= ke < nz):
            raise ValueError("%s: grid %dx%dx%d with rows %d..%d, planes %d..%d in the header" %
                             (path, nx, ny, nz, js, je, ks, ke))
        n = nx * (je - js + 1) * (ke - ks + 1)
        want = rs + 36 + n * cs + (nx + ny + nz) * rs
        if size != want:
            raise ValueError("%s: %d bytes, the header describes %d" % (path, size, want))
        psi = np.frombuffer(_read_exact(fh, n * cs, "field", path), dtype=ct).reshape(ke - ks + 1, je - js + 1, nx)
        x = np.frombuffer(_read_exact(fh, nx * rs, "x", path), dtype=rt)
        y = np.frombuffer(_read_exact(fh, ny * rs, "y", path), dtype=rt)
        z = np.frombuffer(_read_exact(fh, nz * rs, "z", path), dtype=rt)
    return psi.copy(), tt, (nyp, nzp), (js, je, ks, ke), x.copy(), y.copy(), z.copy()


# Section headers of the zeros files, as get_zeros / get_extra_zeros
# (io.f90:919, 945, 971, 1037, 1102, 1165) and get_re_im_zeros
# (io.f90:1263, 1282, 1302; the im file repeats them) write them with
# list-directed output, which puts one blank in front.
ZERO_HEADERS = {
    "zeros": ["# i,j,k --> i+1,j,k", "# i,j,k --> i,j+1,k", "# i,j,k --> i,j,k+1",
              "# i,j,k --> i+1,j,k --> i+1,j+1,k --> i,j+1,k", "# i,j,k --> i+1,j,k --> i+1,j,k+1 --> i,j,k+1",
              "# i,j,k --> i,j,k+1 --> i,j+1,k+1 --> i,j+1,k"],
    "re_zeros": ["# i,j,k --> i+1,j,k", "# i,j,k --> i,j+1,k", "# i,j,k --> i,j,k+1"],
    "im_zeros": ["# i,j,k --> i+1,j,k", "# i,j,k --> i,j+1,k", "# i,j,k --> i,j,k+1"],
}


def fortran_e(v, w=17, d=9):
    """One value in the Fortran edit descriptor Ew.d (the reference writes
    zero points with '(3e17.9)'): 0.<d digits>E+xx, right-justified."""
    v = float(v)
    if v != v:
        s = "NaN"
    elif v in (float("inf"), float("-inf")):
        s = "Infinity" if v > 0 else "-Infinity"
    elif v == 0.0:
        s = ("-" if np.signbit(v) else "") + "0." + "0" * d + "E+00"
    else:
        m = "%.*e" % (d - 1, abs(v))
        mant, ex = m.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
        s = ("-" if v < 0 else "") + "0." + digits + ("E%+03d" % e if abs(e) <= 99 else "%+04d" % e)
    return s.rjust(w)


def write_zeros(proc_dir, which, p, points, nz):
    """proc_dir/<which><itos(p)>.dat with the points of a zeros pass (a
    structured array with fields section, x, y, z in file order).  The z
    sections are written only for a 3-D grid (z_start /= z_end,
    io.f90:970, 1101, 1301)."""
    os.makedirs(proc_dir, exist_ok=True)
    path = os.path.join(proc_dir, "%s%s.dat" % (which, itos(p)))
    heads = ZERO_HEADERS[which]
    sec = np.asarray(points["section"])
    with open(path, "w") as fh:
        for s, h in enumerate(heads):
            if nz <= 1 and ((which == "zeros" and s in (2, 4, 5)) or (which != "zeros" and s == 2)):
                continue
            fh.write(" " + h + "\n")
            for q in np.flatnonzero(sec == s):
                fh.write(fortran_e(points["x"][q]) + fortran_e(points["y"][q]) + fortran_e(points["z"][q]) + "\n")
    return path


def read_zeros(path):
    """Sections of a zeros file: a list of (n, 3) float arrays."""
    out = []
    cur = None
    with open(path) as fh:
        for line in fh:
            if line.strip().startswith("#"):
                cur = []
                out.append(cur)
            elif line.strip():
                cur.append([float(line[q:q + 17].replace("D", "E")) for q in (0, 17, 34)])
    return [np.array(s, dtype=np.float64).reshape(-1, 3) for s in out]
