"""The oracle's linelength (variables.f90:896-1162) pinned on the CPU.

Closed forms: straight zero lines of psi along each axis, crossing every
interior cell through two opposite faces, measure (n-2)*dx exactly (each
cell's two points differ by exactly one unit); a zero line lying in a
shared face is counted half by each of the two cells (the face-of-zeros
branch); tilted lines measure their length to within the discretisation.
A literal transcription of the reference's loop (faces by `select case`,
arrays carried from cell to cell) in pure Python checks the table-driven C
restatement on small random fields, where every branch -- including cells
that find one point and reuse the previous cell's second point -- occurs.
"""
import math

import numpy as np
import pytest

from gpe_amd.config import RunConfig


def _cfg(n=(14, 12, 10), dx=0.5):
    nx, ny, nz = n
    return RunConfig(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2)


def _mesh(cfg):
    x, y, z = cfg.coords()
    return x[None, None, :], y[None, :, None], z[:, None, None]


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_straight_line_exact(oracle, axis):
    c = _cfg()
    X, Y, Z = _mesh(c)
    x0, y0, z0 = 0.3 * c.dx, -0.17 * c.dy, 0.41 * c.dz
    if axis == "z":
        psi = (X - x0) + 1j * (Y - y0) + 0 * Z
        want = (c.nz - 2) * c.dx
    elif axis == "x":
        psi = (Y - y0) + 1j * (Z - z0) + 0 * X
        want = (c.nx - 2) * c.dx
    else:
        psi = (Z - z0) + 1j * (X - x0) + 0 * Y
        want = (c.ny - 2) * c.dx
    psi = np.ascontiguousarray(psi.astype(np.complex128))
    assert oracle.linelength(c, psi) == pytest.approx(want, rel=1e-15)
    assert oracle.linelength(c, psi.astype(np.complex64)) == pytest.approx(want, rel=1e-6)


def test_line_in_a_shared_face_counts_once(oracle):
    # Re psi = 0 on the grid plane z = z_K, Im psi changes sign across y0:
    # every cell on either side of that face adds 1/2
    c = _cfg()
    X, Y, Z = _mesh(c)
    _, _, z = c.coords()
    K = 5
    psi = (Z - z[K]) + 1j * (Y - 0.23 * c.dy) + 0 * X
    assert oracle.linelength(c, np.ascontiguousarray(psi)) == pytest.approx((c.nx - 2) * c.dx, rel=1e-15)


def test_no_zeros_no_length(oracle):
    c = _cfg()
    psi = np.full(c.shape, 1.0 + 0.5j)
    assert oracle.linelength(c, psi) == 0.0


def test_tilted_line(oracle):
    # a line with direction (0.3, 0, 1): length sqrt(1.09) per unit z, less
    # the partial cells at its ends; the piecewise-linear face tracing is
    # exact for a field linear in x, y, z
    c = _cfg(n=(24, 20, 30), dx=0.5)
    X, Y, Z = _mesh(c)
    psi = np.ascontiguousarray((X - 0.3 * Z - 0.11) + 1j * (Y - 0.07) + 0 * Z)
    L = oracle.linelength(c, psi)
    full = math.sqrt(1.09) * (c.nz - 2) * c.dx
    assert abs(L - full) < 2 * math.sqrt(1.09) * c.dx


def _literal(psi, dx):
    """The reference loop, transcribed statement by statement."""
    nz, ny, nx = psi.shape
    x = [0.0] * 8
    y = [0.0] * 8
    z = [0.0] * 8
    L = 0.0
    P = lambda i, j, k: psi[k, j, i]  # noqa: E731
    for k in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                m = 1
                skip = False
                for n in range(1, 7):
                    if n == 1:
                        c = [P(i, j, k), P(i + 1, j, k), P(i + 1, j + 1, k), P(i, j + 1, k)]
                    elif n == 2:
                        c = [P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1)]
                    elif n == 3:
                        c = [P(i + 1, j, k), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k)]
                    elif n == 4:
                        c = [P(i, j, k), P(i, j, k + 1), P(i, j + 1, k + 1), P(i, j + 1, k)]
                    elif n == 5:
                        c = [P(i, j + 1, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1)]
                    else:
                        c = [P(i, j, k), P(i + 1, j, k), P(i + 1, j, k + 1), P(i, j, k + 1)]
                    u = [None] + [float(q.real) for q in c]
                    v = [None] + [float(q.imag) for q in c]
                    u.append(u[1])
                    v.append(v[1])
                    if ((u[1] == 0 and u[2] == 0 and u[3] == 0 and u[4] == 0 and
                         (v[1] * v[2] < 0 or v[2] * v[3] < 0 or v[3] * v[4] < 0 or v[4] * v[1] < 0)) or
                            (v[1] == 0 and v[2] == 0 and v[3] == 0 and v[4] == 0 and
                             (u[1] * u[2] < 0 or u[2] * u[3] < 0 or u[3] * u[4] < 0 or u[4] * u[1] < 0))):
                        L += 0.5
                        skip = True
                        break
                    edge = False
                    for q in range(1, 5):
                        if u[q] == 0 and u[q + 1] == 0 and v[q] == 0 and v[q + 1] == 0:
                            L += 0.25
                            edge = True
                            break
                    if edge:
                        skip = True
                        break

                    def cross(w):
                        xs, ys = [0.0] * 8, [0.0] * 8
                        lw = 1
                        for q in range(1, 5):
                            if w[q] == 0:
                                xs[lw], ys[lw] = {1: (0.0, 0.0), 2: (1.0, 0.0), 3: (1.0, 1.0), 4: (0.0, 1.0)}[q]
                            elif w[q] * w[q + 1] < 0:
                                if q == 1:
                                    xs[lw], ys[lw] = abs(w[q] / (w[q] - w[q + 1])), 0.0
                                elif q == 2:
                                    xs[lw], ys[lw] = 1.0, abs(w[q] / (w[q] - w[q + 1]))
                                elif q == 3:
                                    xs[lw], ys[lw] = abs(w[q + 1] / (w[q] - w[q + 1])), 1.0
                                else:
                                    xs[lw], ys[lw] = 0.0, abs(w[q + 1] / (w[q] - w[q + 1]))
                                if lw == 1 or xs[lw] != xs[lw - 1] or ys[lw] != ys[lw - 1]:
                                    lw += 1
                        return lw, xs, ys

                    lu, xu, yu = cross(u)
                    lv, xv, yv = cross(v)
                    if lu > 2 and lv > 2:
                        den = xv[2] * (yu[1] - yu[2]) + xv[1] * (yu[2] - yu[1]) + (xu[1] - xu[2]) * (yv[1] - yv[2])
                        if den == 0:
                            den = den + 0.0000001
                        xx = (xu[1] * (xv[2] * (yv[1] - yu[2]) + xv[1] * (yu[2] - yv[2])) +
                              xu[2] * (xv[2] * (yu[1] - yv[1]) + xv[1] * (yv[2] - yu[1]))) / den
                        yy = (xv[2] * (yu[1] - yu[2]) * yv[1] + xu[1] * yu[2] * yv[1] - xv[1] * yu[1] * yv[2] -
                              xu[1] * yu[2] * yv[2] + xv[1] * yu[2] * yv[2] + xu[2] * yu[1] * (yv[2] - yv[1])) / den
                        if 0 <= xx <= 1 and 0 <= yy <= 1:
                            x[m], y[m], z[m] = {1: (xx, yy, 0.0), 2: (xx, yy, 1.0), 3: (1.0, yy, xx),
                                                4: (0.0, yy, xx), 5: (xx, 1.0, yy), 6: (xx, 0.0, yy)}[n]
                            m += 1
                if skip:
                    continue
                if m > 1:
                    a = 1.0
                    b01 = lambda t: t == 0 or t == 1  # noqa: E731
                    if (x[1] == x[2] and b01(x[1])) or (y[1] == y[2] and b01(y[1])) or (y[1] == y[2] and b01(y[1])):
                        a *= 0.5
                        if ((x[1] == x[2] and b01(x[1]) and ((y[1] == y[2] and b01(y[1])) or
                                                            (y[1] == y[2] and b01(y[1])))) or
                                (y[1] == y[2] and b01(y[1]) and (z[1] == z[2] and b01(z[1])))):
                            a *= 0.5
                    L += a * math.sqrt((x[1] - x[2]) ** 2 + (y[1] - y[2]) ** 2 + (z[1] - z[2]) ** 2)
    return L * dx


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_literal_transcription(oracle, seed):
    c = _cfg(n=(11, 9, 8))
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    # a few exact zeros and zero edges so the special branches run too
    psi[3, 4, 5] = 0
    psi[3, 4, 6] = 0
    psi[5, 2, 2] = 1j * psi[5, 2, 2].imag
    psi = np.ascontiguousarray(psi)
    got = oracle.linelength(c, psi)
    want = _literal(psi, c.dx)
    assert got == pytest.approx(want, rel=1e-13, abs=1e-13)
