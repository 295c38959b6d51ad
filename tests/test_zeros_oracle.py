"""The oracle's vortex-core points (get_zeros, get_extra_zeros and
get_re_im_zeros, io.f90:886-1393) pinned on the CPU.

Closed forms.  For a field linear in the coordinates the linear
interpolant along an edge is exact, so an edge point sits where the field's
zero plane meets the edge, and the line through a face's two Re crossings
(or Im crossings) is that component's zero line itself, so the face point
is where the vortex line pierces the face.  Straight vortex lines, tilted
within their planes, are placed
  - inside a cell: found on the faces only -- xy faces for a line along z,
    xz for a line along y, yz for a line along x;
  - just off a grid edge, where Re and Im both change sign on one edge:
    found on that edge only (the face test excludes it);
and each case checks the count in every section and the positions.
Multiplying a field by i swaps the two face patterns; the face points must
not change, bit for bit.  The re/im files are checked against the zero
plane of one component, including the z section that the reference's
`.and.` leaves empty.  The '(3e17.9)' writer is checked against flang's
own output (tests/golden/fortran_e_fixture.*)."""
import os

import numpy as np
import pytest

from gpe_amd.config import RunConfig
from gpe_amd.state_io import ZERO_HEADERS, fortran_e, read_zeros, write_zeros

HERE = os.path.dirname(os.path.abspath(__file__))


def _cfg(n=(15, 13, 11), dx=0.5):
    nx, ny, nz = n
    return RunConfig(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2)


def _mesh(cfg):
    x, y, z = cfg.coords()
    return x[None, None, :], y[None, :, None], z[:, None, None]


def _field(re, im):
    return np.ascontiguousarray((re + 1j * im).astype(np.complex128))


def _counts(pts, nsec):
    return [int((pts["section"] == s).sum()) for s in range(nsec)]


def _xyz(pts, sec):
    m = pts["section"] == sec
    return np.stack([pts["x"][m], pts["y"][m], pts["z"][m]], axis=1)


def test_line_inside_a_cell_is_found_on_its_faces(oracle):
    c = _cfg()
    X, Y, Z = _mesh(c)
    x, y, z = c.coords()
    ni = c.nx - 2
    # along z, through cell (6, 5): xy faces
    x0, y0 = x[6] + 0.43 * c.dx, y[5] + 0.57 * c.dy
    psi = _field((X - x0) + 0.1 * (Y - y0) + 0 * Z, (Y - y0) - 0.15 * (X - x0) + 0 * Z)
    pts = oracle.zeros(c, psi, 0)
    assert _counts(pts, 6) == [0, 0, 0, c.nz - 2, 0, 0]
    p = _xyz(pts, 3)
    np.testing.assert_allclose(p[:, 0], x0, atol=1e-13)
    np.testing.assert_allclose(p[:, 1], y0, atol=1e-13)
    np.testing.assert_array_equal(p[:, 2], z[1:-1])
    # along y, through cell (4, ., 6): xz faces
    x0, z0 = x[4] + 0.38 * c.dx, z[6] + 0.61 * c.dz
    psi = _field((X - x0) + 0.1 * (Z - z0) + 0 * Y, (Z - z0) - 0.15 * (X - x0) + 0 * Y)
    pts = oracle.zeros(c, psi, 0)
    assert _counts(pts, 6) == [0, 0, 0, 0, c.ny - 2, 0]
    p = _xyz(pts, 4)
    np.testing.assert_allclose(p[:, 0], x0, atol=1e-13)
    np.testing.assert_array_equal(p[:, 1], y[1:-1])
    np.testing.assert_allclose(p[:, 2], z0, atol=1e-13)
    # along x, through cell (., 7, 3): yz faces (the face's first edge runs in z)
    y0, z0 = y[7] + 0.52 * c.dy, z[3] + 0.36 * c.dz
    psi = _field((Z - z0) + 0.1 * (Y - y0) + 0 * X, (Y - y0) - 0.15 * (Z - z0) + 0 * X)
    pts = oracle.zeros(c, psi, 0)
    assert _counts(pts, 6) == [0, 0, 0, 0, 0, ni]
    p = _xyz(pts, 5)
    np.testing.assert_array_equal(p[:, 0], x[1:-1])
    np.testing.assert_allclose(p[:, 1], y0, atol=1e-13)
    np.testing.assert_allclose(p[:, 2], z0, atol=1e-13)


def test_other_face_pattern_gives_the_same_points(oracle):
    # i*psi has Re' = -Im, Im' = Re: the else-if pattern of get_extra_zeros
    c = _cfg()
    X, Y, Z = _mesh(c)
    x, y, z = c.coords()
    x0, y0, z0 = x[6] + 0.43 * c.dx, y[5] + 0.57 * c.dy, z[4] + 0.29 * c.dz
    for re, im in [((X - x0) + 0.1 * (Y - y0) + 0 * Z, (Y - y0) - 0.15 * (X - x0) + 0 * Z),
                   ((X - x0) + 0.1 * (Z - z0) + 0 * Y, (Z - z0) - 0.15 * (X - x0) + 0 * Y),
                   ((Z - z0) + 0.1 * (Y - y0) + 0 * X, (Y - y0) - 0.15 * (Z - z0) + 0 * X)]:
        a = oracle.zeros(c, _field(re, im), 0)
        b = oracle.zeros(c, _field(-im, re), 0)
        fa, fb = a[a["section"] >= 3], b[b["section"] >= 3]
        assert len(fa) > 0
        for k in ("section", "cell", "x", "y", "z"):
            np.testing.assert_array_equal(fa[k], fb[k])


@pytest.mark.parametrize("edge", [0, 1, 2])
def test_line_next_to_an_edge_is_found_on_that_edge(oracle, edge):
    c = _cfg()
    X, Y, Z = _mesh(c)
    x, y, z = c.coords()
    if edge == 0:
        # along z, 0.01 dy above the x-edges at y(5): Re and Im both cross there
        x0, y0 = x[6] + 0.5 * c.dx, y[5] + 0.01 * c.dy
        psi = _field((X - x0) + 0.1 * (Y - y0) + 0 * Z, 0.2 * (X - x0) + (Y - y0) + 0 * Z)
        n, want = c.nz - 2, lambda p: (x0 - 0.1 * (y[5] - y0), y[5], None)
    elif edge == 1:
        # along z, 0.01 dx beside the y-edges at x(6)
        x0, y0 = x[6] + 0.01 * c.dx, y[5] + 0.5 * c.dy
        psi = _field((Y - y0) + 0.1 * (X - x0) + 0 * Z, 0.2 * (Y - y0) + (X - x0) + 0 * Z)
        n, want = c.nz - 2, lambda p: (x[6], y0 - 0.1 * (x[6] - x0), None)
    else:
        # along y, 0.01 dx beside the z-edges at x(6)
        x0, z0 = x[6] + 0.01 * c.dx, z[4] + 0.5 * c.dz
        psi = _field((Z - z0) + 0.1 * (X - x0) + 0 * Y, 0.2 * (Z - z0) + (X - x0) + 0 * Y)
        n, want = c.ny - 2, lambda p: (x[6], None, z0 - 0.1 * (x[6] - x0))
    pts = oracle.zeros(c, psi, 0)
    cnt = [0] * 6
    cnt[edge] = n
    assert _counts(pts, 6) == cnt
    p = _xyz(pts, edge)
    for ax, v in enumerate(want(p)):
        if v is not None:
            np.testing.assert_allclose(p[:, ax], v, atol=1e-13)
    free = {0: z, 1: z, 2: y}[edge][1:-1]
    np.testing.assert_array_equal(p[:, {0: 2, 1: 2, 2: 1}[edge]], free)


@pytest.mark.parametrize("which", [1, 2])
def test_re_im_zero_planes(oracle, which):
    c = _cfg()
    X, Y, Z = _mesh(c)
    x, y, z = c.coords()
    x0, y0, z0 = 0.137 * c.dx, -0.219 * c.dy, 0.311 * c.dz
    a, b, g = 1.0, 0.3, 0.2
    w = a * (X - x0) + b * (Y - y0) + g * (Z - z0)
    other = 1.0 + 0 * w  # no zeros
    psi = _field(w, other) if which == 1 else _field(other, w)
    pts = oracle.zeros(c, psi, which)
    # x-edges: x* = x0 - (b (y_j - y0) + g (z_k - z0)) / a inside (x_1, x_{nx-1})
    want = []
    for k in range(1, c.nz - 1):
        for j in range(1, c.ny - 1):
            xs = x0 - (b * (y[j] - y0) + g * (z[k] - z0)) / a
            if x[1] < xs < x[-1]:
                want.append((xs, y[j], z[k]))
    p = _xyz(pts, 0)
    assert len(p) == len(want) > 0
    np.testing.assert_allclose(p, np.array(want), atol=1e-13)
    want = []
    for k in range(1, c.nz - 1):
        for i in range(1, c.nx - 1):
            ys = y0 - (a * (x[i] - x0) + g * (z[k] - z0)) / b
            if y[1] < ys < y[-1]:
                want.append((k, ys, i))
    want.sort()
    p = _xyz(pts, 1)
    assert len(p) == len(want) > 0
    np.testing.assert_allclose(p[:, 1], [v[1] for v in want], atol=1e-12)
    np.testing.assert_array_equal(p[:, 0], [x[v[2]] for v in want])
    # w changes sign along z on many edges, yet the z section is empty
    crossings = (w[1:-1, 1:-1, 1:-1] * w[2:, 1:-1, 1:-1] < 0).sum()
    assert crossings > 0
    assert _counts(pts, 3)[2] == 0


def test_file_order_and_single_precision(oracle):
    from gpe_amd import ic
    c = _cfg((24, 22, 20), dx=0.6)
    # a global phase mixes Re and Im, so every file has points
    psi = ic.vortex_ring(c, 3.0) * np.exp(0.7j) * (1 + 0.05 * np.random.default_rng(4).standard_normal(c.shape))
    for which in (0, 1, 2):
        pts = oracle.zeros(c, psi, which)
        key = pts["section"].astype(np.int64) * (1 << 40) + pts["cell"]
        assert len(pts) > 0 and np.all(np.diff(key) > 0)
    # the ring's cores: every zeros point lies within a cell of the circle.
    # Where a component's two crossings of a face share a coordinate (a
    # zero line parallel to a face axis, as on the ring's symmetry planes)
    # the reference's slope m(1) is infinite and the point is NaN.
    pts = oracle.zeros(c, ic.vortex_ring(c, 3.0), 0)
    ok = np.isfinite(pts["x"]) & np.isfinite(pts["y"]) & np.isfinite(pts["z"])
    assert ok.sum() > 20 and (~ok).sum() < ok.sum()
    rho = np.hypot(pts["x"][ok], pts["y"][ok])
    assert np.all(np.hypot(rho - 3.0, pts["z"][ok]) < 1.5 * c.dx)
    # FP32 on the noisy field (no symmetry planes): the same points to
    # single precision
    p64 = oracle.zeros(c, psi, 0)
    p32 = oracle.zeros(c.with_(precision="fp32"), psi.astype(np.complex64), 0)
    both = np.isin(p64["cell"] + (p64["section"].astype(np.int64) << 40),
                   p32["cell"] + (p32["section"].astype(np.int64) << 40))
    assert both.mean() > 0.95
    q64 = p64[both]
    q32 = p32[np.isin(p32["cell"] + (p32["section"].astype(np.int64) << 40),
                      q64["cell"] + (q64["section"].astype(np.int64) << 40))]
    fin = np.isfinite(q64["x"]) & np.isfinite(q32["x"]) & np.isfinite(q64["y"]) & np.isfinite(q32["y"])
    assert fin.mean() > 0.9
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(q32[k][fin], q64[k][fin], atol=1e-4)


def test_zeros_file_format(tmp_path):
    # the E17.9 writer against flang's own output of the same values
    want = open(os.path.join(HERE, "golden", "fortran_e_fixture.txt")).read().splitlines()
    v = [1.0, -12.345678912345, 0.0, 3.2e-120, 9.9999999996, -0.5, 123456.789, 1.0e-5, -7.77777777777e33, 2.5,
         0.15, 1.0e100]
    s = [np.float32(0.1), np.float32(-3.14159274), np.float32(6.0e-8)]
    got = [" # i,j,k --> i+1,j,k"] + ["".join(fortran_e(q) for q in v[r:r + 3]) for r in range(0, 12, 3)]
    got.append("".join(fortran_e(float(q)) for q in s))
    assert got == want
    # a file round trip: headers in order, points per section
    pts = np.zeros(3, dtype=[("section", np.int32), ("reserved", np.int32), ("cell", np.int64),
                             ("x", np.float64), ("y", np.float64), ("z", np.float64)])
    pts["section"] = [0, 3, 5]
    pts["x"], pts["y"], pts["z"] = [1.5, -2.25, 0.125], [0.5, 3.0, -1.0], [2.0, 0.0, 7.5]
    path = write_zeros(str(tmp_path / "proc0000"), "zeros", 12, pts, nz=8)
    assert path.endswith("zeros0000012.dat")
    lines = open(path).read().splitlines()
    assert [ln for ln in lines if "#" in ln] == [" " + h for h in ZERO_HEADERS["zeros"]]
    secs = read_zeros(path)
    assert [len(q) for q in secs] == [1, 0, 0, 1, 0, 1]
    np.testing.assert_array_equal(secs[3], [[-2.25, 3.0, 0.0]])
