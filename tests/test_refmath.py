"""The libm restatements of gpe_amd/csrc/refmath.h, built for the host.

hypot_ref must equal the host libm's hypot (the reference's abs of a
complex) in every bit; pow_cr must be the correctly rounded x**y (checked
against 60-digit decimal arithmetic) and equal the host libm's pow on all
but a small fraction of the step control's arguments.  The header uses only
operations that round identically on the host and the device, so these
host results are the device's."""
import ctypes
import os
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = r"""
#include "refmath.h"
extern "C" void ref_hypot(const double *x, const double *y, long n, double *a, double *b) {
  for (long i = 0; i < n; ++i) { a[i] = gpe::refm::hypot_ref(x[i], y[i]); b[i] = std::hypot(x[i], y[i]); }
}
extern "C" void ref_pow(const double *x, const double *y, long n, double *a, double *b) {
  for (long i = 0; i < n; ++i) { a[i] = gpe::refm::pow_cr(x[i], y[i]); b[i] = std::pow(x[i], y[i]); }
}
"""


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("refmath")
    (d / "shim.cpp").write_text(SHIM)
    so = str(d / "shim.so")
    subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                           "-I" + os.path.join(ROOT, "gpe_amd", "csrc"), str(d / "shim.cpp"), "-o", so])
    L = ctypes.CDLL(so)
    vp = ctypes.c_void_p
    for f in (L.ref_hypot, L.ref_pow):
        f.argtypes = [vp, vp, ctypes.c_long, vp, vp]
        f.restype = None
    return L


def _call(f, x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    a, b = np.empty_like(x), np.empty_like(x)
    p = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    f(p(x), p(y), len(x), p(a), p(b))
    return a, b


def test_hypot_equals_host_libm(shim):
    rng = np.random.default_rng(20261202)
    n = 2_000_000
    cases = {
        "normal": (rng.standard_normal(n), rng.standard_normal(n)),
        "wide": (np.exp(rng.uniform(-700, 700, n)) * rng.choice([-1.0, 1.0], n), np.exp(rng.uniform(-700, 700, n))),
        "close": None,
        "tiny": (np.exp(rng.uniform(-744, -600, n)), -np.exp(rng.uniform(-744, -600, n))),
        "huge": (np.exp(rng.uniform(300, 709, n)), np.exp(rng.uniform(300, 709, n))),
    }
    a = np.exp(rng.uniform(-40, 40, n))
    cases["close"] = (a, a * np.exp(rng.uniform(-40, 0, n)))
    for name, (x, y) in cases.items():
        got, want = _call(shim.ref_hypot, x, y)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), name
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308, 1.0])
    x, y = np.meshgrid(special, special)
    got, want = _call(shim.ref_hypot, x.ravel(), y.ravel())
    assert np.array_equal(got, want, equal_nan=True)


def test_pow_correctly_rounded_and_near_host_libm(shim):
    getcontext().prec = 60
    rng = np.random.default_rng(7)
    n = 1_000_000
    # the step control's arguments: errmax/eps in (errcon, 1] with
    # dt_increase = -0.2, and above 1 with dt_decrease = -0.25
    xs = [np.exp(rng.uniform(np.log(1.89e-4), 0.0, n)), np.exp(rng.uniform(0.0, np.log(1e8), n)),
          np.exp(rng.uniform(-700, 700, n))]
    ys = [np.full(n, -0.2), np.full(n, -0.25), rng.uniform(-1.0, 1.0, n)]
    for x, y in zip(xs, ys):
        got, libm = _call(shim.ref_pow, x, y)
        assert np.all(np.isfinite(got))
        differ = np.nonzero(got != libm)[0]
        # one ulp at most, and rarely
        assert np.all(np.abs(got[differ] - libm[differ]) <= np.spacing(np.abs(libm[differ])))
        assert len(differ) < 2e-3 * n, len(differ)
        # correctly rounded: every disagreement with the host libm, and a
        # random sample of the rest, against 60-digit decimal arithmetic
        pick = np.concatenate([differ[:300], rng.choice(n, 300, replace=False)])
        for i in pick:
            exact = float(Decimal(float(x[i])) ** Decimal(float(y[i])))
            assert got[i] == exact, (x[i], y[i], got[i], exact)
    # exact cases and the fall-back arguments
    x = np.array([16.0, 0.0625, 1.0, 2.0, 0.0, np.inf, -1.0, 3.0])
    y = np.array([-0.25, -0.25, -0.2, 0.0, -0.25, -0.25, -0.2, np.nan])
    got, libm = _call(shim.ref_pow, x, y)
    assert np.array_equal(got, libm, equal_nan=True)
    assert got[0] == 0.5 and got[1] == 2.0


def test_oracle_control_pow_modes(shim, oracle):
    """The oracle's step-control pow: the C library's (the reference's) by
    default; correctly rounded from quad precision in the mode that pins a
    device step control.  Two independent correctly rounded
    implementations (this header's double-double pow_cr and the oracle's
    powq) agree bit for bit, and both differ from the library's pow on the
    same rare arguments."""
    rng = np.random.default_rng(11)
    n = 100_000
    x = np.concatenate([np.exp(rng.uniform(np.log(1.89e-4), 0.0, n)), np.exp(rng.uniform(0.0, np.log(1e8), n))])
    y = np.concatenate([np.full(n, -0.2), np.full(n, -0.25)])
    cr, libm = _call(shim.ref_pow, x, y)
    base = np.array([oracle.control_pow(a, b) for a, b in zip(x, y)])
    assert np.array_equal(base, libm)
    before = oracle.control_pow_ambiguous()
    with oracle.correctly_rounded_control_pow():
        q = np.array([oracle.control_pow(a, b) for a, b in zip(x, y)])
    assert oracle.control_pow_ambiguous() == before
    assert np.array_equal(q, cr)
    assert 0 < np.count_nonzero(q != libm) < 2e-3 * len(x)
    # the mode is restored on leaving the block: the library's pow again
    assert np.array_equal(np.array([oracle.control_pow(a, b) for a, b in zip(x[:1000], y[:1000])]), libm[:1000])
