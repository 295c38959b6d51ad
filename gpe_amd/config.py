"""Run configuration: the reference's run.in namelists for the time-stepping
path, plus the grid that get_grid() derives from them.

Mirrors parameters.f90:24-56 (run_params, eqn_params), the compile-time
adaptive-step constants parameters.f90:90-93, and get_grid (ic.f90:32-69):
x_i = xl + i*dx with xl = -xr and dx = (xr - xl)/(nx-1), so both walls are
grid points.
"""
from dataclasses import dataclass, field, replace

import numpy as np

SCHEMES = ("euler", "rk2", "rk4", "rk45")
PRECISIONS = ("fp64", "fp32", "fp32c")


@dataclass
class RunConfig:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    scheme: str = "rk4"        # run_params%scheme
    eqn: int = 4               # eqn_to_solve
    bcs: int = 2               # 1 periodic, 2 reflective
    order: int = 2             # 2 or 4
    real_time: bool = True
    tau: float = 0.01          # |dt|
    xr: float = 12.6
    yr: float = None
    zr: float = None
    g: float = 1.0
    mu: float = 1.0
    omx: float = 0.0
    omy: float = 0.0
    omz: float = 0.0
    urhs: float = 0.0
    diss_amp: float = 0.0
    nn: float = 1.0            # renormalisation target (imaginary time)
    eps: float = 1e-6          # parameters.f90:90
    safety: float = 0.9        # parameters.f90:91
    dt_decrease: float = -0.25  # parameters.f90:92
    dt_increase: float = -0.20  # parameters.f90:93
    precision: str = "fp64"    # fp64 | fp32 | fp32c (FP32 storage + compensated update)
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.yr is None:
            self.yr = self.xr
        if self.zr is None:
            self.zr = self.xr
        if self.scheme not in SCHEMES:
            raise ValueError("unknown scheme %r" % (self.scheme,))
        if self.precision not in PRECISIONS:
            raise ValueError("unknown precision %r" % (self.precision,))
        if self.bcs not in (1, 2) or self.order not in (2, 4) or self.eqn not in (1, 2, 3, 4):
            raise ValueError("bcs must be 1|2, order 2|4, eqn 1..4")

    # --- get_grid (ic.f90:43-51), evaluated in the storage precision of pr ---
    @property
    def real_dtype(self):
        return np.float32 if self.precision == "fp32" else np.float64

    def _axis(self, r, n):
        t = self.real_dtype
        rr = t(r)
        lo = -rr
        d = (rr - lo) / t(n - 1)
        return float(lo), float(d)

    @property
    def xl(self):
        return self._axis(self.xr, self.nx)[0]

    @property
    def yl(self):
        return self._axis(self.yr, self.ny)[0]

    @property
    def zl(self):
        return self._axis(self.zr, self.nz)[0]

    @property
    def dx(self):
        return self._axis(self.xr, self.nx)[1]

    @property
    def dy(self):
        return self._axis(self.yr, self.ny)[1]

    @property
    def dz(self):
        return self._axis(self.zr, self.nz)[1]

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def dt(self):
        """Initial complex dt (gpe.f90:78-82)."""
        return complex(self.tau, 0.0) if self.real_time else complex(0.0, -self.tau)

    def coords(self):
        t = self.real_dtype
        x = t(self.xl) + np.arange(self.nx, dtype=t) * t(self.dx)
        y = t(self.yl) + np.arange(self.ny, dtype=t) * t(self.dy)
        z = t(self.zl) + np.arange(self.nz, dtype=t) * t(self.dz)
        return x, y, z

    def with_(self, **kw):
        return replace(self, **kw)


def xr_for_spacing(dx, n):
    """Half-width that makes the reference grid have spacing dx with n points
    (both walls on the grid): xr = dx*(n-1)/2."""
    return dx * (n - 1) / 2.0
