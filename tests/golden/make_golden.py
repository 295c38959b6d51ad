"""Write tests/golden/ref_derivs.npz from the reference's own derivs module.

Runs oracle/_ref/ref_derivs_{d,s} (the reference's parameters.f90 and
derivs.f90 compiled unchanged, grid 12 x 10 x 9, see oracle/Makefile) on
seeded random fields for every (bcs, order) and stores input and output:
the laplacian as deriv_xx + deriv_yy + deriv_zz (variables.f90:68) and
deriv_x.  These vectors let the CPU tests pin the oracle's stencils and
boundary maps where the reference cannot be built.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from gpe_amd.config import RunConfig  # noqa: E402

NX, NY, NZ = 12, 10, 9


def main():
    out = {}
    rng = np.random.default_rng(20261202)
    for prec, ct, exe in (("fp64", np.complex128, "ref_derivs_d"), ("fp32", np.complex64, "ref_derivs_s")):
        path = os.path.join(ROOT, "oracle", "_ref", exe)
        for bcs in (1, 2):
            for order in (2, 4):
                cfg = RunConfig(nx=NX, ny=NY, nz=NZ, bcs=bcs, order=order, xr=3.3, yr=2.1, zr=1.7, precision=prec)
                f = (rng.standard_normal(cfg.shape) + 1j * rng.standard_normal(cfg.shape)).astype(ct)
                with tempfile.TemporaryDirectory() as d:
                    f.tofile(os.path.join(d, "in.bin"))
                    subprocess.check_call([path, str(bcs), str(order), repr(cfg.dx), repr(cfg.dy), repr(cfg.dz),
                                           os.path.join(d, "in.bin"), os.path.join(d, "out.bin")])
                    res = np.fromfile(os.path.join(d, "out.bin"), dtype=ct)
                key = "%s_b%d_o%d" % (prec, bcs, order)
                out[key + "_in"] = f
                out[key + "_lap"] = res[: f.size].reshape(cfg.shape)
                out[key + "_dx"] = res[f.size:].reshape(cfg.shape)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ref_derivs.npz"), **out)
    print("wrote ref_derivs.npz with", len(out), "arrays")


if __name__ == "__main__":
    main()
