"""Write tests/golden/line_pair_256_rk4_2000.npz: the oracle's run of the
256^3 line pair in a trap for its full 2000 RK4 steps.

Workload: 256^3 grid, dx = 0.25, reflective walls, 4th-order FD, RK4 with
dt = 0.005, eqn 4 with g = mu = 1 and the trap omx = omy = 0.1,
omz = sqrt(0.02); initial state gpe_amd.ic.line_pair (two straight lines
along z at (+-3, 0) in the Thomas-Fermi background sqrt(max(0, 10 - V))).
The oracle (oracle/gpe_oracle.c, FP64) advances it 2000 solver() calls,
the time loop of gpe.f90:136 with the RK4 of solve.f90:109-154.

Stored (a few MiB, instead of the 256 MiB field):
  norm        get_norm after every step (gpe.f90:243-249)        (2000,)
  energy      energy() every 250 steps, index 0 = the initial field (9,)
  sample      psi at 2^18 sites: site s_i = (i*SAMPLE_MUL + SAMPLE_ADD) mod 2^24,
              a bijection of the 2^24 sites, so the samples are distinct
              and spread over the whole box                       (2^18,)
  core_*      the zeros-file core points (io.f90:886-1393) of the final
              field whose cell lies in a plane k = 0 mod 8: section, cell,
              x, y, z
  core_count  the number of core points of the whole final field

The oracle takes about 15-25 minutes on 8 host threads:

    make -C oracle && python tests/golden/make_line_pair_256.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from gpe_amd import ic  # noqa: E402
from gpe_amd.config import RunConfig, xr_for_spacing  # noqa: E402
from oracle import pyoracle  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "line_pair_256_rk4_2000.npz")
N = 256
STEPS = 2000
ENERGY_EVERY = 250
NSAMPLE = 1 << 18
SAMPLE_MUL = 0x9E3779B1  # odd: i -> i*MUL + ADD is a bijection mod 2^24
SAMPLE_ADD = 12345
CORE_PLANE_STRIDE = 8


def workload():
    return RunConfig(nx=N, ny=N, nz=N, xr=xr_for_spacing(0.25, N), bcs=2, order=4, scheme="rk4", tau=0.005,
                     g=1.0, mu=1.0, omx=0.1, omy=0.1, omz=float(np.sqrt(0.02)))


def sample_sites():
    i = np.arange(NSAMPLE, dtype=np.uint64)
    return ((i * np.uint64(SAMPLE_MUL) + np.uint64(SAMPLE_ADD)) & np.uint64(N ** 3 - 1)).astype(np.int64)


def core_subset(z):
    """Core points whose cell lies in a plane k = 0 mod CORE_PLANE_STRIDE."""
    k = z["cell"] // (N * N)
    return z[k % CORE_PLANE_STRIDE == 0]


def main():
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = workload()
    psi = ic.line_pair(cfg)
    st = pyoracle.make_state(cfg.dt)
    energy = [pyoracle.energy(cfg, psi, nt)]
    norms = []
    t0 = time.perf_counter()
    done = 0
    while done < STEPS:
        psi, wn = pyoracle.run(cfg, st, psi, ENERGY_EVERY, nt, with_norms=True)
        norms.append(wn)
        done += ENERGY_EVERY
        energy.append(pyoracle.energy(cfg, psi, nt))
        print("oracle %d/%d steps, %.0f s" % (done, STEPS, time.perf_counter() - t0), flush=True)
    z = pyoracle.zeros(cfg, psi, 0)
    zs = core_subset(z)
    np.savez_compressed(OUT, norm=np.concatenate(norms), energy=np.array(energy),
                        sample=psi.reshape(-1)[sample_sites()], t=np.array([st.t]),
                        core_section=zs["section"], core_cell=zs["cell"], core_x=zs["x"], core_y=zs["y"],
                        core_z=zs["z"], core_count=np.array([len(z)]))
    print("wrote %s: %d core points (%d stored), t = %.6f" % (OUT, len(z), len(zs), st.t))


if __name__ == "__main__":
    main()
