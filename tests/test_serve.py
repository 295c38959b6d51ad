"""Serving: continuous batching under a scripted arrival trace and a fake
clock.  Every delivered result must equal the same request run alone
(bitwise: the per-site arithmetic and each instance's own reductions do not
depend on the mix), whatever else was resident while it ran."""
import sys

import numpy as np
import pytest

from gpe_amd.config import RunConfig
from gpe_amd.workloads import Request


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


class FakeClock:
    def __init__(self, step=0.05):
        self.t = 0.0
        self.step = step

    def __call__(self):
        self.t += self.step
        return self.t


def _cube(n, dx, **kw):
    xr = dx * (n - 1) / 2.0
    return RunConfig(nx=n, ny=n, nz=n, xr=xr, yr=xr, zr=xr, **kw)


def trace():
    """A small mixed trace: sizes, BCs, orders, schemes and precisions."""
    spec = [
        (0.0, 24, 2, 2, "rk4", "fp64", "ring", 30),
        (0.1, 32, 1, 4, "rk45", "fp32c", "random_phase", 6),
        (0.2, 16, 2, 4, "rk4", "fp32c", "line_pair", 40),
        (0.3, 20, 1, 2, "rk45", "fp64", "ring", 5),
        (0.9, 28, 2, 2, "rk4", "fp32c", "ring", 12),
        (1.0, 16, 1, 4, "rk4", "fp64", "random_phase", 25),
        (1.5, 24, 2, 4, "rk45", "fp32c", "line_pair", 4),
        (2.0, 32, 1, 2, "rk4", "fp64", "line_pair", 9),
    ]
    reqs = []
    for i, (t, n, bcs, order, scheme, prec, kind, steps) in enumerate(spec):
        dx = {"ring": 0.4, "line_pair": 0.25, "random_phase": 0.5}[kind]
        om = (0.1, 0.1, float(np.sqrt(0.02))) if kind == "line_pair" else (0.0, 0.0, 0.0)
        cfg = _cube(n, dx, scheme=scheme, order=order, bcs=bcs, tau=0.01, eqn=4, g=1.0, mu=1.0, omx=om[0],
                    omy=om[1], omz=om[2], precision=prec)
        reqs.append(Request(i, t, cfg, steps, kind, 1000 + i))
    return reqs


def test_request_stream_shape():
    from gpe_amd.workloads import request_stream
    reqs = request_stream(rate=40.0, duration=10.0, seed=5)
    assert 300 < len(reqs) < 500
    assert all(48 <= r.cfg.nx <= 256 and r.cfg.nx % 16 == 0 for r in reqs)
    assert all(200 <= r.steps <= 5000 for r in reqs)
    assert {r.cfg.scheme for r in reqs} == {"rk4", "rk45"}
    assert {r.cfg.precision for r in reqs} == {"fp64", "fp32c"}
    assert all(a.arrival < b.arrival for a, b in zip(reqs, reqs[1:]))
    # same seed, same stream
    again = request_stream(rate=40.0, duration=10.0, seed=5)
    assert [(r.arrival, r.cfg.nx, r.steps) for r in reqs] == [(r.arrival, r.cfg.nx, r.steps) for r in again]


@pytest.mark.gpu
@pytest.mark.parametrize("budget_fields", [0, 3])
def test_served_results_equal_alone(budget_fields):
    """budget_fields=3: an HBM budget so small that requests wait in the
    queue for memory; 0: everything admitted at once."""
    from gpe_amd.batch import Batch, footprint
    from gpe_amd.serve import Server
    reqs = trace()
    budget = 0
    if budget_fields:
        budget = budget_fields * max(footprint(r.cfg) for r in reqs)
    say("serve trace", len(reqs), "budget", budget)
    got = {}

    def deliver(r, psi, info):
        got[r.rid] = (psi, info)

    with Server(hbm_budget=budget, target_chunk_s=0.001, deliver=deliver) as s:
        m = s.serve(reqs, clock=FakeClock())
    assert m["completed"] == len(reqs) and m["ok"] == len(reqs), m
    for r in reqs:
        with Batch() as b:
            i = r.add_to(b)  # the same initial-state path as the server's
            b.run_to_completion(chunk=5)
            want = b.download(i)
            q = b.query(i)
        g, info = got[r.rid]
        assert info["steps"] == r.steps
        assert np.array_equal(g, want), r
        assert info["t"] == q["t"]


@pytest.mark.gpu
def test_checkpoint_and_resume_equal_alone(tmp_path):
    """A serving window that ends with requests running checkpoints them;
    a new server resumes them; every result still equals its run alone."""
    from gpe_amd.batch import Batch
    from gpe_amd.serve import Server
    reqs = trace()
    ck = str(tmp_path / "ck")
    say("serve checkpoint / resume")
    got = {}

    def deliver(r, psi, info):
        got[r.rid] = psi

    with Server(target_chunk_s=0.001, deliver=deliver) as s:
        m1 = s.serve(reqs, duration=1.2, drain=False, clock=FakeClock(), checkpoint_dir=ck)
    assert m1["unfinished"] > 0
    # the resumed server also takes the requests the first window did not
    # admit, under a resident-site cap that the restored requests count toward
    later = [Request(r.rid, r.arrival - 1.2, r.cfg, r.steps, r.ic_kind, r.seed) for r in reqs if r.arrival >= 1.2]
    cap = max(r.cfg.nx * r.cfg.ny * r.cfg.nz for r in reqs)
    with Server(target_chunk_s=0.001, deliver=deliver, max_sites=cap) as s:
        m2 = s.serve(later, clock=FakeClock(), resume_dir=ck)
    assert sorted(got) == sorted(r.rid for r in reqs), (sorted(got), m1, m2)
    for r in reqs:
        with Batch() as b:
            i = r.add_to(b)
            b.run_to_completion(chunk=5)
            want = b.download(i)
        assert np.array_equal(got[r.rid], want), r
