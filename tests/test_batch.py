"""Batches: every instance of a batch against the same configuration run
alone (single-simulation path and a batch of one) and against the oracle.

The per-site arithmetic of a batched sweep is the single-simulation
kernels', and each instance reduces only its own partial sums in a fixed
order, so:
  - FP64 and FP32C rk4/rk2/euler instances equal their single-simulation
    runs bitwise, and the same instance alone and in a mix bitwise;
  - rk45 instances run the step-size control of rkqs on the device, whose
    pow() can differ in the last bit from the host libm pow() the
    single-simulation path (and the oracle) use; the error norm's
    cancellation (err is a difference of nearly equal stage sums) amplifies
    that to ~1e-12 relative in dt after a few steps, so t and dt agree to
    1e-10 and psi to 1e-12, and alone vs in a mix is still bitwise (both
    run the same device control).
FP64 instances are also held to the oracle (<= 1e-12 relative L2).
"""
import os
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig

NT = min(16, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def cfg_of(n, scheme="rk4", order=2, bcs=2, precision="fp64", om=0.15, tau=0.01, **kw):
    dims = n if isinstance(n, tuple) else (n, n, n)
    nx, ny, nz = dims
    dx = 0.4
    base = dict(nx=nx, ny=ny, nz=nz, xr=dx * (nx - 1) / 2, yr=dx * (ny - 1) / 2, zr=dx * (nz - 1) / 2,
                eqn=4, g=1.0, mu=1.0, omx=om, omy=om, omz=2 * om, tau=tau, scheme=scheme, order=order,
                bcs=bcs, precision=precision)
    base.update(kw)
    return RunConfig(**base)


def ring(cfg, frac=0.25):
    return ic.trapped_ring(cfg, frac * cfg.nx * cfg.dx)


# A mix of uneven grids, boundary conditions, orders and traps.
MIX = [
    dict(n=(24, 20, 18), bcs=2, order=2, om=0.10),
    dict(n=(70, 21, 19), bcs=1, order=4, om=0.20),
    dict(n=(33, 37, 16), bcs=2, order=4, om=0.05),
    dict(n=(16, 16, 16), bcs=1, order=2, om=0.12),
    dict(n=(40, 9, 25), bcs=2, order=2, om=0.18),
]


def _single(cfg, psi, nsteps):
    from gpe_amd.sim import Simulation
    with Simulation(cfg, psi) as sim:
        sim.advance(nsteps)
        return sim.download(np.complex128), sim.state(), sim.norm()


def _batch(entries, chunk=7):
    """entries: [(cfg, psi, nsteps)] -> [(psi, query)] in order."""
    from gpe_amd.batch import Batch
    with Batch() as b:
        ids = [b.add(c, p, n) for c, p, n in entries]
        done = b.run_to_completion(chunk=chunk)
        assert all(done[i] == 1 for i in ids), done
        return [(b.download(i), b.query(i)) for i in ids]


def test_footprint_and_validation(gpu):
    # host-side only: sizes and parameter checks (no device memory touched)
    from gpe_amd import batch
    c = cfg_of(32, scheme="rk4", precision="fp64")
    assert batch.footprint(c) == 5 * 32 ** 3 * 16
    assert batch.footprint(c.with_(precision="fp32c")) == 7 * 32 ** 3 * 8
    assert batch.footprint(c.with_(scheme="rk45")) == 8 * 32 ** 3 * 16
    assert batch.footprint(c.with_(scheme="euler", precision="fp32c")) == 4 * 32 ** 3 * 8
    with pytest.raises(Exception, match="fp64 or fp32c|GPE_FP32C"):
        batch.footprint(c.with_(precision="fp32"))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4"])
def test_batch_equals_single(precision, scheme):
    entries = []
    for k, m in enumerate(MIX):
        c = cfg_of(m["n"], scheme=scheme, order=m["order"], bcs=m["bcs"], precision=precision, om=m["om"],
                   tau=0.005)
        entries.append((c, ring(c), 6 + 3 * k))  # uneven step counts
    say("batch vs single", precision, scheme, [e[0].shape for e in entries])
    got = _batch(entries)
    for (c, p, n), (g, q) in zip(entries, got):
        want, st, nrm = _single(c, p, n)
        assert q["steps"] == n and q["status"] == 1
        assert np.array_equal(g, want), (c.shape, rel_l2(g, want))
        assert q["t"] == st["t"]
        assert abs(q["norm"] - nrm) <= 1e-12 * abs(nrm)


@pytest.mark.gpu
def test_batch_fp64_rk4_vs_oracle(oracle):
    entries = []
    for m in MIX[:3]:
        c = cfg_of(m["n"], order=m["order"], bcs=m["bcs"], om=m["om"])
        entries.append((c, ring(c), 12))
    say("batch fp64 rk4 vs oracle", [e[0].shape for e in entries])
    got = _batch(entries)
    for (c, p, n), (g, q) in zip(entries, got):
        want, _ = oracle.run(c, oracle.make_state(c.dt), p, n, NT)
        assert rel_l2(g, want) < 1e-12
        assert abs(q["norm"] - oracle.norm(c, want)) < 1e-12 * abs(q["norm"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_rk45_equals_single(precision):
    entries = []
    for k, m in enumerate(MIX):
        c = cfg_of(m["n"], scheme="rk45", order=m["order"], bcs=m["bcs"], precision=precision, om=m["om"],
                   tau=0.05)
        entries.append((c, ring(c), 4 + 2 * k))
    say("batch rk45 vs single", precision, [e[0].shape for e in entries])
    got = _batch(entries, chunk=5)
    for (c, p, n), (g, q) in zip(entries, got):
        want, st, _ = _single(c, p, n)
        assert q["steps"] == n
        assert abs(q["t"] - st["t"]) <= 1e-10 * abs(st["t"])
        assert abs(q["dt"] - st["dt"]) <= 1e-10 * abs(st["dt"])
        tol = 1e-12 if precision == "fp64" else 1e-9
        assert rel_l2(g, want) < tol, (c.shape, rel_l2(g, want))


@pytest.mark.gpu
def test_batch_fp64_rk45_vs_oracle(oracle):
    entries = []
    for m in MIX[:3]:
        c = cfg_of(m["n"], scheme="rk45", order=m["order"], bcs=m["bcs"], om=m["om"], tau=0.05)
        entries.append((c, ring(c), 5))
    say("batch fp64 rk45 vs oracle", [e[0].shape for e in entries])
    got = _batch(entries, chunk=4)
    for (c, p, n), (g, q) in zip(entries, got):
        st = oracle.make_state(c.dt)
        want, _ = oracle.run(c, st, p, n, NT)
        assert rel_l2(g, want) < 1e-12
        assert abs(q["t"] - st.t) <= 1e-10 * st.t


@pytest.mark.gpu
@pytest.mark.parametrize("precision,scheme", [("fp64", "rk4"), ("fp32c", "rk4"), ("fp64", "rk45"),
                                              ("fp32c", "rk45")])
def test_alone_equals_mix(precision, scheme):
    """The same instance alone and among others: bitwise equal, including
    an instance added while the others are already running."""
    from gpe_amd.batch import Batch
    tau = 0.05 if scheme == "rk45" else 0.005
    cfgs = [cfg_of(m["n"], scheme=scheme, order=m["order"], bcs=m["bcs"], precision=precision, om=m["om"],
                   tau=tau) for m in MIX]
    psis = [ring(c) for c in cfgs]
    steps = [9, 5, 7, 11, 6]
    say("alone vs mix", precision, scheme)
    alone = [_batch([(c, p, n)])[0] for c, p, n in zip(cfgs, psis, steps)]
    with Batch() as b:
        ids = [b.add(c, p, n) for c, p, n in zip(cfgs[:3], psis[:3], steps[:3])]
        b.run(3)
        ids += [b.add(c, p, n) for c, p, n in zip(cfgs[3:], psis[3:], steps[3:])]
        done = b.run_to_completion(chunk=2)
        assert all(done[i] == 1 for i in ids)
        mix = [(b.download(i), b.query(i)) for i in ids]
    for (ga, qa), (gm, qm) in zip(alone, mix):
        assert np.array_equal(ga, gm)
        assert qa["t"] == qm["t"] and qa["dt"] == qm["dt"] and qa["norm"] == qm["norm"]


@pytest.mark.gpu
def test_graph_equals_eager_and_remove():
    """Graph replays equal direct launches; removing an instance mid-run
    leaves the others unchanged."""
    from gpe_amd.batch import Batch
    cfgs = [cfg_of(m["n"], scheme="rk45", order=m["order"], bcs=m["bcs"], precision="fp32c", om=m["om"],
                   tau=0.05) for m in MIX[:4]]
    psis = [ring(c) for c in cfgs]
    say("graph vs eager, remove")
    res = []
    for graph in (True, False):
        with Batch() as b:
            b.use_graph(graph)
            ids = [b.add(c, p, 8) for c, p in zip(cfgs, psis)]
            b.run(4)
            b.sync()
            b.remove(ids[1])
            done = b.run_to_completion(chunk=3)
            assert set(done) == {ids[0], ids[2], ids[3]}
            res.append([b.download(i) for i in (ids[0], ids[2], ids[3])])
    for a, c in zip(res[0], res[1]):
        assert np.array_equal(a, c)
    want = _batch([(cfgs[0], psis[0], 8)])[0][0]
    assert np.array_equal(res[0][0], want)


@pytest.mark.gpu
def test_autotune_commits_nothing():
    from gpe_amd.batch import Batch
    cfgs = [cfg_of(m["n"], scheme="rk4", order=m["order"], bcs=m["bcs"], precision="fp32c", om=m["om"])
            for m in MIX[:3]]
    psis = [ring(c) for c in cfgs]
    say("autotune commits nothing")
    ref = _batch([(c, p, 10) for c, p in zip(cfgs, psis)])
    with Batch() as b:
        ids = [b.add(c, p, 10) for c, p in zip(cfgs, psis)]
        b.run(4)
        b.autotune(reps=1)
        q = b.query(ids[0])
        assert q["steps"] == 4
        b.run_to_completion(chunk=3)
        for i, (g, _) in zip(ids, ref):
            assert np.array_equal(b.download(i), g)


@pytest.mark.gpu
def test_budget_admission():
    from gpe_amd.batch import Batch, footprint
    c = cfg_of(32, precision="fp64")
    fp = footprint(c)
    say("budget admission", fp)
    with Batch(hbm_budget=2 * fp + fp // 2) as b:
        a = b.add(c, ring(c), 3)
        b2 = b.add(c, ring(c), 3)
        assert a is not None and b2 is not None
        assert b.add(c, ring(c), 3) is None  # third does not fit
        assert b.stats()["bytes_in_use"] == 2 * fp
        b.run_to_completion()
        b.remove(a)
        assert b.add(c, ring(c), 3) is not None  # freed memory is reused


@pytest.mark.gpu
@pytest.mark.parametrize("precision,scheme,variant", [("fp32c", "rk4", 0), ("fp32c", "rk4", 1), ("fp32c", "rk4", 2),
                                                      ("fp32c", "rk45", 0), ("fp32c", "rk45", 1), ("fp32c", "rk45", 2)])
def test_batch_fused_pairs_equal_unfused(precision, scheme, variant):
    """Batched fused stage pairs give bitwise the batch's stage-by-stage
    results (mixed orders: two groups)."""
    from gpe_amd.batch import Batch
    tau = 0.05 if scheme == "rk45" else 0.005
    cfgs = [cfg_of(m["n"], scheme=scheme, order=m["order"], bcs=m["bcs"], precision=precision, om=m["om"],
                   tau=tau) for m in MIX]
    psis = [ring(c) for c in cfgs]
    pairs = ["rk4_12", "rk4_34"] if scheme == "rk4" else ["ck_12", "ck_34", "ck_56"]
    say("batch fused pairs", precision, scheme, variant)
    res = []
    for fused in (False, True):
        with Batch() as b:
            if fused:
                for order in (2, 4):
                    for pr in pairs:
                        b.tune_set(precision, order, scheme, pr, variant, 16)
            ids = [b.add(c, p, 6) for c, p in zip(cfgs, psis)]
            b.run_to_completion(chunk=4)
            res.append([b.download(i) for i in ids])
            if fused:
                assert b.tune_get(precision, 2, scheme, pairs[0]) == (variant, 16)
    for a, c in zip(res[0], res[1]):
        assert np.array_equal(a, c)


# Instances at the edges of the tile geometry: 5 points per axis, smaller
# than one tile in every direction, one point past a tile multiple, z shorter
# than the prefetch window, beside an ordinary grid.
EDGE = [
    dict(n=(5, 5, 5), bcs=2, order=4, om=0.30),
    dict(n=(6, 40, 5), bcs=1, order=2, om=0.20),
    dict(n=(65, 17, 9), bcs=2, order=2, om=0.15),
    dict(n=(5, 7, 64), bcs=1, order=4, om=0.25),
    dict(n=(32, 32, 32), bcs=2, order=4, om=0.12),
]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_edge_shapes_rk4_equal_single(precision):
    entries = []
    for k, m in enumerate(EDGE):
        c = cfg_of(m["n"], order=m["order"], bcs=m["bcs"], precision=precision, om=m["om"], tau=0.005)
        entries.append((c, ring(c), 5 + 2 * k))
    say("batch edge shapes rk4 vs single", precision, [e[0].shape for e in entries])
    got = _batch(entries, chunk=4)
    for (c, p, n), (g, q) in zip(entries, got):
        want, st, nrm = _single(c, p, n)
        assert q["steps"] == n and q["status"] == 1
        assert np.array_equal(g, want), (c.shape, rel_l2(g, want))
        assert q["t"] == st["t"]
        assert abs(q["norm"] - nrm) <= 1e-12 * abs(nrm)


@pytest.mark.gpu
def test_batch_edge_shapes_fp64_rk45_vs_oracle(oracle):
    # the batch's device step control rounds pow() correctly; so does the
    # oracle in this mode, so every instance matches it bit for bit
    entries = []
    for m in EDGE:
        c = cfg_of(m["n"], scheme="rk45", order=m["order"], bcs=m["bcs"], om=m["om"], tau=0.05)
        entries.append((c, ring(c), 6))
    say("batch edge shapes fp64 rk45 vs oracle", [e[0].shape for e in entries])
    got = _batch(entries, chunk=4)
    with oracle.correctly_rounded_control_pow():
        for (c, p, n), (g, q) in zip(entries, got):
            st = oracle.make_state(c.dt)
            want, _ = oracle.run(c, st, p, n, NT)
            assert np.array_equal(g, want), (c.shape, rel_l2(g, want))
            assert q["t"] == st.t and q["dt"] == st.dt_re
