"""GPU parity: the HIP library against the oracle on the same seeded inputs.

FP64 is the drop-in parity path: <= 1e-12 relative L2 in psi, norm and
energy (the bar the reference's FP64 path is held to).  The FP32 forms are
held to 1e-5 relative L2 against the FP64 oracle.  Every test prints the
case before touching the GPU.
"""
import os
import sys

import numpy as np
import pytest

from gpe_amd import ic
from gpe_amd.config import RunConfig, xr_for_spacing

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def small_cfg(**kw):
    # non-cubic, x spans two 64-wide tiles (the second partial), odd y and z
    base = dict(nx=70, ny=21, nz=19, xr=7.0, yr=4.0, zr=3.6, eqn=4, g=1.0, mu=1.0,
                omx=0.3, omy=0.2, omz=0.4, tau=0.005)
    base.update(kw)
    return RunConfig(**base)


def ring_ic(cfg):
    psi = ic.vortex_ring(cfg, 2.0)
    rng = np.random.default_rng(7)
    return psi * (1 + 0.05 * rng.standard_normal(cfg.shape))


@pytest.mark.parametrize("eqn", [1, 2, 3, 4])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_rhs_fp64(gpu, oracle, eqn, order, bcs):
    from gpe_amd.sim import Simulation
    extra = {}
    if eqn in (1, 3):
        extra = dict(diss_amp=0.3, urhs=0.2)
    elif eqn == 2:
        extra = dict(urhs=-0.1)
    cfg = small_cfg(eqn=eqn, order=order, bcs=bcs, **extra)
    psi = ring_ic(cfg)
    say("rhs fp64", eqn, order, bcs, cfg.shape)
    with Simulation(cfg, psi) as sim:
        got = sim.rhs()
    want = oracle.rhs(cfg, psi, NT)
    err = rel_l2(got, want)
    say("rel L2 %.3e, bitwise-equal sites %.4f" % (err, np.mean(got == want)))
    assert err < 1e-14


@pytest.mark.parametrize("scheme", ["euler", "rk2", "rk4", "rk45"])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("bcs", [1, 2])
def test_steps_fp64(gpu, oracle, scheme, order, bcs):
    from gpe_amd.sim import Simulation
    cfg = small_cfg(scheme=scheme, order=order, bcs=bcs, tau=0.002 if scheme == "euler" else 0.005)
    psi = ring_ic(cfg)
    n = 30
    say("steps fp64", scheme, order, bcs, n)
    st = oracle.make_state(cfg.dt)
    want, wnorms = oracle.run(cfg, st, psi, n, NT, with_norms=True)
    with Simulation(cfg, psi) as sim:
        gnorms = sim.advance(n, record_norm=True)
        got = sim.download()
        gst = sim.state()
    assert rel_l2(got, want) < 1e-12
    assert np.max(np.abs(gnorms - wnorms) / wnorms) < 1e-12
    # rk45: errmax is a heavily cancelling difference of stage values, so
    # ulp-level differences (hypot) move dt at the 1e-11 level; psi above is
    # the parity check
    tt = 1e-9 if scheme == "rk45" else 1e-14
    assert gst["t"] == pytest.approx(st.t, rel=tt, abs=1e-15)
    assert gst["dt"].real == pytest.approx(st.dt_re, rel=1e-9 if scheme == "rk45" else 1e-14)
    assert gst["p"] == n


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("scheme", ["rk4", "rk45"])
def test_tile_variants_fp64(gpu, oracle, variant, scheme):
    # every compiled tile variant, with short z-chunks so chunk seams are
    # crossed, must give the same field as the oracle
    from gpe_amd import _lib
    from gpe_amd.sim import Simulation
    cfg = small_cfg(nx=70, ny=37, nz=29, scheme=scheme, order=4, bcs=1, tau=0.005)
    psi = ring_ic(cfg)
    n = 12
    say("variant", variant, scheme)
    want, _ = oracle.run(cfg, oracle.make_state(cfg.dt), psi, n, NT)
    with Simulation(cfg, psi) as sim:
        for k in _lib.STAGE_KINDS:
            if k.startswith("ck_") and variant in (0, 1):
                # FP64 Cash-Karp stages refuse the (4,4) and (8,2) tiles
                # and keep their default
                with pytest.raises(_lib.GpeError):
                    sim.tune_set(k, variant, 8)
                sim.tune_set(k, 4, 8)
                continue
            sim.tune_set(k, variant, 8)
        sim.advance(n)
        got = sim.download()
    assert rel_l2(got, want) < 1e-12


def test_rk45_step_control_fp64(gpu, oracle):
    # a too-large first step: rejections and dt adaptation must match
    from gpe_amd.sim import Simulation
    cfg = small_cfg(scheme="rk45", order=4, bcs=2, tau=0.5)
    psi = ring_ic(cfg)
    say("rk45 control", cfg.shape)
    st = oracle.make_state(cfg.dt)
    rejects, dts = [], []
    cur = psi
    for _ in range(6):
        cur = oracle.solver(cfg, st, cur, NT)
        rejects.append(st.n_rejects)
        dts.append(st.dt_re)
    with Simulation(cfg, psi) as sim:
        g_rej, g_dts = [], []
        for _ in range(6):
            sim.advance(1)
            s = sim.state()
            g_rej.append(s["n_rejects"])
            g_dts.append(s["dt"].real)
        got = sim.download()
    assert sum(rejects) > 0
    assert g_rej == rejects
    assert np.allclose(g_dts, dts, rtol=1e-9, atol=0)
    assert rel_l2(got, cur) < 1e-11


def test_imaginary_time_rk2_renormalise(gpu, oracle):
    # trap_vortex-like: imaginary time, rk2, renormalisation by the norm
    from gpe_amd.sim import Simulation
    cfg = small_cfg(scheme="rk2", order=2, bcs=2, real_time=False, tau=0.001, g=5.0, mu=8.0)
    psi = ring_ic(cfg)
    say("imaginary time rk2")
    st = oracle.make_state(cfg.dt)
    cur = psi
    with Simulation(cfg, psi) as sim:
        for _ in range(5):
            cur = oracle.solver(cfg, st, cur, NT)
            nrm = oracle.norm(cfg, cur)
            cur = oracle.renormalise(cfg, cur, 1.0, nrm)
            sim.advance(1)
            gn = sim.norm()
            assert gn == pytest.approx(nrm, rel=1e-12)
            sim.renormalise(1.0, gn)
        got = sim.download()
        s = sim.state()
    assert rel_l2(got, cur) < 1e-12
    assert s["im_t"] == pytest.approx(st.im_t, rel=1e-14)


@pytest.mark.parametrize("order", [2, 4])
def test_reductions_fp64(gpu, oracle, order):
    from gpe_amd.sim import Simulation
    cfg = small_cfg(order=order, bcs=2)
    psi = ring_ic(cfg)
    say("reductions", order)
    with Simulation(cfg, psi) as sim:
        assert sim.norm() == pytest.approx(oracle.norm(cfg, psi), rel=1e-13)
        assert sim.energy() == pytest.approx(oracle.energy(cfg, psi), rel=1e-12)
        assert sim.mass() == pytest.approx(oracle.mass(cfg, psi), rel=1e-13)


def test_dropin_solver_step_with_halo(gpu, oracle):
    # the Fortran-facing call: psi%old(0:nx1, js-2:je+2, ks-2:ke+2) in, interior out
    from gpe_amd.sim import Simulation
    cfg = small_cfg(scheme="rk4", order=4, bcs=1)
    psi = ring_ic(cfg)
    say("drop-in solver_step")
    h = 2
    old = np.full((cfg.nz + 2 * h, cfg.ny + 2 * h, cfg.nx), np.nan + 0j)
    st = oracle.make_state(cfg.dt)
    cur = psi
    with Simulation(cfg) as sim:
        for _ in range(3):
            old[h:-h, h:-h, :] = cur
            out = sim.solver_step(old, halo=h)
            cur = oracle.solver(cfg, st, cur, NT)
            assert rel_l2(out, cur) < 1e-13
        assert sim.state()["t"] == pytest.approx(st.t, rel=1e-15)


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_fp32_forms_track_fp64_rk4(gpu, oracle, precision):
    from gpe_amd.sim import Simulation
    cfg64 = small_cfg(scheme="rk4", order=4, bcs=1)
    cfg = cfg64.with_(precision=precision)
    psi = ring_ic(cfg64)
    n = 40
    say("fp32 form rk4", precision)
    want, _ = oracle.run(cfg64, oracle.make_state(cfg64.dt), psi, n, NT)
    with Simulation(cfg, psi) as sim:
        sim.advance(n)
        got = sim.download(np.complex128)
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_fp32_forms_track_fp64_rk45(gpu, oracle, precision):
    # adaptive steps: the FP32 and FP64 runs choose their own dt, so each step
    # starts the GPU run from the oracle's (dt, t); steps where both accept
    # the same dt are compared
    from gpe_amd.sim import Simulation
    cfg64 = small_cfg(scheme="rk45", order=4, bcs=1, tau=0.01)
    cfg = cfg64.with_(precision=precision)
    psi = ring_ic(cfg64)
    say("fp32 form rk45", precision)
    st = oracle.make_state(cfg64.dt)
    cur = psi
    same = 0
    with Simulation(cfg, psi) as sim:
        for _ in range(12):
            dt0, t0 = complex(st.dt_re, st.dt_im), st.t
            sim.set_state(dt0, t=t0)
            sim.upload(cur)
            cur = oracle.solver(cfg64, st, cur, NT)
            sim.advance(1)
            s = sim.state()
            # plain FP32 runs its step control in single precision
            if abs(s["t"] - st.t) <= (1e-6 if precision == "fp32" else 1e-12) * abs(st.t):
                same += 1
                got = sim.download(np.complex128)
                assert rel_l2(got, cur) < 1e-6
    assert same >= 8


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_fp32_forms_at_fp64_steps(gpu, oracle, precision):
    # every one of the oracle's accepted steps, taken by the FP32 form from
    # the oracle's field with the oracle's accepted dt (tolerance 1e20: the
    # first attempt is accepted): psi within 1e-6 of the oracle's on all 12,
    # and the unscaled error norm max|err/scale| close to the oracle's (its
    # k fields carry single-precision rounding into err, a difference of
    # fifth- and fourth-order solutions ~1e-6 of the scale)
    from gpe_amd.sim import Simulation
    cfg64 = small_cfg(scheme="rk45", order=4, bcs=1, tau=0.01)
    cfg = cfg64.with_(precision=precision, eps=1e20)
    psi = ring_ic(cfg64)
    say("fp32 form at fp64 steps", precision)
    # plain FP32 runs its step control (t, dt) in single precision
    t_rel = 1e-6 if precision == "fp32" else 1e-14
    st = oracle.make_state(cfg64.dt)
    cur = psi
    worst_psi = worst_err = 0.0
    rejected = compared = 0
    with Simulation(cfg, psi) as sim:
        for i in range(12):
            t0, start = st.t, cur
            cur = oracle.solver(cfg64, st, cur, NT)
            h = st.t - t0
            rejected += st.n_rejects
            sim.set_state(h, t=t0)
            sim.upload(start)
            sim.advance(1)
            s = sim.state()
            assert s["n_rejects"] == 0
            assert s["t"] == pytest.approx(st.t, rel=t_rel)
            got = sim.download(np.complex128)
            e_psi = rel_l2(got, cur)
            worst_psi = max(worst_psi, e_psi)
            e_norm = abs(s["errmax"] * cfg.eps / (st.errmax * cfg64.eps) - 1.0)
            say("step %d dt %.6e rejections %d: psi rel L2 %.3e, error norm rel diff %.3e"
                % (i, h, st.n_rejects, e_psi, e_norm))
            # after a rejection the oracle's accepted attempt still uses the
            # error scale formed with the first attempt's dt (solve.f90:176-178),
            # the forced single attempt its own: compare steps without one
            if st.n_rejects == 0:
                worst_err = max(worst_err, e_norm)
                compared += 1
    say("worst psi %.3e, worst error norm %.3e over %d steps, oracle rejections %d"
        % (worst_psi, worst_err, compared, rejected))
    assert worst_psi < 1e-6
    assert compared >= 8 and worst_err < 0.05


def test_ring_64cube_rk4_1000_steps(gpu, oracle):
    # 64^3, dx = 0.4, reflective, 2nd order, RK4 dt = 0.01, 1000 steps, V = 0
    cfg = RunConfig(nx=64, ny=64, nz=64, xr=xr_for_spacing(0.4, 64), bcs=2, order=2, scheme="rk4",
                    tau=0.01, g=1.0, mu=1.0)
    psi = ic.vortex_ring(cfg, 4.0)
    say("64^3 ring rk4 x1000")
    from gpe_amd.sim import Simulation
    st = oracle.make_state(cfg.dt)
    want, wnorms = oracle.run(cfg, st, psi, 1000, NT, with_norms=True)
    with Simulation(cfg, psi) as sim:
        gn = sim.advance(1000, record_norm=True)
        got = sim.download()
        ge = sim.energy()
    err = rel_l2(got, want)
    say("64^3 ring rel L2 = %.3e" % err)
    assert err < 1e-12
    assert abs(gn[-1] - wnorms[-1]) / wnorms[-1] < 1e-12
    assert ge == pytest.approx(oracle.energy(cfg, want, NT), rel=1e-12)
    with Simulation(cfg, psi) as sim:
        assert sim.linelength() == pytest.approx(oracle.linelength(cfg, psi), rel=1e-12)
    # the compensated FP32 form after the same 1000 RK4 steps: <= 1e-5
    with Simulation(cfg.with_(precision="fp32c"), psi) as sim:
        sim.advance(1000)
        got32 = sim.download(np.complex128)
        z32 = sim.zeros()
    err32 = rel_l2(got32, want)
    say("64^3 ring fp32c rel L2 = %.3e" % err32)
    assert err32 < 1e-5
    # vortex-core positions after 1000 steps (the reference's zeros file):
    # the GPU's FP64 field and the oracle's give the same core points to
    # 1e-9 dx; the compensated form's cores sit within 1e-4 dx of them
    from gpe_amd.accuracy import core_offsets
    with Simulation(cfg, got) as sim:
        z64 = sim.zeros()
    zo = oracle.zeros(cfg, want, 0)
    d64 = core_offsets(z64, zo, cfg.dx)
    d32 = core_offsets(z32, zo, cfg.dx)
    say("64^3 ring core offsets fp64 %s fp32c %s" % (d64, d32))
    assert d64["points"][0] == d64["points"][1] > 50
    assert d64["max_dx"] < 1e-9
    assert d32["mean_dx"] < 1e-4 and d32["frac_over_tenth_dx"] == 0.0


def test_line_pair_256cube(gpu, oracle):
    # 256^3, dx = 0.25, 4th order, reflective, trap, RK4 dt = 0.005 (a few steps)
    cfg = RunConfig(nx=256, ny=256, nz=256, xr=xr_for_spacing(0.25, 256), bcs=2, order=4, scheme="rk4",
                    tau=0.005, g=1.0, mu=1.0, omx=0.1, omy=0.1, omz=float(np.sqrt(0.02)))
    psi = ic.line_pair(cfg)
    n = 4
    say("256^3 line pair rk4 x%d" % n)
    from gpe_amd.sim import Simulation
    want, _ = oracle.run(cfg, oracle.make_state(cfg.dt), psi, n, NT)
    with Simulation(cfg, psi) as sim:
        sim.advance(n)
        got = sim.download()
    assert rel_l2(got, want) < 1e-12


def test_random_phase_512_rk45_fp64_matches_oracle(gpu, oracle):
    # the headline workload at full size (512^3 random phase, periodic,
    # 4th order, RKF45 tol 1e-6): one accepted step of the GPU's FP64 path
    # against the oracle on the same field, with the same dt and t
    from gpe_amd.sim import Simulation
    cfg = RunConfig(nx=512, ny=512, nz=512, xr=xr_for_spacing(0.5, 512), bcs=1, order=4, scheme="rk45",
                    tau=0.05, g=1.0, mu=1.0)
    say("512^3 random phase rk45 fp64 vs oracle")
    with Simulation(cfg, None) as sim:
        sim.random_phase(seed=20261202, kmax=0.5)
        psi = sim.download()
        sim.advance(1)
        got = sim.download()
        s = sim.state()
    st = oracle.make_state(cfg.dt)
    want = oracle.solver(cfg, st, psi, NT)
    del psi
    err = rel_l2(got, want)
    say("rel L2 %.3e, t %.12f / %.12f, rejects %d / %d" % (err, s["t"], st.t, s["n_rejects"], st.n_rejects))
    assert err < 1e-12
    assert s["t"] == pytest.approx(st.t, rel=1e-12)
    assert s["dt"].real == pytest.approx(st.dt_re, rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_plane_wave_512_rk45_matches_scalar_recursion(gpu, precision):
    # full config-2 size: a plane wave stays a plane wave; every site must
    # follow the scalar Cash-Karp recursion of its amplitude
    from gpe_amd.sim import Simulation
    from tests.test_oracle import _ck_scalar
    cfg = RunConfig(nx=512, ny=512, nz=512, xr=xr_for_spacing(0.5, 512), bcs=1, order=4, scheme="rk45",
                    tau=0.05, g=1.0, mu=1.0, precision=precision)
    m = (3, -2, 5)
    lam = 0.0
    for ma, nn, d in ((m[0], cfg.nx, cfg.dx), (m[1], cfg.ny, cfg.dy), (m[2], cfg.nz, cfg.dz)):
        th = 2 * np.pi * ma / nn
        lam += (-2 * np.cos(2 * th) + 32 * np.cos(th) - 30) / (12 * d ** 2)
    a0 = 0.9 + 0.1j
    f = lambda a: -1j * (-0.5 * lam + abs(a) ** 2 - 1.0) * a  # noqa: E731
    psi = a0 * ic.plane_wave(cfg, m)
    say("512^3 plane wave rk45", precision)
    ts = [0.0]
    with Simulation(cfg, psi) as sim:
        for _ in range(3):
            sim.advance(1)
            ts.append(sim.state()["t"])
        got = sim.download(np.complex128)
    # scalar Cash-Karp with the same accepted steps: every site must follow it
    a = a0
    for t0, t1 in zip(ts[:-1], ts[1:]):
        a, _ = _ck_scalar(f, a, t1 - t0)
    # and the scalar step-size control from the same initial dt lands on the
    # same times (up to the error norm's cancellation, see test_steps_fp64)
    b, dt, t = a0, 0.05, 0.0
    for _ in range(3):
        scale = abs(b) + dt * abs(f(b)) + 1e-30
        while True:
            y5, e = _ck_scalar(f, b, dt)
            em = abs(e) / scale / 1e-6
            if em <= 1:
                break
            dt = max(abs(0.9 * dt * em ** -0.25), 0.1 * abs(dt))
        t += dt
        dt = 0.9 * dt * em ** -0.2 if em > (5 / 0.9) ** (1 / -0.2) else 5 * dt
        b = y5
    # steps reach dt ~ 0.6: the lattice Laplacian's rounding (a ~1e3-fold
    # cancellation for this long wave) dominates the departure from the
    # scalar recursion
    err = rel_l2(got, a * ic.plane_wave(cfg, m))
    say("t %.9f (scalar control %.9f), rel L2 vs scalar recursion %.3e" % (ts[-1], t, err))
    if precision == "fp64":
        assert ts[-1] == pytest.approx(t, rel=1e-6)
        assert err < 1e-9
    else:
        # FP32 storage: the steps the exact control takes (0.05, 0.25, 0.6)
        # lie far outside the explicit scheme's stability region for the
        # lattice's shortest waves: the 4th-order stencil's largest symbol is
        # 16/(3 dx^2) per axis, so |z| = dt * 0.5 * 3 * 16/(3 dx^2) = 32 dt,
        # about 19 at dt = 0.6.  The plane wave has no content there in FP64
        # (rounding ~1e-16 grows to ~1e-10, the FP64 case above), but FP32
        # rounding (~6e-8) grows to ~5e-3 over those three steps (measured
        # with the steps forced on the compensated form: 5.4e-3).  The error
        # norm sees that growth, so the compensated form takes smaller steps,
        # as it should; psi follows the exact recursion for the steps it took
        assert 0.2 * t < ts[-1] <= t * (1 + 1e-6)
        assert err < 1e-5
