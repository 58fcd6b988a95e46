"""Output series and checkpoints of the QD mediator on the GPU.

The reference's timeMonitor (strikeSlip_linearElastic_qd.cpp:417-463) writes
fault and boundary series every stride1D steps, body fields every stride2D
steps and checkpoints every `interval` steps (then stops); a run restarted
from a checkpoint reloads the state and continues.  The files are read back
here with a plain numpy reader of the PETSc binary format (big-endian class
id 1211214, length, doubles) and checked against the driver's own state, the
closed forms the mediator fixes (bcR = vL t / faultTypeScale + bcRShift, bcL =
slip / faultTypeScale), the oracle's muxDy for the stress series, and, for
the restart, the oracle's RK run from the checkpointed state.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import OracleSbp, lib as olib, ptr as optr
from test_gpu_fault_qd import _oracle_qd, _Run, _small_cfg

pytestmark = pytest.mark.gpu

CLASS_ID = 1211214


def read_vecs(path):
    """Every record of a PETSc binary Vec file."""
    raw = open(path, "rb").read()
    out, p = [], 0
    while p < len(raw):
        cid, n = np.frombuffer(raw, dtype=">i4", count=2, offset=p)
        assert cid == CLASS_ID, path
        out.append(np.frombuffer(raw, dtype=">f8", count=n, offset=p + 8).astype(np.float64))
        p += 8 + 8 * n
    return out


def read_lines(path):
    return [float(x) for x in open(path).read().split()]


def _mediator(cfg, reuse=2):
    from scycle_amd import QdMediator
    g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0, precond="fdiag64", reuse=reuse)
    g.set_state(cfg.psi0, cfg.slip0, np.zeros(cfg.Nz))
    return g


def test_qd_output_series(tmp_path):
    """stride1D = 3, stride2D = 5 over 20 RK32 steps: the step-0 record, then
    one per stride; every series has one record per line of med_time1D.txt."""
    cfg = _small_cfg(25, 25, 4, True)
    Nz, Ny, fts = cfg.Nz, cfg.Ny, cfg.fault_type_scale
    d = str(tmp_path) + "/"
    g = _mediator(cfg)
    g.set_output(d, stride1D=3, stride2D=5)
    steps = 20
    tEnd, n, dts = g.integrate(steps, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    assert n == steps
    st, u = g.get_state(), g.get_u()
    t = read_lines(d + "med_time1D.txt")
    times = np.cumsum(dts)
    rec1 = [2, 5, 8, 11, 14, 17]  # steps 3, 6, ..., 18
    assert len(t) == 1 + len(rec1)
    assert t[0] == 0.0 and np.allclose(t[1:], times[rec1], rtol=1e-14)
    dt1 = read_lines(d + "med_dt1D.txt")
    assert np.allclose(dt1, [1e-3] + list(dts[rec1]), rtol=1e-14)
    series = {k: read_vecs(d + k) for k in ("slip", "slipVel", "tauP", "tauQSP", "strength", "psi",
                                            "sNEff", "surfDisp", "momBal_bcR", "momBal_bcL",
                                            "momBal_bcT", "momBal_bcB", "momBal_bcRShift")}
    for k, v in series.items():
        assert len(v) == len(t), k
        assert len(v[0]) == (Ny if k in ("surfDisp", "momBal_bcT", "momBal_bcB") else Nz), k
    # step 0: the state as set
    assert np.array_equal(series["psi"][0], cfg.psi0) and np.array_equal(series["slip"][0], cfg.slip0)
    # sigma_N is a parameter; top and bottom data are zero in this mediator
    for r in series["sNEff"]:
        assert np.array_equal(r, cfg.sN)
    for k in ("momBal_bcT", "momBal_bcB"):
        assert all(not r.any() for r in series[k])
    shift = np.zeros(Nz) if cfg.bcRShift is None else cfg.bcRShift
    for i in range(1, len(t)):
        ti = times[rec1[i - 1]]
        # remote loading and the fault datum of the d_dt at that state
        assert np.allclose(series["momBal_bcR"][i], cfg.vL * ti / fts + shift, rtol=1e-13, atol=0)
        assert np.allclose(series["momBal_bcL"][i] * fts, series["slip"][i], rtol=1e-15, atol=0)
        # the bcRShift series holds bcRShift first and bcR after it (LinearElastic::writeStep1D)
        assert np.array_equal(series["momBal_bcRShift"][i], series["momBal_bcR"][i])
        # tau = tauQS - eta V
        assert np.allclose(series["tauP"][i], series["tauQSP"][i] - cfg.eta * series["slipVel"][i],
                           rtol=1e-14, atol=1e-14 * np.abs(series["tauQSP"][i]).max())
    assert np.array_equal(series["momBal_bcRShift"][0], shift)
    # 2-D series at steps 0, 5, 10, 15, 20; the last equals the final state
    t2 = read_lines(d + "med_time2D.txt")
    assert len(t2) == 5 and np.allclose(t2[1:], times[[4, 9, 14, 19]], rtol=1e-14)
    us, sxy = read_vecs(d + "momBal_u"), read_vecs(d + "momBal_sxy")
    assert len(us) == 5 and len(sxy) == 5
    assert np.array_equal(us[-1], u)
    o = OracleSbp(cfg.order, Ny, Nz, cfg.Ly, cfg.Lz, cfg.mu, cfg.y, cfg.z, var_grid=cfg.var_grid,
                  bc=cfg.mat_bc)
    ref = o.muxdy(u)
    err = np.abs(sxy[-1] - ref).max() / np.abs(ref).max()
    print("sxy vs oracle muxDy", err, flush=True)
    assert err <= 1e-13
    # the 1-D series at the last stride step hold the state of that step: a
    # fresh run of 18 steps ends there.  Reading the deferred-check flag at an
    # output step shifts when later solves defer their check, and a deferred
    # solve forms the fault traction from the basis tractions, so the two
    # runs agree to rounding, not bitwise.
    g2 = _mediator(cfg)
    g2.integrate(18, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    s18 = g2.get_state()
    for k, kk in (("psi", "psi"), ("slip", "slip"), ("slipVel", "V"), ("tauP", "tau")):
        assert np.allclose(series[k][-1], s18[kk], rtol=1e-11, atol=1e-11 * np.abs(s18[kk]).max()), k
    u18 = g2.get_u()[::Nz]
    assert np.allclose(series["surfDisp"][-1], u18, rtol=1e-11, atol=1e-11 * np.abs(u18).max())
    # the final state is the last 2-D record and, step 20 not being a
    # stride1D step, not in the 1-D series
    assert not np.array_equal(series["psi"][-1], st["psi"])
    g.close()
    g2.close()


def test_qd_output_final_time(tmp_path):
    """A record is also written when the run reaches maxTime exactly, whatever
    the stride."""
    cfg = _small_cfg(21, 21, 2, False)
    d = str(tmp_path) + "/"
    g0 = _mediator(cfg)
    _, _, dts = g0.integrate(12, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    g0.close()
    T = float(np.sum(dts[:9]) + 0.5 * dts[9])  # inside step 10
    g = _mediator(cfg)
    g.set_output(d, stride1D=1000, stride2D=0)
    tEnd, n, _ = g.integrate(100, tol=1e-8, init_dt=1e-3, max_dt=1e10, max_time=T)
    assert tEnd == T
    t = read_lines(d + "med_time1D.txt")
    assert len(t) == 2 and t[0] == 0.0 and abs(t[1] / T - 1) <= 1e-15
    assert not os.path.exists(d + "med_time2D.txt")
    g.close()


def _ascii(d, name):
    return read_lines(d + name)[0]


@pytest.mark.parametrize("reuse", [0, 2])
def test_qd_checkpoint_restart(tmp_path, reuse):
    """enableCheckpointing with interval 6: a run of up to 15 steps writes a
    checkpoint at step 0 and at step 6 and stops there.  A new mediator loads
    it and runs on; its step count restarts at 0 (the reference's step-0
    monitor call resets it), so it stops at its own step 6.  The restarted
    run matches the oracle's RK run from the checkpointed state (same
    initial step, PID history and time) as closely as a free adaptive run
    does, and the series continue in append mode."""
    cfg = _small_cfg(25, 25, 4, True)
    Nz, N, fts = cfg.Nz, cfg.Ny * cfg.Nz, cfg.fault_type_scale
    d = str(tmp_path) + "/"
    g = _mediator(cfg, reuse)
    g.set_output(d, stride1D=2, stride2D=0, checkpoint=1, interval=6)
    assert g.load_checkpoint() is None
    tEnd, n, dtsA = g.integrate(15, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    assert n == 6
    stA = g.get_state()
    assert _ascii(d, "ckptNumber") == 2 and _ascii(d, "chkpt_stepCount") == 6
    assert abs(_ascii(d, "chkpt_currT") / tEnd - 1) <= 1e-15
    for k, v in (("psi", stA["psi"]), ("slip", stA["slip"]), ("slipVel", stA["V"]), ("tau", stA["tau"])):
        assert np.array_equal(read_vecs(d + "chkpt_" + k)[0], v), k
    # bcL = slip / faultTypeScale of the d_dt at that state
    assert np.allclose(read_vecs(d + "chkpt_momBal_bcL")[0] * fts, stA["slip"], rtol=1e-15, atol=0)
    # chkpt_deltaT is the step the run would try next: an uninterrupted run
    # (no output) tries it for step 7 and, unless it rejects it, accepts it
    # (to rounding: the output steps' flag reads shift the deferral policy)
    gd = _mediator(cfg, reuse)
    _, _, dtsD = gd.integrate(7, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    assert np.allclose(dtsD[:6], dtsA, rtol=1e-10, atol=0)
    if gd.counters()["rejected_steps"] == g.counters()["rejected_steps"]:
        assert abs(dtsD[6] / _ascii(d, "chkpt_deltaT") - 1) <= 1e-10
    gd.close()
    g.close()
    # the state the restart loads (the restart's own checkpoints overwrite it)
    ck = {k: read_vecs(d + "chkpt_" + k)[0].copy() for k in ("psi", "momBal_bcL", "slipVel")}
    ck_t = {k: _ascii(d, "chkpt_" + k) for k in ("currT", "deltaT", "currErr", "prevErr")}

    # restart
    g2 = _mediator(cfg, reuse)
    g2.set_state(np.zeros(Nz), np.zeros(Nz), np.zeros(Nz))  # overwritten by the checkpoint
    g2.set_output(d, stride1D=2, stride2D=0, checkpoint=1, interval=6)
    info = g2.load_checkpoint()
    assert info is not None and info["ckpt_number"] == 2
    assert info["init_time"] == ck_t["currT"] and info["init_dt"] == ck_t["deltaT"]
    assert info["curr_err"] == ck_t["currErr"] and info["prev_err"] == ck_t["prevErr"]
    s0 = g2.get_state()
    assert np.array_equal(s0["psi"], ck["psi"]) and np.array_equal(s0["slip"], ck["momBal_bcL"] * fts)
    assert np.array_equal(s0["V"], ck["slipVel"])
    tEnd2, n2, dts2 = g2.integrate(15, tol=1e-8, init_dt=info["init_dt"], max_dt=1e10,
                                   init_time=info["init_time"], curr_err=info["curr_err"],
                                   prev_err=info["prev_err"])
    assert n2 == 6
    st2, u2 = g2.get_state(), g2.get_u()
    assert _ascii(d, "ckptNumber") == 4 and _ascii(d, "chkpt_stepCount") == 6
    assert np.array_equal(read_vecs(d + "chkpt_psi")[0], st2["psi"])
    # series: steps 0, 2, 4, 6 of the first run, then 0 (the checkpoint state
    # again, its time through the "%.15e" file), 2, 4, 6 of the restart
    t = read_lines(d + "med_time1D.txt")
    assert len(t) == 8 and len(read_vecs(d + "slip")) == 8
    assert abs(t[4] / t[3] - 1) <= 1e-15 and abs(t[-1] / tEnd2 - 1) <= 1e-15
    g2.close()

    # oracle RK run from the checkpointed state, with the loaded step, PID
    # history and time, six steps with the step count starting at 0
    h, keep, _ = _oracle_qd(cfg)
    r = _Run(0, 1, 0, 6, info["init_time"], 1e15, info["init_dt"], 0.0, 1e10, 1e-8, 0, None,
             info["curr_err"], info["prev_err"])
    ref = dict(psi=ck["psi"].copy(), slip=ck["momBal_bcL"] * fts, V=ck["slipVel"].copy(), tau=np.zeros(Nz),
               u=np.zeros(N))
    tE = C.c_double()
    odts = np.zeros(6)
    ns = olib().orc_qd_integrate(h, C.byref(r), optr(ref["psi"]), optr(ref["slip"]), optr(ref["V"]),
                                 optr(ref["tau"]), optr(ref["u"]), C.byref(tE), optr(odts), 6, None)
    olib().orc_qd_free(h)
    assert ns == 6
    dt_err = np.abs(dts2 / odts - 1).max()
    e = {k: np.abs(st2[k] - ref[k]).max() / np.abs(ref[k]).max() for k in ("slip", "psi", "V", "tau")}
    e["u"] = np.abs(u2 - ref["u"]).max() / np.abs(ref["u"]).max()
    print("reuse", reuse, "restart vs oracle: dt", dt_err, e, flush=True)
    assert dt_err <= 1e-6 and abs(tEnd2 / tE.value - 1) <= 1e-6
    assert max(e.values()) <= 1e-7, e



def test_powerlaw_output_series(tmp_path):
    """The power-law mediator's series (StrikeSlip_PowerLaw_qd::timeMonitor,
    strikeSlip_powerLaw_qd.cpp:352-408): every step below step 50, then every
    100th, whatever the strides; the 2-D files of PowerLaw::writeStep2D
    (powerLaw.cpp:2312-2372).  The last record is checked against the
    mediator's own state and viscous fields (the d_dt kernels, not the
    output kernel), the oracle's Dy/Dz of the written u, and the creep law
    (powerLaw.cpp:438-474) on the written stresses."""
    from scycle_amd import PowerLawMediator
    from test_gpu_powerlaw import _case
    c, gxy, gxz, slip = _case(21, 23, 4, 9)
    Ny, Nz, N = c.Ny, c.Nz, c.Ny * c.Nz
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi")
    m.set_state(c.psi0, slip, np.zeros(Nz))
    m.set_visc(gxy, gxz)
    d = str(tmp_path) + "/"
    m.set_output(d, stride1D=7, stride2D=3)
    steps = 49
    tEnd, n, dts = m.integrate(steps, tol=1e-8, init_dt=1e5, init_time=3e7)
    assert n == steps
    times = read_lines(d + "med_time1D.txt")
    want_t = 3e7 + np.concatenate([[0.0], np.cumsum(dts[:n])])
    assert len(times) == steps + 1 and len(read_lines(d + "med_time2D.txt")) == steps + 1
    assert np.allclose(times, want_t, rtol=1e-14)
    st, v, u = m.get_state(), m.get_visc(), m.get_u()
    last = {k: read_vecs(d + k)[-1] for k in ("slip", "psi", "slipVel", "tauQSP", "momBal_u", "momBal_gxy",
                                              "momBal_gxz", "momBal_dgVxy", "momBal_dgVxz", "momBal_sxy",
                                              "momBal_sxz", "momBal_sdev", "momBal_gTxy", "momBal_gTxz",
                                              "momBal_effVisc", "momBal_T", "disl_invEffVisc",
                                              "disl_dgVdev")}
    assert len(read_vecs(d + "momBal_effVisc")) == steps + 1
    for k, ref in (("slip", st["slip"]), ("psi", st["psi"]), ("slipVel", st["V"]), ("momBal_u", u),
                   ("momBal_gxy", v["gVxy"]), ("momBal_gxz", v["gVxz"]), ("momBal_dgVxy", v["dgVxy"]),
                   ("momBal_dgVxz", v["dgVxz"]), ("momBal_sxy", v["sxy"]), ("momBal_effVisc", v["effVisc"]),
                   ("momBal_T", c.T)):
        err = np.abs(last[k] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < 1e-13, (k, err)
    o = OracleSbp(c.order, Ny, Nz, c.Ly, c.Lz, c.mu, c.y, c.z, var_grid=c.var_grid)
    for k, ref in (("momBal_gTxy", o.dy(last["momBal_u"])), ("momBal_gTxz", o.dz(last["momBal_u"]))):
        assert np.abs(last[k] - ref).max() <= 1e-12 * np.abs(ref).max(), k
    sxy = c.mu * (last["momBal_gTxy"] - last["momBal_gxy"])
    sxz = c.mu * (last["momBal_gTxz"] - last["momBal_gxz"])
    sdev = np.sqrt(sxy ** 2 + sxz ** 2)
    inv = 1e3 * c.A * sdev ** (c.n - 1.0) * np.exp(-c.B / c.T)
    for k, ref in (("momBal_sxz", sxz), ("momBal_sdev", sdev), ("disl_invEffVisc", inv),
                   ("disl_dgVdev", sdev * inv)):
        assert np.abs(last[k] - ref).max() <= 1e-12 * np.abs(ref).max(), k
    m.close()
    # from step 50 on the schedule is every 100th step: a 60-step run writes
    # steps 0..49 only, with the same times
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi")
    m.set_state(c.psi0, slip, np.zeros(Nz))
    m.set_visc(gxy, gxz)
    d2 = str(tmp_path) + "/long_"
    m.set_output(d2, stride1D=7, stride2D=0)
    _, n2, dts2 = m.integrate(60, tol=1e-8, init_dt=1e5, init_time=3e7)
    assert n2 == 60
    times2 = read_lines(d2 + "med_time1D.txt")
    assert len(times2) == 50 and np.allclose(times2, times, rtol=1e-14)
    assert not os.path.exists(d2 + "med_time2D.txt")
    m.close()


def test_fd_output_series(tmp_path):
    """The FD mediator's integrate loop and series (OdeSolver_WaveEq::integrate,
    strikeSlip_linearElastic_fd::timeMonitor): records at step 0 and every
    stride, time t0 + k dt, the last record equal to the driver's state
    after the same number of plain steps, sxy the oracle's muxDy of the
    written u, tauQSP = tauP, surfDisp the initial surface."""
    from scycle_amd import FdMediator
    from scycle_amd.configs import fd_rupture
    cfg = fd_rupture(33, 29, order=4, Ly=24.0, Lz=24.0)
    Ny, Nz, N = cfg.Ny, cfg.Nz, cfg.Ny * cfg.Nz
    rng = np.random.default_rng(3)
    u = 1e-3 * rng.standard_normal(N)
    u_prev = u - cfg.dt * 1e-2 * rng.random(N)
    psi = cfg.psi0 + 0.01 * rng.random(Nz)
    state = (u_prev, u, psi.copy(), psi, 2.0 * u_prev[:Nz], 2.0 * u[:Nz])
    for bulk in ("fp64", "fp32"):
        d = str(tmp_path) + f"/{bulk}_"
        m = FdMediator(cfg, bulk=bulk)
        m.set_state(*state)
        m.set_output(d, stride1D=5, stride2D=10)
        steps = 23
        tEnd, n = m.integrate(steps, cfg.dt, init_time=0.5)
        assert n == steps and abs(tEnd - (0.5 + steps * cfg.dt)) < 1e-12
        got = m.get_state()
        m.close()
        ref_m = FdMediator(cfg, bulk=bulk)
        ref_m.set_state(*state)
        ref_m.steps(steps, cfg.dt)
        ref = ref_m.get_state()
        ref_m.close()
        times = read_lines(d + "med_time1D.txt")
        assert len(times) == 1 + steps // 5 and times[0] == 0.5
        assert np.allclose(times, 0.5 + 5 * cfg.dt * np.arange(len(times)), rtol=1e-13)
        assert len(read_lines(d + "med_time2D.txt")) == 1 + steps // 10
        # step 23 is not written: compare the record of step 20 with 20 plain steps
        m2 = FdMediator(cfg, bulk=bulk)
        m2.set_state(*state)
        m2.steps(20, cfg.dt)
        s20 = m2.get_state()
        m2.close()
        for k, ref_k in (("slip", "slip"), ("psi", "psi"), ("slipVel", "V"), ("tauP", "tau"),
                         ("tauQSP", "tau"), ("strength", "strength")):
            rec = read_vecs(d + k)
            assert len(rec) == 1 + steps // 5
            assert np.abs(rec[-1] - s20[ref_k]).max() <= 1e-14 * np.abs(s20[ref_k]).max(), k
        assert np.abs(got["slip"] - ref["slip"]).max() == 0.0  # batched = step by step
        surf = read_vecs(d + "surfDisp")
        assert all(np.array_equal(s, u[::Nz]) for s in surf)
        assert all(np.all(r == 0) for r in read_vecs(d + "momBal_bcR"))
        u2 = read_vecs(d + "momBal_u")[-1]
        assert np.abs(u2 - s20["u"]).max() <= 1e-15 * np.abs(s20["u"]).max()
        o = OracleSbp(cfg.order, Ny, Nz, cfg.Ly, cfg.Lz, cfg.mu, cfg.y, cfg.z, var_grid=cfg.var_grid,
                      bc=(1, 1, 1, 1))
        sxy = read_vecs(d + "momBal_sxy")[-1]
        ref_sxy = o.muxdy(u2)
        assert np.abs(sxy - ref_sxy).max() <= 1e-12 * np.abs(ref_sxy).max()
