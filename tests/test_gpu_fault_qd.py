"""GPU rate-and-state kernel and quasi-dynamic step against the oracle.

Tolerances: slip velocity from both root solves satisfies |F(V)| < rootTol =
1e-12 MPa, so V agrees to ~1e-11 relative; state rates and stresses follow.
One d_dt agrees to 1e-9 relative (the linear solve is CG to 1e-12 on the GPU
and a direct Cholesky solve in the oracle).  Integrated fault slip after the
RK steps agrees to 1e-8 relative, the project's parity bar for slip.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import lib as olib, ptr as optr, vel_qd

pytestmark = pytest.mark.gpu


def _fault_inputs(n, seed, creeping=True):
    rng = np.random.default_rng(seed)
    a = 0.010 + 0.015 * rng.random(n)
    b = np.full(n, 0.02)
    sN = 50.0 * np.ones(n)
    eta = np.full(n, 0.5 * np.sqrt(36 * 3.0))
    Dc = np.full(n, 8e-3)
    psi = 0.6 + 0.2 * rng.random(n)
    # tauQS spanning creep (V ~ 1e-20..1e-9) to seismic slip (V ~ 1 m/s)
    V = 10.0 ** rng.uniform(-20, 1, n)
    tau = a * sN * np.arcsinh(V / 2e-6 * np.exp(psi / a)) + eta * V
    return dict(a=a, b=b, sN=sN, eta=eta, Dc=Dc, psi=psi, tauQS=tau, Vtrue=V)


def _gpu_fault(inp, jac_mode, state_law=0, V0=None):
    from scycle_amd import DeviceArray, load
    from scycle_amd._lib import FaultQdArgs, check
    n = len(inp["a"])
    d = {k: DeviceArray(n, host=inp[k]) for k in ("a", "b", "sN", "eta", "Dc", "psi", "tauQS")}
    V = DeviceArray(n, host=np.zeros(n) if V0 is None else V0)
    out = {k: DeviceArray(n) for k in ("dpsi", "tau", "strength")}
    st = DeviceArray(2, np.int32)
    # device-side context for the stream
    from scycle_amd import SbpOps
    ops = SbpOps(2, 9, 9, 1.0, 1.0, np.ones(81), var_grid=False)
    args = FaultQdArgs(n, 1e-6, 1e-9, 0.6, 1e-12, 10000, state_law, jac_mode,
                       d["eta"].ptr, d["sN"].ptr, d["a"].ptr, d["b"].ptr, d["Dc"].ptr, None, None,
                       d["psi"].ptr, d["tauQS"].ptr, V.ptr, out["dpsi"].ptr, out["tau"].ptr,
                       out["strength"].ptr, st.ptr)
    check(load().sc_fault_qd_rates(ops.h, C.byref(args)))
    ops.sync()
    return dict(V=V.to_host(), dpsi=out["dpsi"].to_host(), tau=out["tau"].to_host(),
                strength=out["strength"].to_host(), status=int(st.to_host()[0]),
                tauQS=d["tauQS"].to_host())


@pytest.mark.parametrize("jac_mode", [0, 1])
@pytest.mark.parametrize("state_law", [0, 1])
def test_fault_kernel_parity(jac_mode, state_law):
    inp = _fault_inputs(1001, 42)
    g = _gpu_fault(inp, jac_mode, state_law)
    assert g["status"] == 0
    Vo, bad, _ = vel_qd(inp["eta"], inp["tauQS"], inp["sN"], inp["psi"], inp["a"], 1e-6,
                        jac_mode=jac_mode)
    assert bad == 0
    rel = np.abs(g["V"] - Vo) / np.maximum(np.abs(Vo), 1e-300)
    assert rel.max() < 1e-10
    # both satisfy the residual tolerance of the reference root finder
    F = inp["a"] * inp["sN"] * np.arcsinh(g["V"] / 2e-6 * np.exp(inp["psi"] / inp["a"])) \
        - (inp["tauQS"] - inp["eta"] * g["V"])
    assert np.abs(F).max() < 1e-11
    # the strength comes from the root solve's own asinh (log-based, no library
    # call): against numpy's asinh over z from ~1e-4 to ~1e30 (both of its
    # branches), to the rounding of z and of the exponential
    z = g["V"] / 2e-6 * np.exp(inp["psi"] / inp["a"])
    print("z range", z.min(), z.max(), "below 1/2:", int((z < 0.5).sum()), flush=True)
    sref = inp["a"] * inp["sN"] * np.arcsinh(z)
    err = np.abs(g["strength"] - sref) / np.abs(sref)
    print("strength rel err max", err.max(), flush=True)
    assert err.max() < 1e-14
    L = olib()
    dref = np.array([L.orc_aging(p, v, b, 0.6, 1e-6, dc) if state_law == 0 else
                     L.orc_sliplaw(p, v, a, b, 0.6, 1e-6, dc)
                     for p, v, a, b, dc in zip(inp["psi"], g["V"], inp["a"], inp["b"], inp["Dc"])])
    # Both laws are differences of nearly equal terms near steady state, so the
    # bound is relative to the magnitude of the terms (libm ulps differ between
    # the device and host exp/log/asinh): aging (b v0/Dc)(|e^((f0-psi)/b)| + V/v0),
    # slip |V|/Dc (|f| + |fss|).
    V, a, b, Dc, psi = g["V"], inp["a"], inp["b"], inp["Dc"], inp["psi"]
    if state_law == 0:
        scale = (b * 1e-6 / Dc) * (np.exp((0.6 - psi) / b) + V / 1e-6)
    else:
        f = a * np.arcsinh(V / 2e-6 * np.exp(psi / a))
        fss = 0.6 + (a - b) * np.log(np.maximum(V, 1e-300) / 1e-6)
        scale = V / Dc * (np.abs(f) + np.abs(fss))
    assert np.all(np.abs(g["dpsi"] - dref) <= 1e-13 * scale + 1e-300)
    assert np.allclose(g["tau"], inp["tauQS"] - inp["eta"] * g["V"], rtol=1e-12)


def _small_cfg(Ny, Nz, order, var_grid):
    from scycle_amd.configs import strike_slip
    return strike_slip(Ny, Nz, order=order, var_grid=var_grid,
                       zstretch=3.0 if var_grid else None, layered_mu=var_grid)


def _oracle_qd(cfg, solver=0):
    class Cfg(C.Structure):
        _fields_ = [("order", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("var_grid", C.c_int),
                    ("Ly", C.c_double), ("Lz", C.c_double), ("bcR", C.c_int), ("bcT", C.c_int),
                    ("bcL", C.c_int), ("bcB", C.c_int), ("remoteLoadingR", C.c_int),
                    ("faultTypeScale", C.c_double), ("vL", C.c_double), ("f0", C.c_double),
                    ("v0", C.c_double), ("rootTol", C.c_double), ("stateLaw", C.c_int),
                    ("jac_mode", C.c_int), ("solver", C.c_int), ("cg_rtol", C.c_double)]
    c = Cfg(cfg.order, cfg.Ny, cfg.Nz, int(cfg.var_grid), cfg.Ly, cfg.Lz, *cfg.mat_bc, 1,
            cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0, cfg.root_tol, cfg.state_law, 0,
            solver, 1e-12)
    keep = [np.ascontiguousarray(x) for x in (cfg.mu, cfg.y, cfg.z, cfg.a, cfg.b, cfg.Dc,
                                              cfg.sN, cfg.eta)]
    h = olib().orc_qd_new(C.byref(c), *[optr(x) for x in keep], None, None, None)
    assert h
    return h, keep, c


@pytest.mark.parametrize("order,var_grid,N", [(2, False, 21), (4, True, 31)])
def test_qd_ddt_parity(order, var_grid, N):
    from scycle_amd import QdMediator
    cfg = _small_cfg(N, N, order, var_grid)
    g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0)
    h, keep, _ = _oracle_qd(cfg)
    L = olib()
    slip = 1e-3 * np.linspace(1, 0, N)
    for t in (0.0, 3.7e7):
        g.set_state(cfg.psi0, slip, np.full(N, 1e-9))
        dpsi, dslip = g.d_dt(t)
        Vo = np.full(N, 1e-9)
        dp_o, ds_o, tau_o = np.zeros(N), np.zeros(N), np.zeros(N)
        u_o = np.zeros(N * N)
        assert L.orc_qd_ddt(h, t, optr(cfg.psi0), optr(slip), optr(Vo), optr(dp_o), optr(ds_o),
                            optr(tau_o), optr(u_o)) == 0
        st = g.get_state()
        assert np.abs(dslip - ds_o).max() <= 1e-9 * np.abs(ds_o).max()
        assert np.abs(dpsi - dp_o).max() <= 1e-9 * np.abs(dp_o).max()
        assert np.abs(st["tau"] - tau_o).max() <= 1e-9 * np.abs(tau_o).max()
        u = g.get_u()
        assert np.abs(u - u_o).max() <= 1e-9 * np.abs(u_o).max()
    L.orc_qd_free(h)


class _Run(C.Structure):
    _fields_ = [("integrator", C.c_int), ("control", C.c_int), ("normType", C.c_int),
                ("maxSteps", C.c_int), ("initTime", C.c_double), ("maxTime", C.c_double),
                ("initDt", C.c_double), ("minDt", C.c_double), ("maxDt", C.c_double),
                ("tol", C.c_double), ("nReplay", C.c_int), ("replayDt", C.c_void_p),
                ("currErr", C.c_double), ("prevErr", C.c_double)]


def _fields_err(gs, ug, ref):
    out = {}
    for name in ("slip", "psi", "V", "tau"):
        out[name] = np.abs(gs[name] - ref[name]).max() / np.abs(ref[name]).max()
    out["u"] = np.abs(ug - ref["u"]).max() / np.abs(ref["u"]).max()
    return out


@pytest.mark.parametrize("integrator", ["RK32", "RK43"])
@pytest.mark.parametrize("solve", ["cg64", "ir32"])
def test_qd_integrate_parity(integrator, solve):
    """40 adaptive RK steps from the same state against the oracle (direct
    Cholesky solves).  Replay mode forces the oracle's step sizes on the GPU
    run: fault slip, state, slip velocity, shear stress and off-fault
    displacement then agree to 1e-9 relative.  In a free adaptive run the PID
    controller maps rounding-level differences of the error estimate
    (itself a difference of two nearly equal RK solutions) into the step
    sizes; they agree to 1e-6 and the fields to 1e-7."""
    from scycle_amd import QdMediator
    N = 25
    cfg = _small_cfg(N, N, 4, True)
    steps = 40
    h, keep, _ = _oracle_qd(cfg)
    r = _Run(0 if integrator == "RK32" else 1, 1, 0, steps, 0.0, 1e15, 1e-3, 0.0, 1e10, 1e-8, 0,
             None)
    ref = dict(psi=cfg.psi0.copy(), slip=cfg.slip0.copy(), V=np.zeros(N), tau=np.zeros(N),
               u=np.zeros(N * N))
    tE = C.c_double()
    odts = np.zeros(steps)
    ns = olib().orc_qd_integrate(h, C.byref(r), optr(ref["psi"]), optr(ref["slip"]),
                                 optr(ref["V"]), optr(ref["tau"]), optr(ref["u"]), C.byref(tE),
                                 optr(odts), steps, None)
    olib().orc_qd_free(h)
    assert ns == steps
    # replay: identical step sizes
    g = QdMediator(cfg, solve=solve, rtol=1e-13, jac_mode=0)
    g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    tEnd, nst, dts = g.integrate(steps, integrator=integrator, tol=1e-8, init_dt=1e-3,
                                 max_dt=1e10, replay_dt=odts)
    assert nst == steps and abs(tEnd - tE.value) <= 1e-14 * tE.value
    e = _fields_err(g.get_state(), g.get_u(), ref)
    print(integrator, solve, "replay", e, flush=True)
    assert max(e.values()) <= 1e-9, e
    # free adaptive run
    g2 = QdMediator(cfg, solve=solve, rtol=1e-13, jac_mode=0)
    g2.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    tEnd2, nst2, dts2 = g2.integrate(steps, integrator=integrator, tol=1e-8, init_dt=1e-3,
                                     max_dt=1e10)
    assert nst2 == steps
    print(integrator, solve, "free dt err", np.abs(dts2 / odts - 1).max(), flush=True)
    assert np.allclose(dts2, odts, rtol=1e-6) and abs(tEnd2 / tE.value - 1) <= 1e-6
    e2 = _fields_err(g2.get_state(), g2.get_u(), ref)
    print(integrator, solve, "free", e2, flush=True)
    assert max(e2.values()) <= 1e-7, e2


@pytest.mark.parametrize("reuse,defer", [(1, "always"), (4, "always"), (4, "auto"), (4, "never")])
@pytest.mark.parametrize("integrator", ["RK32", "RK43"])
def test_qd_integrate_projected_guess(integrator, reuse, defer):
    """The projected initial guess inside sc_qd_integrate: its residual check
    runs on the device and is read back with each attempt's error estimate;
    an attempt in which any guess missed the tolerance is repeated with
    checked solves.  defer "always" defers every solve with a basis (and so
    meets misses and the repeat), "auto" only after a run of hits, "never"
    checks every solve.  In replay mode the fields match the oracle's direct
    solves to 1e-9 in every case."""
    from scycle_amd import QdMediator
    N = 25
    cfg = _small_cfg(N, N, 4, True)
    steps = 40
    h, keep, _ = _oracle_qd(cfg)
    r = _Run(0 if integrator == "RK32" else 1, 1, 0, steps, 0.0, 1e15, 1e-3, 0.0, 1e10, 1e-8, 0,
             None)
    ref = dict(psi=cfg.psi0.copy(), slip=cfg.slip0.copy(), V=np.zeros(N), tau=np.zeros(N),
               u=np.zeros(N * N))
    tE = C.c_double()
    odts = np.zeros(steps)
    ns = olib().orc_qd_integrate(h, C.byref(r), optr(ref["psi"]), optr(ref["slip"]),
                                 optr(ref["V"]), optr(ref["tau"]), optr(ref["u"]), C.byref(tE),
                                 optr(odts), steps, None)
    olib().orc_qd_free(h)
    assert ns == steps
    g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0, precond="fdiag64", reuse=reuse)
    g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    tEnd, nst, dts = g.integrate(steps, integrator=integrator, tol=1e-8, init_dt=1e-3,
                                 max_dt=1e10, replay_dt=odts, defer=defer)
    c = g.counters()
    assert nst == steps and abs(tEnd - tE.value) <= 1e-14 * tE.value
    e = _fields_err(g.get_state(), g.get_u(), ref)
    print(integrator, "reuse", reuse, defer, "replay", e, "deferred", c["deferred_checks"], "missed",
          c["missed_guesses"], "solves", c["solves"], "iters", c["cg_iters"], flush=True)
    assert max(e.values()) <= 1e-9, e
    if defer == "never":
        assert c["deferred_checks"] == 0
    else:
        assert c["deferred_checks"] > 0
    if defer == "always":
        assert c["missed_guesses"] > 0
    g.close()


@pytest.mark.parametrize("defer", ["auto", "always"])
def test_qd_device_step_control(defer):
    """RK32 with the step control on the device (the host enqueues attempts
    in batches and reads the control block once per batch) takes the steps
    the host-controlled loop takes: the controller is the same arithmetic,
    pow aside, so step sizes agree to rounding and the fields to 1e-10.
    With defer "always" the first guesses miss, the device loop halts and
    the host redoes those steps with checked solves, then resumes.  (Step
    sizes to 1e-9 and fields to 1e-9: the loops evaluate the guess's
    residual test by different sums, so a guess on the threshold may be
    accepted by one and solved by the other.)  Both runs also match the
    oracle's free adaptive run as test_qd_integrate_parity does (step sizes
    1e-6, fields 1e-7)."""
    from scycle_amd import QdMediator
    N = 25
    cfg = _small_cfg(N, N, 4, True)
    steps = 80
    out = {}
    for mode in ("host", "auto"):
        g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0, precond="fdiag64", reuse=2)
        g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
        tEnd, nst, dts = g.integrate(steps, tol=1e-8, init_dt=1e-3, max_dt=1e10, defer=defer, step_control=mode)
        out[mode] = dict(t=tEnd, n=nst, dts=dts, st=g.get_state(), u=g.get_u(), c=g.counters())
        g.close()
    h, d = out["host"], out["auto"]
    print(defer, "device steps", d["c"]["device_steps"], "missed", d["c"]["missed_guesses"], "host missed",
          h["c"]["missed_guesses"], flush=True)
    assert h["c"]["device_steps"] == 0 and d["c"]["device_steps"] > steps // 2
    assert h["n"] == d["n"] == steps
    assert np.allclose(d["dts"], h["dts"], rtol=1e-9, atol=0)
    assert abs(d["t"] / h["t"] - 1) <= 1e-9
    for k in ("psi", "slip", "V", "tau"):
        ref = h["st"][k]
        assert np.abs(d["st"][k] - ref).max() <= 1e-9 * np.abs(ref).max(), k
    assert np.abs(d["u"] - h["u"]).max() <= 1e-9 * np.abs(h["u"]).max()
    # the oracle's free run
    hq, keep, _ = _oracle_qd(cfg)
    r = _Run(0, 1, 0, steps, 0.0, 1e15, 1e-3, 0.0, 1e10, 1e-8, 0, None)
    ref = dict(psi=cfg.psi0.copy(), slip=cfg.slip0.copy(), V=np.zeros(N), tau=np.zeros(N), u=np.zeros(N * N))
    tE = C.c_double()
    odts = np.zeros(steps)
    ns = olib().orc_qd_integrate(hq, C.byref(r), optr(ref["psi"]), optr(ref["slip"]), optr(ref["V"]),
                                 optr(ref["tau"]), optr(ref["u"]), C.byref(tE), optr(odts), steps, None)
    olib().orc_qd_free(hq)
    assert ns == steps
    assert np.allclose(d["dts"], odts, rtol=1e-6) and abs(d["t"] / tE.value - 1) <= 1e-6
    e = _fields_err(d["st"], d["u"], ref)
    assert max(e.values()) <= 1e-7, e


@pytest.mark.parametrize("case", ["l2_relative", "p_control", "final_time", "forced_min_dt", "rejections"])
def test_qd_device_step_control_branches(case):
    """The device loop's step control on the branches the default run does
    not take, against the host-controlled loop from the same state: the
    L2_relative norm (the error sums of hi^2 as well), P control
    throughout, a run that ends on maxTime (the last step clamped to it),
    and, continuing a run with a tighter tolerance, rejected attempts and
    acceptances forced at minDt (a tolerance no step down to minDt meets).
    Step sizes agree to 1e-9, the end time to 1e-9 (exactly on maxTime) and
    the fields to 1e-9: the controller is the same arithmetic, but the two
    loops evaluate the guess's residual test differently (the device loop
    on the band rows plus the off-band Gram form, the host loop over the
    grid), so a guess whose residual sits on rtol ||b|| to rounding can be
    accepted by one and solved by the other, both then meeting rtol."""
    from scycle_amd import QdMediator
    N = 25
    cfg = _small_cfg(N, N, 4, True)
    steps = 60
    kw = dict(tol=1e-8, init_dt=1e-3, max_dt=1e10, defer="auto")
    slip0 = cfg.slip0
    if case == "l2_relative":  # a relative norm needs a slip away from zero
        kw["norm"] = "L2_relative"
        slip0 = 1e-3 * np.linspace(1.0, 0.5, N)
    elif case == "p_control":
        kw["control"] = "P"
    elif case == "final_time":
        g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0, precond="fdiag64", reuse=2)
        g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
        t60, _, _ = g.integrate(steps, step_control="host", **kw)
        g.close()
        kw["max_time"] = 0.37 * t60
    warm = None
    if case in ("forced_min_dt", "rejections"):
        # both runs continue from the same state after 30 host-controlled
        # steps (the basis then predicts the solutions, so the attempts run
        # on the device): with a tolerance the current step size cannot meet
        # (rejected attempts), or one no step down to min_dt meets (forced
        # acceptances at min_dt)
        warm = dict(tol=1e-8, init_dt=1e-3, max_dt=1e10, defer="auto", step_control="host")
    out = {}
    for mode in ("host", "auto"):
        g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0, precond="fdiag64", reuse=2)
        g.set_state(cfg.psi0, slip0, np.zeros(N))
        if warm:
            t1, _, d1 = g.integrate(30, **warm)
            kw = dict(max_dt=1e10, defer="always", init_time=t1, init_dt=float(d1[-1]))
            if case == "forced_min_dt":
                kw.update(tol=1e-14, min_dt=0.5 * float(d1[-1]))
            else:
                kw.update(tol=1e-10)
        tEnd, nst, dts = g.integrate(steps, step_control=mode, **kw)
        out[mode] = dict(t=tEnd, n=nst, dts=dts, st=g.get_state(), u=g.get_u(), c=g.counters())
        g.close()
    h, d = out["host"], out["auto"]
    print(case, "steps", d["n"], "device steps", d["c"]["device_steps"], "rejected", d["c"]["rejected_steps"],
          h["c"]["rejected_steps"], "t", d["t"], h["t"], flush=True)
    assert h["c"]["device_steps"] == 0 and d["c"]["device_steps"] > 0
    assert h["n"] == d["n"] and d["c"]["rejected_steps"] == h["c"]["rejected_steps"]
    assert np.allclose(d["dts"], h["dts"], rtol=1e-9, atol=0)
    assert abs(d["t"] / h["t"] - 1) <= 1e-9
    for k in ("psi", "slip", "V", "tau"):
        ref = h["st"][k]
        assert np.abs(d["st"][k] - ref).max() <= 1e-9 * np.abs(ref).max(), k
    assert np.abs(d["u"] - h["u"]).max() <= 1e-9 * np.abs(h["u"]).max()
    if case == "final_time":
        assert d["n"] < steps and d["t"] == kw["max_time"] and h["t"] == kw["max_time"]
    elif case == "forced_min_dt":
        # every accepted step after the first rejections is at the minimum:
        # the tolerance is never met
        assert d["dts"][-1] == kw["min_dt"] and d["c"]["rejected_steps"] >= d["n"] // 2
    elif case == "rejections":
        assert d["c"]["rejected_steps"] > 0 and d["dts"][0] < kw["init_dt"]
    else:
        assert d["n"] == steps


@pytest.mark.parametrize("N,order,var_grid", [(25, 4, True), (33, 4, True), (29, 2, False)])
def test_qd_stage_kernel_rates(N, order, var_grid):
    """The device loop's stage kernel (k_qd_stage: stage state, projections
    from the per-basis plan, fault rates, projected guess) in its
    accepted-state form against a host-checked d_dt at the same state and
    time: slip rate, state rate (through a second evaluation) and the
    displacement formed from the guess.  The guess meets the solve tolerance
    (1e-13) at a state the basis was built on, so the rates agree to the
    root solve's and the solve's tolerance."""
    from scycle_amd import QdMediator
    from scycle_amd._lib import check, load
    cfg = _small_cfg(N, N, order, var_grid)
    g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=1, precond="fdiag64", reuse=2)
    g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    tEnd, nst, _ = g.integrate(40, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    assert nst == 40
    ms, words = C.c_double(), C.c_double()
    check(load().sc_qd_bench_stage(g.h, tEnd, 3, C.byref(ms), C.byref(words)))
    M = 2
    assert words.value == (2 * M + 19) * N + (M + 2) * 2 * 9 * N + 2 * M * M
    st = g.get_state()
    u_stage = g.get_u()
    dpsi, dslip = g.d_dt(tEnd)  # host-checked solve at the same state
    st2 = g.get_state()
    u_ref = g.get_u()
    print("V", np.abs(st["V"] - dslip).max() / np.abs(dslip).max(),
          "u", np.abs(u_stage - u_ref).max() / np.abs(u_ref).max(), flush=True)
    assert np.abs(st["V"] - dslip).max() <= 1e-9 * np.abs(dslip).max()
    assert np.abs(st["tau"] - st2["tau"]).max() <= 1e-9 * np.abs(st2["tau"]).max()
    assert np.abs(u_stage - u_ref).max() <= 1e-9 * np.abs(u_ref).max()
    g.close()


@pytest.mark.parametrize("N,order,var_grid", [(101, 4, True), (61, 2, False)])
def test_stage_guess_residual_gram_form(N, order, var_grid):
    """The stage kernel's residual test forms ||b - A x||^2 of the projected
    guess on the band rows node by node and off them (b = 0) as c^T G c with
    the off-band Gram matrix G = <A y, A y> in double-double (no pass over
    the grid).  Against the same two sums by an explicit pass (x formed, b
    from the boundary data, A x applied): ||b||^2 to 1e-12, ||r|| to 1e-3 of
    itself where it stands above 1e-14 ||b|| (the explicit pass's rounding
    lies below that), both under 2e-14 ||b|| otherwise, and the KSP test
    ||r|| < rtol ||b|| decided alike for rtol 1e-12 and 1e-8.  States: the
    accepted state the basis was built on, times ahead of it, and slips
    perturbed by 1e-9 .. 1e-3 (guesses from good to useless)."""
    from scycle_amd import QdMediator
    from scycle_amd._lib import check, load
    cfg = _small_cfg(N, N, order, var_grid)
    g = QdMediator(cfg, solve="cg64", rtol=1e-12, jac_mode=1, precond="fdiag64", reuse=2)
    g.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    tEnd, nst, dts = g.integrate(40, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    st = g.get_state()
    rng = np.random.default_rng(7)
    out = np.zeros(4)
    cases = [(ahead, 0.0) for ahead in (0.0, 1.0, 10.0, 1e3)] + [(0.0, e) for e in (1e-9, 1e-6, 1e-3)]
    seen_above = 0
    for ahead, eps in cases:
        if eps:
            g.set_state(st["psi"], st["slip"] * (1.0 + eps * rng.standard_normal(N)), st["V"])
        check(load().sc_qd_guess_residual(g.h, tEnd + ahead * float(dts[nst - 1]),
                                          out.ctypes.data_as(C.POINTER(C.c_double))))
        rg, bg, re, be = out
        print(f"ahead {ahead} eps {eps}: gram r/b {np.sqrt(rg / bg):.4e} explicit {np.sqrt(re / be):.4e} "
              f"b.b {bg:.6e} {be:.6e}", flush=True)
        assert bg > 0 and abs(bg - be) <= 1e-12 * be
        if np.sqrt(re) > 1e-14 * np.sqrt(be):
            seen_above += 1
            assert abs(np.sqrt(rg) - np.sqrt(re)) <= 1e-3 * np.sqrt(re)
        else:
            assert np.sqrt(rg) <= 2e-14 * np.sqrt(bg)
        for rtol in (1e-12, 1e-8):
            assert (rg <= rtol ** 2 * bg) == (re <= rtol ** 2 * be)
    assert seen_above >= 5
    g.close()


def test_qd_device_step_control_wide_fault():
    """A fault longer than the stage kernel's tables (Nz = 2049 > 2048): the
    device loop runs the rhs, guess and step-end kernels instead of one
    kernel per stage, and takes the host loop's steps.  Checks are deferred
    from the first guess on (the default waits for a run of 8 hits, which 30
    steps of this start do not reach), and the solves are held to 1e-10 (a
    basis of 3 solutions does not reach 1e-13 at this shape): the two loops'
    step sizes and fields then agree to 1e-8 (measured 2e-9 for the step
    sizes)."""
    from scycle_amd import QdMediator
    Ny, Nz = 17, 2049
    cfg = _small_cfg(Ny, Nz, 4, True)
    out = {}
    for mode in ("host", "auto"):
        g = QdMediator(cfg, solve="cg64", rtol=1e-10, jac_mode=0, precond="fdiag64", reuse=3)
        g.set_state(cfg.psi0, cfg.slip0, np.zeros(Nz))
        tEnd, nst, dts = g.integrate(30, tol=1e-8, init_dt=1e-3, max_dt=1e10, defer="always",
                                     step_control=mode)
        out[mode] = dict(t=tEnd, n=nst, dts=dts, st=g.get_state(), u=g.get_u(), c=g.counters())
        g.close()
    h, d = out["host"], out["auto"]
    print("device steps", d["c"]["device_steps"], "missed", d["c"]["missed_guesses"], "host missed",
          h["c"]["missed_guesses"], "deferred", h["c"]["deferred_checks"], flush=True)
    assert h["c"]["device_steps"] == 0 and d["c"]["device_steps"] > 10
    assert h["n"] == d["n"] == 30
    assert np.allclose(d["dts"], h["dts"], rtol=1e-8, atol=0)
    for k in ("psi", "slip", "V", "tau"):
        ref = h["st"][k]
        assert np.abs(d["st"][k] - ref).max() <= 1e-8 * np.abs(ref).max(), k
    assert np.abs(d["u"] - h["u"]).max() <= 1e-8 * np.abs(h["u"]).max()
