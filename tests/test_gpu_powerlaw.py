"""GPU power-law viscoelastic QD step (dislocation creep) against the oracle.

Tolerances: the pointwise strain-rate update in FP64 agrees to 1e-13
relative (same formula, pow/exp from two math libraries), with FP32 arrays
to 1e-6; one d_dt (GPU CG to 1e-12 vs the oracle's direct solve) to 1e-9 in
the fault rates and the viscous strain rates; RK32 steps with the oracle's
step sizes to 1e-8 in slip and the viscous strains.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _rates_inputs(N, seed):
    rng = np.random.default_rng(seed)
    return dict(sxy=20 * rng.standard_normal(N), sxz=10 * rng.standard_normal(N),
                T=300 + 700 * rng.random(N), A=np.full(N, 1e-3), B=np.full(N, 15635.2),
                n=3 + rng.random(N), mu=30 + 6 * rng.random(N), sat=rng.standard_normal(N))


@pytest.mark.parametrize("prec", [0, 1])
def test_powerlaw_rates_pointwise(prec):
    from scycle_amd import DeviceArray, SbpOps, load
    from scycle_amd._lib import PowerLawArgs, check
    N = 4099
    d = _rates_inputs(N, 5)
    ev, gxy, gxz = np.zeros(N), np.zeros(N), np.zeros(N)
    arrs = [np.ascontiguousarray(d[k]) for k in ("sxy", "sxz", "T", "A", "B", "n", "mu")]
    ol.lib().orc_powerlaw(N, *[ol.ptr(x) for x in arrs], ol.ptr(d["sat"]), 1e30, ol.ptr(ev),
                          ol.ptr(gxy), ol.ptr(gxz))
    dt = np.float64 if prec == 0 else np.float32
    dev = {k: DeviceArray(N, dt, d[k]) for k in d}
    out = {k: DeviceArray(N, dt) for k in ("ev", "gxy", "gxz")}
    ops = SbpOps(2, 9, 9, 1.0, 1.0, np.ones(81), var_grid=False)
    a = PowerLawArgs(N, 1e30, dev["sxy"].ptr, dev["sxz"].ptr, dev["T"].ptr, dev["A"].ptr,
                     dev["B"].ptr, dev["n"].ptr, dev["mu"].ptr, dev["sat"].ptr, out["ev"].ptr,
                     out["gxy"].ptr, out["gxz"].ptr)
    check(load().sc_powerlaw_rates(ops.h, prec, C.byref(a)))
    ops.sync()
    if prec == 1:  # the FP32 run sees the inputs rounded to FP32
        r = {k: d[k].astype(np.float32).astype(np.float64) for k in d}
        arrs = [np.ascontiguousarray(r[k]) for k in ("sxy", "sxz", "T", "A", "B", "n", "mu")]
        ol.lib().orc_powerlaw(N, *[ol.ptr(x) for x in arrs], ol.ptr(r["sat"]), 1e30, ol.ptr(ev),
                              ol.ptr(gxy), ol.ptr(gxz))
    tol = 1e-13 if prec == 0 else 1e-6
    for name, g, ref in (("effVisc", out["ev"], ev), ("gxy", out["gxy"], gxy), ("gxz", out["gxz"], gxz)):
        err = np.max(np.abs(g.to_host().astype(np.float64) - ref) / np.abs(ref))
        print(f"powerlaw prec={prec} {name} max rel err {err:.2e}", flush=True)
        assert err < tol, name


def _case(Ny, Nz, order, seed):
    from scycle_amd.configs import power_law
    c = power_law(Ny, Nz, order=order, Ly=20.0, Lz=24.0)
    rng = np.random.default_rng(seed)
    N = Ny * Nz
    c.mu = 30 + 6 * rng.random(N)
    # a hotter geotherm so that creep matters on this small domain
    c.T = 600.0 + 400.0 * np.tile(np.arange(Nz) / (Nz - 1.0), Ny)
    gxy = 1e-5 * rng.standard_normal(N)
    gxz = 1e-5 * rng.standard_normal(N)
    slip = 1e-2 * rng.random(Nz)
    return c, gxy, gxz, slip


@pytest.mark.parametrize("order,Ny,Nz", [(4, 21, 23), (2, 17, 19), (4, 41, 37)])
def test_powerlaw_ddt_parity(order, Ny, Nz):
    from scycle_amd import PowerLawMediator
    c, gxy, gxz, slip = _case(Ny, Nz, order, Ny + order)
    t = 3e7
    o = ol.PowerLawOracle(c, jac_mode=1)
    ref = o.d_dt(t, c.psi0, slip, gxy, gxz)
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi")
    m.set_state(c.psi0, slip, np.zeros(Nz))
    m.set_visc(gxy, gxz)
    dpsi, dslip = m.d_dt(t)
    v = m.get_visc()
    for name, g, r in (("dpsi", dpsi, ref["dpsi"]), ("dslip", dslip, ref["dslip"]),
                       ("dgxy", v["dgVxy"], ref["dgxy"]), ("dgxz", v["dgVxz"], ref["dgxz"]),
                       ("effVisc", v["effVisc"], ref["effVisc"]), ("sxy", v["sxy"], ref["sxy"])):
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"powerlaw ddt order={order} {Ny}x{Nz} {name} rel err {err:.2e}", flush=True)
        assert err < 1e-9, name
    m.close()


def test_powerlaw_source_parity():
    """rhs += B gxy + C gxz: through d_dt with slip = 0 at t = 0 the boundary
    data vanish, so the displacement solves A u = B gxy + C gxz."""
    from scycle_amd import PowerLawMediator
    c, gxy, gxz, _ = _case(29, 31, 4, 3)
    o = ol.PowerLawOracle(c, jac_mode=1)
    ref = o.d_dt(0.0, c.psi0, np.zeros(c.Nz), gxy, gxz)
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi")
    m.set_state(c.psi0, np.zeros(c.Nz), np.zeros(c.Nz))
    m.set_visc(gxy, gxz)
    m.d_dt(0.0)
    u = m.get_u()
    err = np.abs(u - ref["u"]).max() / np.abs(ref["u"]).max()
    print("powerlaw source: u rel err", err, flush=True)
    assert err < 1e-9
    m.close()


@pytest.mark.parametrize("err_inds", [("psi", "slip"), ("psi", "slip", "gVxy", "gVxz")])
def test_powerlaw_integrate_matches_oracle(err_inds):
    """RK32 from the same state: replaying the oracle's step sizes the fields
    agree to 1e-8; the free-running controller picks the same step sizes."""
    from scycle_amd import PowerLawMediator
    c, gxy, gxz, slip = _case(21, 23, 4, 9)
    mask = sum({"psi": 1, "slip": 2, "gVxy": 4, "gVxz": 8}[k] for k in err_inds)
    o = ol.PowerLawOracle(c, jac_mode=1)
    steps = 12
    ref = o.integrate(steps, c.psi0, slip, gxy, gxz, tol=1e-8, init_dt=1e5, init_time=3e7,
                      max_dt=1e10, err_mask=mask)
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi", err_inds=err_inds)
    m.set_state(c.psi0, slip, np.zeros(c.Nz))
    m.set_visc(gxy, gxz)
    tEnd, n, dts = m.integrate(steps, tol=1e-8, init_dt=1e5, init_time=3e7, replay_dt=ref["dts"])
    st, v = m.get_state(), m.get_visc()
    for name, g, r in (("psi", st["psi"], ref["psi"]), ("slip", st["slip"], ref["slip"]),
                       ("gVxy", v["gVxy"], ref["gVxy"]), ("gVxz", v["gVxz"], ref["gVxz"])):
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"powerlaw RK32 replay {err_inds} {name} rel err {err:.2e}", flush=True)
        assert err < 1e-8, name
    m.close()
    m = PowerLawMediator(c, solve="cg64", rtol=1e-13, precond="jacobi", err_inds=err_inds)
    m.set_state(c.psi0, slip, np.zeros(c.Nz))
    m.set_visc(gxy, gxz)
    tEnd, n, dts = m.integrate(steps, tol=1e-8, init_dt=1e5, init_time=3e7)
    assert n == steps
    assert np.allclose(dts, ref["dts"], rtol=1e-6)
    m.close()


def test_powerlaw_c3_full_size_runs():
    """801^2 C3: a few RK32 steps with the FD preconditioner converge, stay
    finite and keep eta within (0, cap]."""
    from scycle_amd import PowerLawMediator
    from scycle_amd.configs import config_c3
    c = config_c3()
    N = c.Ny * c.Nz
    m = PowerLawMediator(c, solve="cg64", rtol=1e-12, precond="fdiag64", reuse=4)
    slip = np.linspace(1e-2, 0.0, c.Nz)
    m.set_state(c.psi0, slip, np.zeros(c.Nz))
    m.set_visc(np.zeros(N), np.zeros(N))
    tEnd, n, dts = m.integrate(5, tol=1e-8, init_dt=1e3, init_time=3e7)
    st, v = m.get_state(), m.get_visc()
    print("C3 5 steps: dt", dts, "max gVxy", np.abs(v["gVxy"]).max(), flush=True)
    assert n == 5 and np.all(np.isfinite(st["slip"])) and np.all(np.isfinite(v["gVxy"]))
    assert np.all(v["effVisc"] > 0) and np.all(v["effVisc"] <= 1e30)
    m.close()
