"""GPU fully dynamic step (wave update of propagateWaves + Fault_fd::d_dt)
against the oracle, plus size-independent properties at full size.

Tolerances: the FP64 wave update agrees with the oracle to 1e-13 of the
magnitude of its terms (dt^2/rho |Jinv Hinv| |A| |u| + 2|u| + |c2 uPrev|,
node by node: the two sides sum the same products in a different order); the
FP32 bulk update to 2e-6 of the same magnitude.  Fault quantities come from
root solves with |F| < rootTol = 1e-12 on both sides, so V, psi and slip
agree to 1e-9 relative.  Several coupled steps in FP64 agree to 1e-9; the
FP32-bulk run is compared to the FP64 oracle at 1e-4 in slip (single
precision body, FP64 fault), the stated accuracy of that mode.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import FdFaultOracle, OracleSbp, fd_ay, fd_propagate

pytestmark = pytest.mark.gpu


def _grid(Ny, Nz, Ly, Lz, by=0.0, bz=0.0):
    q = np.repeat(np.arange(Ny) / (Ny - 1.0), Nz)
    r = np.tile(np.arange(Nz) / (Nz - 1.0), Ny)
    y = Ly * (np.sinh(by * q) / np.sinh(by) if by else q)
    z = Lz * (np.sinh(bz * r) / np.sinh(bz) if bz else r)
    return y, z


def _magnitude(o, dt, u, u_prev, rho, ay, var_grid):
    """Per-node magnitude of the terms of the update (for the error scale)."""
    import scipy.sparse as sp  # noqa: F401  (csr from the oracle)
    A = o.csr()
    absAu = abs(A) @ np.abs(u)
    d2 = o.Hinv(absAu)
    if var_grid:
        yq, zr = o.metric()
        d2 = d2 / np.abs(yq * zr)
    c1 = dt * dt / rho
    return c1 * d2 + 2 * np.abs(u) + np.abs(dt * ay - 1.0) * np.abs(u_prev), c1 * d2


CASES = [(o, v, Ny, Nz) for o in (2, 4) for v in (False, True)
         for (Ny, Nz) in ((21, 25), (41, 37), (17, 33), (9, 12))]


@pytest.mark.parametrize("order,var_grid,Ny,Nz", CASES)
def test_wave_step_parity(order, var_grid, Ny, Nz):
    from scycle_amd import DeviceArray, SbpOps, load
    from scycle_amd._lib import check
    rng = np.random.default_rng(Ny * 5 + Nz + order)
    Ly, Lz = 30.0, 24.0
    N = Ny * Nz
    mu = 30 + 6 * rng.random(N)
    rho = 2.5 + rng.random(N)
    cs = np.sqrt(mu / rho)
    y, z = _grid(Ny, Nz, Ly, Lz, 1.2, 3.0) if var_grid else (None, None)
    bc = (1, 1, 1, 1)
    o = OracleSbp(order, Ny, Nz, Ly, Lz, mu, y, z, var_grid=var_grid, bc=bc)
    g = SbpOps(order, Ny, Nz, Ly, Lz, mu, y, z, var_grid=var_grid, bc=bc)
    ay = fd_ay(o, cs, 1, 0, 1, 1)
    dt = 1e-3
    u, up = rng.standard_normal(N), rng.standard_normal(N)
    un_ref, d2u_ref = fd_propagate(o, dt, u, up, rho, ay)
    mag, d2mag = _magnitude(o, dt, u, up, rho, ay, var_grid)
    for prec, dtype, tol in ((0, np.float64, 1e-13), (1, np.float32, 2e-6)):
        du, dup, drho, day = (DeviceArray(N, dtype, x) for x in (u, up, rho, ay))
        dun, dd2 = DeviceArray(N, dtype), DeviceArray(Nz, dtype)
        check(load().sc_wave_step(g.h, prec, dt, du.ptr, dup.ptr, dun.ptr, drho.ptr, day.ptr, dd2.ptr))
        g.sync()
        err = np.max(np.abs(dun.to_host() - un_ref) / mag)
        errf = np.max(np.abs(dd2.to_host() - d2u_ref[:Nz]) * (dt * dt / rho[:Nz]) / d2mag[:Nz].clip(1e-300))
        print(f"wave order={order} var={var_grid} {Ny}x{Nz} prec={prec} err={err:.2e} fault={errf:.2e}",
              flush=True)
        assert err < tol and errf < tol


def _fault_case(Nz, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 24.0, Nz)
    a = np.where((z > 8) & (z < 14), 0.015, 0.025)
    b = np.full(Nz, 0.02)
    Dc = np.full(Nz, 8e-3)
    sN = np.full(Nz, 50.0)
    rho = np.full(Nz, 3.0)
    alphay = np.full(Nz, 0.0083)
    tau0 = 30 + 10 * np.exp(-(z - 11) ** 2 / 2)
    slip0 = np.zeros(Nz)
    psi = 0.6 + 0.15 * rng.random(Nz)
    psi_prev = psi + 1e-4 * rng.standard_normal(Nz)
    slip = 1e-3 * rng.random(Nz)
    slip_prev = slip - 1e-4 * rng.random(Nz)
    d2u = 1e-1 * rng.standard_normal(Nz)
    V = 10.0 ** rng.uniform(-12, 0, Nz)
    return dict(a=a, b=b, Dc=Dc, sNEff=sN, rho=rho, alphay=alphay, tau0=tau0, slip0=slip0,
                psi=psi, psi_prev=psi_prev, slip=slip, slip_prev=slip_prev, d2u=d2u, V=V)


@pytest.mark.parametrize("state_law", [0, 1])
def test_fault_fd_step_parity(state_law):
    from scycle_amd import DeviceArray, SbpOps, load
    from scycle_amd._lib import FaultFdArgs, check
    Nz, dt = 513, 2e-3
    c = _fault_case(Nz, 7 + state_law)
    orc = FdFaultOracle(Nz, 1e-6, 0.6, 1e-9, 1e-12, 2.0, state_law, c["a"], c["b"], c["Dc"],
                        c["sNEff"], c["rho"], c["alphay"], c["tau0"], c["slip0"])
    Vr = c["V"].copy()
    psi_r, slip_r, uf_r, Vr, bad = orc.step(dt, c["d2u"], Vr, c["psi"], c["psi_prev"], c["slip"],
                                            c["slip_prev"])
    assert bad == 0
    ops = SbpOps(2, 9, 9, 1.0, 1.0, np.ones(81), var_grid=False)
    d = {k: DeviceArray(Nz, host=c[k]) for k in ("a", "b", "Dc", "sNEff", "rho", "alphay", "tau0",
                                                 "slip0", "d2u", "psi", "psi_prev", "slip",
                                                 "slip_prev", "V")}
    out = {k: DeviceArray(Nz) for k in ("psiN", "slipN", "strength", "uF")}
    st = DeviceArray(2, np.int32)
    args = FaultFdArgs(Nz, dt, 1e-6, 1e-9, 0.6, 1e-12, 2.0, 10000, state_law,
                       d["a"].ptr, d["b"].ptr, d["Dc"].ptr, d["sNEff"].ptr, d["rho"].ptr,
                       d["alphay"].ptr, d["tau0"].ptr, d["slip0"].ptr, None, d["d2u"].ptr,
                       d["psi"].ptr, d["psi_prev"].ptr, d["slip"].ptr, d["slip_prev"].ptr,
                       d["V"].ptr, out["psiN"].ptr, out["slipN"].ptr, out["strength"].ptr,
                       out["uF"].ptr, st.ptr)
    check(load().sc_fault_fd_step(ops.h, C.byref(args)))
    ops.sync()
    assert int(st.to_host()[0]) == 0
    Vg = d["V"].to_host()
    for name, g, r in (("V", Vg, Vr), ("psi", out["psiN"].to_host(), psi_r),
                       ("slip", out["slipN"].to_host(), slip_r), ("uF", out["uF"].to_host(), uf_r)):
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"fault fd law={state_law} {name} rel err {err:.2e}", flush=True)
        assert err < 1e-9, name


def _oracle_fd_run(cfg, steps, u_prev, u, psi_prev, psi, slip_prev, slip):
    """The reference's d_dt sequence on the oracle: propagateWaves, Fault_fd
    step, fault displacement written into the fault column, rotate."""
    Ny, Nz = cfg.Ny, cfg.Nz
    o = OracleSbp(cfg.order, Ny, Nz, cfg.Ly, cfg.Lz, cfg.mu, cfg.y, cfg.z, var_grid=cfg.var_grid,
                  bc=(1, 1, 1, 1))
    ay = fd_ay(o, cfg.cs, cfg.outR, cfg.outT, cfg.outB, cfg.faultL)
    h11y = o.scalars()["h11y"]
    yq, _ = o.metric()
    alphay = np.full(Nz, h11y) * (yq[:Nz] if cfg.var_grid else 1.0)
    f = FdFaultOracle(Nz, cfg.v0, cfg.f0, cfg.vL, cfg.root_tol, cfg.fault_type_scale, cfg.state_law,
                      cfg.a, cfg.b, cfg.Dc, cfg.sNEff, cfg.rho[:Nz], alphay, cfg.tau0, cfg.slip0)
    V = np.zeros(Nz)
    for _ in range(steps):
        un, d2u = fd_propagate(o, cfg.dt, u, u_prev, cfg.rho, ay)
        psi_n, slip_n, uf, V, bad = f.step(cfg.dt, d2u[:Nz], V, psi, psi_prev, slip, slip_prev)
        assert bad == 0
        un[:Nz] = uf
        u_prev, u = u, un
        psi_prev, psi = psi, psi_n
        slip_prev, slip = slip, slip_n
    return dict(u=u, psi=psi, slip=slip, V=V)


@pytest.mark.parametrize("order,var_grid,Ny,Nz", [(4, False, 41, 37), (2, True, 33, 29)])
def test_fd_driver_matches_oracle(order, var_grid, Ny, Nz):
    from scycle_amd import FdMediator
    from scycle_amd.configs import fd_rupture
    cfg = fd_rupture(Ny, Nz, order=order, Ly=24.0, Lz=24.0)
    if var_grid:
        y, z = _grid(Ny, Nz, 24.0, 24.0, 0.0, 2.0)
        cfg.var_grid, cfg.y, cfg.z = True, y, z
    rng = np.random.default_rng(3)
    N = Ny * Nz
    # a rupture already under way: nonzero slip rate on the patch
    u = 1e-3 * rng.standard_normal(N)
    u_prev = u - cfg.dt * 1e-2 * rng.random(N)
    psi = cfg.psi0 + 0.01 * rng.random(Nz)
    psi_prev = psi.copy()
    slip = 2.0 * u[:Nz]
    slip_prev = 2.0 * u_prev[:Nz]
    steps = 20
    ref = _oracle_fd_run(cfg, steps, u_prev, u, psi_prev, psi, slip_prev, slip)
    for bulk, tol_slip, tol_u in (("fp64", 1e-9, 1e-9), ("fp32", 1e-4, 1e-4)):
        m = FdMediator(cfg, bulk=bulk)
        m.set_state(u_prev, u, psi_prev, psi, slip_prev, slip)
        m.steps(steps, cfg.dt)
        got = m.get_state()
        errs = {k: np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300)
                for k in ("u", "psi", "slip", "V")}
        print(f"fd driver {bulk} order={order} var={var_grid}: {errs}", flush=True)
        assert errs["slip"] < tol_slip and errs["u"] < tol_u
        assert errs["psi"] < tol_slip and errs["V"] < 10 * tol_slip
        m.close()


def test_wave_energy_conserved_full_size():
    """1001^2, 4th order, all-Neumann body, no absorbing terms: the discrete
    energy  E = (du)^T diag(rho J H) (du)/dt^2 - u_{n+1}^T A u_n  of the
    leapfrog update is conserved to round-off (A symmetric), which pins the
    dt^2/rho, Hinv and Jinv factors and the sign of the update."""
    from scycle_amd import DeviceArray, SbpOps, load
    from scycle_amd._lib import check
    from scycle_amd.configs import _grid as cgrid
    n = 1001
    y, z = cgrid(n, n, 30.0, 30.0, 5.0)
    N = n * n
    mu = np.full(N, 36.0)
    rho = np.full(N, 3.0)
    g = SbpOps(4, n, n, 30.0, 30.0, mu, y, z, var_grid=True, bc=(1, 1, 1, 1))
    yy, zz = y.reshape(n, n), z.reshape(n, n)
    u0 = (np.exp(-((yy - 15) ** 2 + (zz - 10) ** 2) / 4.0)).ravel()
    dmin = min(z[1] - z[0], 30.0 / (n - 1))
    dt = 0.1 * dmin / np.sqrt(12.0)
    bufs = [DeviceArray(N, host=u0), DeviceArray(N, host=u0), DeviceArray(N)]
    drho, day, dd2 = DeviceArray(N, host=rho), DeviceArray(N), DeviceArray(n)
    # weights rho * J * H (H = Hy (x) Hz, J = yq zr) from the library's own H and metric
    Hw = g.H(np.ones(N))
    yq, zr = g.metric()
    W = rho * Hw * np.repeat(yq, n) * np.tile(zr, n)

    def energy(a, b):
        return float(((b - a) ** 2 * W).sum() / dt ** 2 - b @ g.apply(a))

    es = []
    prev, cur, nxt = 0, 1, 2
    for it in range(40):
        check(load().sc_wave_step(g.h, 0, dt, bufs[cur].ptr, bufs[prev].ptr, bufs[nxt].ptr,
                                  drho.ptr, day.ptr, dd2.ptr))
        if it % 10 == 9:
            g.sync()
            es.append(energy(bufs[cur].to_host(), bufs[nxt].to_host()))
        prev, cur, nxt = cur, nxt, prev
    es = np.array(es)
    print("energies", es, flush=True)
    assert np.all(np.isfinite(es)) and es[0] > 0
    assert np.abs(es - es[0]).max() <= 1e-9 * abs(es[0])


def test_fd_full_size_c4_runs():
    """2001^2 C4 rupture: 20 steps in FP64 and FP32 bulk stay finite, the
    nucleation patch slips and the two precisions agree on slip to 1e-4."""
    from scycle_amd import FdMediator
    from scycle_amd.configs import config_c4
    cfg = config_c4()
    N, Nz = cfg.Ny * cfg.Nz, cfg.Nz
    res = {}
    for bulk in ("fp64", "fp32"):
        m = FdMediator(cfg, bulk=bulk)
        z0 = np.zeros(N)
        psi = np.full(Nz, 1.0) * cfg.psi0
        m.set_state(z0, z0, psi, psi, np.zeros(Nz), np.zeros(Nz))
        m.steps(20, cfg.dt)
        res[bulk] = m.get_state()
        m.close()
    s64, s32 = res["fp64"]["slip"], res["fp32"]["slip"]
    assert np.all(np.isfinite(s64)) and np.all(np.isfinite(res["fp64"]["u"]))
    assert s64.max() > 0
    err = np.abs(s32 - s64).max() / np.abs(s64).max()
    print("C4 20 steps: max slip", s64.max(), "fp32 vs fp64", err, flush=True)
    assert err < 1e-4
