"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product library never
imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(os.path.dirname(_HERE), "oracle", "liboracle.so")

_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_SO)
        L.orc_sbp_new.restype = C.c_void_p
        L.orc_sbp_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _dp, _dp, _dp]
        L.orc_sbp_free.argtypes = [C.c_void_p]
        L.orc_sbp_change_bc.argtypes = [C.c_void_p] + [C.c_int] * 4
        L.orc_sbp_nnz.restype = C.c_long
        L.orc_sbp_nnz.argtypes = [C.c_void_p, C.c_int]
        L.orc_sbp_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_sbp_csr.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_int), _dp]
        for name in ("orc_sbp_apply", "orc_sbp_muxdy", "orc_sbp_muxdz", "orc_sbp_dy",
                     "orc_sbp_dz", "orc_sbp_h", "orc_sbp_hinv", "orc_sbp_dyxmu", "orc_sbp_dzxmu"):
            getattr(L, name).argtypes = [C.c_void_p, _dp, _dp]
        L.orc_sbp_bnd.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_sbp_set_rhs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_sbp_diag.argtypes = [C.c_void_p, _dp]
        L.orc_sbp_metric.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_sbp_scalars.argtypes = [C.c_void_p, _dp]
        L.orc_sbp1d.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int] + [_dp] * 10
        L.orc_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_chol_solve.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_strength.restype = C.c_double
        L.orc_strength.argtypes = [C.c_double] * 5
        L.orc_aging.restype = C.c_double
        L.orc_aging.argtypes = [C.c_double] * 6
        L.orc_sliplaw.restype = C.c_double
        L.orc_sliplaw.argtypes = [C.c_double] * 7
        L.orc_vel_qd.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, _dp, C.POINTER(C.c_long)]
        L.orc_qd_new.restype = C.c_void_p
        L.orc_qd_new.argtypes = [C.c_void_p] + [_dp] * 11
        L.orc_qd_free.argtypes = [C.c_void_p]
        L.orc_qd_ddt.argtypes = [C.c_void_p, C.c_double] + [_dp] * 7
        L.orc_qd_integrate.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                       C.c_int, _dp]
        L.orc_qd_counter.restype = C.c_long
        L.orc_qd_counter.argtypes = [C.c_void_p, C.c_int]
        L.orc_fd_propagate.argtypes = [C.c_void_p, C.c_double] + [_dp] * 6
        L.orc_fd_ay.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_fdfault_new.restype = C.c_void_p
        L.orc_fdfault_new.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int] + [_dp] * 8
        L.orc_fdfault_free.argtypes = [C.c_void_p]
        L.orc_fdfault_step.argtypes = [C.c_void_p, C.c_double] + [_dp] * 9
        L.orc_powerlaw.argtypes = [C.c_int] + [_dp] * 8 + [C.c_double] + [_dp] * 3
        L.orc_pl_new.restype = C.c_void_p
        L.orc_pl_new.argtypes = [C.c_void_p] + [_dp] * 15 + [C.c_double]
        L.orc_pl_free.argtypes = [C.c_void_p]
        L.orc_pl_source.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_pl_sat.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_pl_ddt.argtypes = [C.c_void_p, C.c_double] + [_dp] * 12
        L.orc_pl_integrate.argtypes = [C.c_void_p, C.c_void_p, C.c_int] + [_dp] * 7 + [C.c_int]
        L.orc_qdfd_new.restype = C.c_void_p
        L.orc_qdfd_new.argtypes = [C.c_void_p, C.c_void_p] + [_dp] * 5
        L.orc_qdfd_free.argtypes = [C.c_void_p]
        L.orc_qdfd_dt.restype = C.c_double
        L.orc_qdfd_dt.argtypes = [C.c_void_p]
        L.orc_qdfd_ay.argtypes = [C.c_void_p, _dp]
        L.orc_qdfd_integrate.argtypes = [C.c_void_p, C.c_void_p] + [_dp] * 5 + \
            [C.POINTER(C.c_int)] * 3 + [C.c_void_p, C.c_int]
        L.orc_rk_linear.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_double, C.c_double, _dp, _dp,
                                    C.POINTER(C.c_long)]
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


DIRICHLET, NEUMANN = 0, 1


class OracleSbp:
    """Assembled SBP-SAT operator (SbpOps_m_varGrid / SbpOps_m_constGrid)."""

    def __init__(self, order, Ny, Nz, Ly, Lz, mu, y=None, z=None, var_grid=True,
                 bc=(DIRICHLET, NEUMANN, DIRICHLET, NEUMANN), compat=0, laplace=0):
        self.Ny, self.Nz, self.N = Ny, Nz, Ny * Nz
        self._mu = np.ascontiguousarray(mu, dtype=np.float64)
        self._y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self._z = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
        self.h = lib().orc_sbp_new(order, Ny, Nz, Ly, Lz, int(var_grid), compat, laplace,
                                   bc[0], bc[1], bc[2], bc[3], ptr(self._mu), ptr(self._y),
                                   ptr(self._z))
        if not self.h:
            raise RuntimeError("orc_sbp_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sbp_free(self.h)
            self.h = None

    def change_bc(self, bcR, bcT, bcL, bcB):
        lib().orc_sbp_change_bc(self.h, bcR, bcT, bcL, bcB)

    def csr(self, which=0):
        import scipy.sparse as sp
        nr, nc = C.c_int(), C.c_int()
        lib().orc_sbp_shape(self.h, which, C.byref(nr), C.byref(nc))
        nnz = lib().orc_sbp_nnz(self.h, which)
        rp = np.zeros(nr.value + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz, dtype=np.float64)
        lib().orc_sbp_csr(self.h, which, rp.ctypes.data_as(C.POINTER(C.c_long)),
                          ci.ctypes.data_as(C.POINTER(C.c_int)), ptr(v))
        return sp.csr_matrix((v, ci, rp), shape=(nr.value, nc.value))

    def _unary(self, fn, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.N)
        fn(self.h, ptr(x), ptr(out))
        return out

    def apply(self, u):
        return self._unary(lib().orc_sbp_apply, u)

    def muxdy(self, u):
        return self._unary(lib().orc_sbp_muxdy, u)

    def muxdz(self, u):
        return self._unary(lib().orc_sbp_muxdz, u)

    def dy(self, u):
        return self._unary(lib().orc_sbp_dy, u)

    def dz(self, u):
        return self._unary(lib().orc_sbp_dz, u)

    def dyxmu(self, u):
        return self._unary(lib().orc_sbp_dyxmu, u)

    def dzxmu(self, u):
        return self._unary(lib().orc_sbp_dzxmu, u)

    def bnd(self, which, x):
        """Hy^-1 e0y, Hy^-1 eNy (x: Nz face values), Hy^-1 E0y, Hy^-1 ENy,
        Hz^-1 E0z, Hz^-1 ENz (x: a grid field), for which = 0..5."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.N)
        lib().orc_sbp_bnd(self.h, int(which), ptr(x), ptr(out))
        return out

    def H(self, u):
        return self._unary(lib().orc_sbp_h, u)

    def Hinv(self, u):
        return self._unary(lib().orc_sbp_hinv, u)

    def set_rhs(self, bcL, bcR, bcT, bcB):
        out = np.zeros(self.N)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (bcL, bcR, bcT, bcB)]
        lib().orc_sbp_set_rhs(self.h, ptr(out), *[ptr(x) for x in a])
        return out

    def diag(self):
        out = np.zeros(self.N)
        lib().orc_sbp_diag(self.h, ptr(out))
        return out

    def metric(self):
        yq, zr = np.zeros(self.N), np.zeros(self.N)
        lib().orc_sbp_metric(self.h, ptr(yq), ptr(zr))
        return yq, zr

    def scalars(self):
        out = np.zeros(7)
        lib().orc_sbp_scalars(self.h, ptr(out))
        return dict(zip(("dy", "dz", "alphaDy", "alphaDz", "h11y", "h11z", "alphaT"), out))

    def pcg(self, b, u0=None, rtol=1e-12, maxit=100000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        u = np.zeros(self.N) if u0 is None else np.array(u0, dtype=np.float64)
        rr = C.c_double()
        it = lib().orc_pcg(self.h, ptr(b), ptr(u), rtol, maxit, C.byref(rr))
        return u, it, rr.value

    def chol_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        u = np.zeros(self.N)
        if lib().orc_chol_solve(self.h, ptr(b), ptr(u)):
            raise RuntimeError("banded Cholesky failed")
        return u


def sbp1d(order, N, scale, compat=0):
    mats = {k: np.zeros((N, N)) for k in ("H", "Hinv", "D1", "BS", "D2", "C2", "D3", "D4", "C3", "C4")}
    e = lib().orc_sbp1d(order, N, scale, compat, *[ptr(mats[k]) for k in
                        ("H", "Hinv", "D1", "BS", "D2", "C2", "D3", "D4", "C3", "C4")])
    if e:
        raise RuntimeError(f"orc_sbp1d failed ({e})")
    return mats


def vel_qd(eta, tauQS, sN, psi, a, v0, vL=1e-9, tol=1e-12, maxits=10000, jac_mode=0,
           V0=None, locked=None):
    n = len(eta)
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (eta, tauQS, sN, psi, a)]
    lk = None if locked is None else np.ascontiguousarray(locked, dtype=np.float64)
    V = np.zeros(n) if V0 is None else np.array(V0, dtype=np.float64)
    its = C.c_long(0)
    bad = lib().orc_vel_qd(n, *[ptr(x) for x in arrs], ptr(lk), v0, vL, tol, maxits, jac_mode,
                           ptr(V), C.byref(its))
    return V, bad, its.value


def fd_propagate(o, dt, u, u_prev, rho, ay):
    """propagateWaves on the oracle operator `o`: (uNext, D2u)."""
    N = len(u)
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, u_prev, rho, ay)]
    u_next, d2u = np.zeros(N), np.zeros(N)
    lib().orc_fd_propagate(o.h, dt, *[ptr(x) for x in arrs], ptr(u_next), ptr(d2u))
    return u_next, d2u


def fd_ay(o, cs, outR, outT, outB, faultL):
    ay = np.zeros(len(cs))
    lib().orc_fd_ay(o.h, ptr(np.ascontiguousarray(cs, dtype=np.float64)), int(outR), int(outT),
                    int(outB), int(faultL), ptr(ay))
    return ay


class FdFaultOracle:
    """Fault_fd::d_dt of the oracle (one step at a time)."""

    def __init__(self, N, v0, f0, vL, rootTol, fts, stateLaw, a, b, Dc, sNEff, rho, alphay,
                 tau0, slip0):
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                      (a, b, Dc, sNEff, rho, alphay, tau0, slip0)]
        self.N = N
        self.h = lib().orc_fdfault_new(N, v0, f0, vL, rootTol, fts, stateLaw,
                                       *[ptr(x) for x in self._keep])

    def __del__(self):
        try:
            lib().orc_fdfault_free(self.h)
        except Exception:
            pass

    def step(self, dt, d2u, V, psi, psi_prev, slip, slip_prev):
        """Returns (psiNext, slipNext, uFault, V, bad); V is updated in place."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (d2u, psi, psi_prev, slip, slip_prev)]
        psi_n, slip_n, uf = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        bad = lib().orc_fdfault_step(self.h, dt, ptr(arrs[0]), ptr(V), ptr(arrs[1]), ptr(arrs[2]),
                                     ptr(arrs[3]), ptr(arrs[4]), ptr(psi_n), ptr(slip_n), ptr(uf))
        return psi_n, slip_n, uf, V, bad


def qd_solve_ss(h, Nz, N, psi, have_state_vals=False):
    """solveSS on an oracle QD model (orc_qd_new): psi (in: used when
    have_state_vals), slip, V, bcRShift and the shifted u."""
    f = lib().orc_qd_solve_ss
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp, _dp]
    out = dict(psi=np.array(psi, dtype=np.float64), slip=np.zeros(Nz), V=np.zeros(Nz),
               bcRShift=np.zeros(Nz), u=np.zeros(N))
    rc = f(h, ptr(out["psi"]), int(bool(have_state_vals)), ptr(out["slip"]), ptr(out["V"]),
           ptr(out["bcRShift"]), ptr(out["u"]))
    if rc:
        raise RuntimeError(f"orc_qd_solve_ss failed ({rc})")
    return out


def rk_linear_trace(order, lam, dt0, final_t, tol, y0=1.0, control=1, cap=100000):
    """Adaptive run of y' = lam*y by the oracle's RK solver; per accepted step
    (size, error estimate, next step size chosen)."""
    L = lib()
    f = L.orc_rk_linear_trace
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                  _dp, _dp, _dp, C.c_int]
    dts, errs, nd = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    n = f(order, lam, dt0, final_t, tol, y0, control, ptr(dts), ptr(errs), ptr(nd), cap)
    if n < 0:
        raise RuntimeError("orc_rk_linear_trace failed")
    n = min(n, cap)
    return dts[:n], errs[:n], nd[:n]


def rk_linear(order, lam, dt, n, lam2=None, final_t=0.0, tol=1e-8, y0=1.0):
    """y' = lam*y integrated by the oracle's RK32 (order 3) / RK43 (order 4):
    n fixed steps of dt, or adaptive to final_t when n < 0.
    Returns (y[2], t, steps, d_dt calls)."""
    y, t, calls = np.zeros(2), C.c_double(), C.c_long()
    steps = lib().orc_rk_linear(order, lam, lam if lam2 is None else lam2, dt, n, final_t, tol,
                                y0, ptr(y), C.byref(t), C.byref(calls))
    return y, t.value, steps, calls.value


class OrcQdCfg(C.Structure):
    _fields_ = [("order", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("var_grid", C.c_int),
                ("Ly", C.c_double), ("Lz", C.c_double), ("bcR", C.c_int), ("bcT", C.c_int),
                ("bcL", C.c_int), ("bcB", C.c_int), ("remoteLoadingR", C.c_int),
                ("faultTypeScale", C.c_double), ("vL", C.c_double), ("f0", C.c_double),
                ("v0", C.c_double), ("rootTol", C.c_double), ("stateLaw", C.c_int),
                ("jac_mode", C.c_int), ("solver", C.c_int), ("cg_rtol", C.c_double)]


class OrcQdRun(C.Structure):
    _fields_ = [("integrator", C.c_int), ("control", C.c_int), ("normType", C.c_int),
                ("maxSteps", C.c_int), ("initTime", C.c_double), ("maxTime", C.c_double),
                ("initDt", C.c_double), ("minDt", C.c_double), ("maxDt", C.c_double),
                ("tol", C.c_double), ("nReplay", C.c_int), ("replayDt", _dp),
                ("currErr", C.c_double), ("prevErr", C.c_double)]


class PowerLawOracle:
    """StrikeSlip_PowerLaw_qd of the oracle (dislocation creep).  cfg is a
    scycle_amd.configs.PowerLawConfig."""

    def __init__(self, cfg, solver=0, jac_mode=1):
        c = OrcQdCfg(cfg.order, cfg.Ny, cfg.Nz, int(cfg.var_grid), cfg.Ly, cfg.Lz, *cfg.mat_bc,
                     int(cfg.remote_loading_R), cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0,
                     cfg.root_tol, cfg.state_law, jac_mode, solver, 1e-12)
        self._keep = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in
                      (cfg.mu, cfg.y, cfg.z, cfg.a, cfg.b, cfg.Dc, cfg.sN, cfg.eta, cfg.locked,
                       cfg.prestress, cfg.bcRShift, cfg.T, cfg.A, cfg.B, cfg.n)]
        self.N, self.Nz = cfg.Ny * cfg.Nz, cfg.Nz
        self.h = lib().orc_pl_new(C.byref(c), *[ptr(x) for x in self._keep], cfg.visc_cap)
        if not self.h:
            raise RuntimeError("orc_pl_new failed")

    def __del__(self):
        try:
            lib().orc_pl_free(self.h)
        except Exception:
            pass

    def source(self, gxy, gxz):
        out = np.zeros(self.N)
        lib().orc_pl_source(self.h, ptr(np.ascontiguousarray(gxy, float)),
                            ptr(np.ascontiguousarray(gxz, float)), ptr(out))
        return out

    def sat(self, u, gL, gR):
        out = np.zeros(self.N)
        lib().orc_pl_sat(self.h, *[ptr(np.ascontiguousarray(x, float)) for x in (u, gL, gR)], ptr(out))
        return out

    def d_dt(self, t, psi, slip, gxy, gxz, V=None):
        N, Nz = self.N, self.Nz
        V = np.zeros(Nz) if V is None else np.array(V, dtype=np.float64)
        out = dict(dpsi=np.zeros(Nz), dslip=np.zeros(Nz), dgxy=np.zeros(N), dgxz=np.zeros(N),
                   u=np.zeros(N), sxy=np.zeros(N), effVisc=np.zeros(N))
        e = lib().orc_pl_ddt(self.h, t, *[ptr(np.ascontiguousarray(x, float)) for x in (psi, slip, gxy, gxz)],
                             ptr(V), *[ptr(out[k]) for k in ("dpsi", "dslip", "dgxy", "dgxz", "u", "sxy",
                                                              "effVisc")])
        if e:
            raise RuntimeError(f"orc_pl_ddt failed ({e})")
        out["V"] = V
        return out

    def integrate(self, steps, psi, slip, gxy, gxz, integrator=0, tol=1e-8, init_dt=1e-3,
                  init_time=0.0, max_time=1e30, max_dt=1e30, replay=None, err_mask=3, V=None):
        N, Nz = self.N, self.Nz
        st = [np.array(x, dtype=np.float64) for x in (psi, slip, gxy, gxz)]
        Vv = np.zeros(Nz) if V is None else np.array(V, dtype=np.float64)
        rp = None if replay is None else np.ascontiguousarray(replay, dtype=np.float64)
        r = OrcQdRun(integrator, 1, 0, steps, init_time, max_time, init_dt, 0.0, max_dt, tol,
                     0 if rp is None else len(rp), ptr(rp))
        tEnd = C.c_double()
        dts = np.zeros(max(steps, 1))
        n = lib().orc_pl_integrate(self.h, C.byref(r), err_mask, *[ptr(x) for x in st], ptr(Vv),
                                   C.byref(tEnd), ptr(dts), len(dts))
        if n < 0:
            raise RuntimeError(f"orc_pl_integrate failed ({n})")
        return dict(psi=st[0], slip=st[1], gVxy=st[2], gVxz=st[3], V=Vv, t=tEnd.value, steps=n,
                    dts=dts[:n])


class OracleQd:
    """StrikeSlip_LinearElastic_qd of the oracle (direct Cholesky solves by
    default) for a scycle_amd.configs.QdConfig, with a, b, Dc replaceable."""

    def __init__(self, cfg, a=None, b=None, Dc=None, solver=0, jac_mode=1):
        c = OrcQdCfg(cfg.order, cfg.Ny, cfg.Nz, int(cfg.var_grid), cfg.Ly, cfg.Lz, *cfg.mat_bc,
                     int(cfg.remote_loading_R), cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0,
                     cfg.root_tol, cfg.state_law, jac_mode, solver, 1e-12)
        fr = [cfg.a if a is None else a, cfg.b if b is None else b, cfg.Dc if Dc is None else Dc]
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                      (cfg.mu, cfg.y, cfg.z, *fr, cfg.sN, cfg.eta)]
        self.N, self.Nz = cfg.Ny * cfg.Nz, cfg.Nz
        self._opt = [None if x is None else np.ascontiguousarray(x, dtype=np.float64)
                     for x in (cfg.locked, cfg.prestress, cfg.bcRShift)]
        self.h = lib().orc_qd_new(C.byref(c), *[ptr(x) for x in self._keep], *[ptr(x) for x in self._opt])
        if not self.h:
            raise RuntimeError("orc_qd_new failed")

    def __del__(self):
        try:
            lib().orc_qd_free(self.h)
        except Exception:
            pass

    def replay(self, dts, psi, slip, V=None, t0=0.0):
        """RK32 with the step sizes dts; returns the final psi, slip, V and the
        displacement of the last d_dt (at the final state)."""
        Nz = self.Nz
        st = [np.array(x, dtype=np.float64) for x in (psi, slip)]
        Vv = np.zeros(Nz) if V is None else np.array(V, dtype=np.float64)
        rp = np.ascontiguousarray(dts, dtype=np.float64)
        r = OrcQdRun(0, 1, 0, len(rp), t0, 1e30, rp[0], 0.0, 1e30, 1e-8, len(rp), ptr(rp))
        tau, u = np.zeros(Nz), np.zeros(self.N)
        tEnd = C.c_double()
        out_dts = np.zeros(len(rp))
        n = lib().orc_qd_integrate(self.h, C.byref(r), ptr(st[0]), ptr(st[1]), ptr(Vv), ptr(tau),
                                   ptr(u), C.byref(tEnd), ptr(out_dts), len(out_dts), None)
        if n != len(rp):
            raise RuntimeError(f"orc_qd_integrate failed ({n})")
        return dict(psi=st[0], slip=st[1], V=Vv, u=u, t=tEnd.value)


class OrcQdFdCfg(C.Structure):
    _fields_ = [("fdBcR", C.c_int), ("fdBcT", C.c_int), ("fdBcL", C.c_int), ("fdBcB", C.c_int),
                ("outR", C.c_int), ("outT", C.c_int), ("outB", C.c_int), ("ayCompat", C.c_int),
                ("deltaT_fd", C.c_double), ("CFL", C.c_double), ("trigger_qd2fd", C.c_double),
                ("trigger_fd2qd", C.c_double), ("limit_qd", C.c_double), ("limit_fd", C.c_double),
                ("maxNumCycles", C.c_int), ("fdStop", C.c_int)]


class OrcQdFdRun(C.Structure):
    _fields_ = [("integrator", C.c_int), ("control", C.c_int), ("normType", C.c_int),
                ("maxStepCount", C.c_int), ("initTime", C.c_double), ("maxTime", C.c_double),
                ("minDt", C.c_double), ("maxDt", C.c_double), ("tol", C.c_double)]


class OrcQdFdPhase(C.Structure):
    _fields_ = [("regime", C.c_int), ("stepStart", C.c_int), ("stepEnd", C.c_int),
                ("tStart", C.c_double), ("tEnd", C.c_double), ("R", C.c_double)]


class OracleQdFd:
    """strikeSlip_linearElastic_qd_fd of the oracle: QD phases with direct
    Cholesky solves, FD phases with the assembled operator."""

    def __init__(self, cfg, rho, fd_bc=(1, 1, 1, 1), out=(1, 0, 1), ay_compat=True, deltaT_fd=-1.0,
                 cfl=-1.0, trigger_qd2fd=1e-3, trigger_fd2qd=1e-3, limit_qd=None, limit_fd=1e-1,
                 max_num_cycles=1000, fd_stop=False, jac_mode=1):
        self.qd = OracleQd(cfg, jac_mode=jac_mode)
        N = cfg.Ny * cfg.Nz
        self.N, self.Nz = N, cfg.Nz
        rho_a = np.broadcast_to(np.asarray(rho, dtype=np.float64), (N,)).copy()
        cs = np.sqrt(np.asarray(cfg.mu, dtype=np.float64) / rho_a)
        c = OrcQdFdCfg(*[int(x) for x in fd_bc], *[int(x) for x in out], int(ay_compat), deltaT_fd, cfl,
                       trigger_qd2fd, trigger_fd2qd, 10 * cfg.vL if limit_qd is None else limit_qd,
                       limit_fd, int(max_num_cycles), int(fd_stop))
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (cfg.mu, cfg.y, cfg.z, rho_a, cs)]
        self.h = lib().orc_qdfd_new(self.qd.h, C.byref(c), *[ptr(x) for x in self._keep])
        if not self.h:
            raise RuntimeError("orc_qdfd_new failed")

    def __del__(self):
        try:
            lib().orc_qdfd_free(self.h)
        except Exception:
            pass

    @property
    def deltaT_fd(self):
        return lib().orc_qdfd_dt(self.h)

    def ay(self):
        out = np.zeros(self.N)
        lib().orc_qdfd_ay(self.h, ptr(out))
        return out

    def integrate(self, psi, slip, V, max_step_count, max_time=1e15, init_time=0.0, integrator=0,
                  tol=1e-8, min_dt=1e-3, max_dt=1e10, control=1, log_cap=4096):
        st = [np.array(x, dtype=np.float64) for x in (psi, slip, V)]
        u = np.zeros(self.N)
        r = OrcQdFdRun(integrator, control, 0, int(max_step_count), init_time, max_time, min_dt,
                       max_dt, tol)
        log = (OrcQdFdPhase * log_cap)()
        tEnd = C.c_double()
        steps, cycles, regime = C.c_int(), C.c_int(), C.c_int()
        n = lib().orc_qdfd_integrate(self.h, C.byref(r), *[ptr(x) for x in st], ptr(u), C.byref(tEnd),
                                     C.byref(steps), C.byref(cycles), C.byref(regime), log, log_cap)
        if n < 0:
            raise RuntimeError(f"orc_qdfd_integrate failed ({n})")
        names = {0: "qd", 1: "fd", 2: "qd2fd_step"}
        phases = [dict(regime=names[p.regime], step_start=p.stepStart, step_end=p.stepEnd,
                       t_start=p.tStart, t_end=p.tEnd, R=p.R) for p in log[:min(n, log_cap)]]
        return dict(psi=st[0], slip=st[1], V=st[2], u=u, t=tEnd.value, steps=steps.value,
                    cycles=cycles.value, regime=names[regime.value], phases=phases)
