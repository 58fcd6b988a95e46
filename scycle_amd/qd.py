"""Python mirror of StrikeSlip_LinearElastic_qd (strikeSlip_linearElastic_qd.cpp)
driven through the C ABI: d_dt and the RK32/RK43 integrate loop run in
libscycle_hip.so."""
import ctypes as C

import numpy as np

from ._lib import (AdjDesc, AdjStats, GridDesc, QdCounters, QdDesc, QdOutput, QdRun, SolveOpts, check,
                   dptr, load)

SOLVE_MODES = {"cg64": 0, "ir32": 1}
GREEN_MODES = {"auto": 0, "always": 1, "never": 2}


class QdMediator:
    def __init__(self, cfg, solve="cg64", rtol=1e-12, inner_rtol=1e-5, jac_mode=1,
                 check_every=64, max_iters=400000, precond="jacobi", reuse=0):
        """cfg: a scycle_amd.configs.QdConfig; solve "cg64" or "ir32";
        precond "jacobi", "fdiag64", "fdiag32" or "fdiag_bf16" (cg64 only);
        reuse: number of previous solutions projected for the initial guess."""
        from .sbp_ops import PRECONDS
        self.cfg = cfg
        g = GridDesc(cfg.Ny, cfg.Nz, cfg.Ly, cfg.Lz, cfg.order, int(cfg.var_grid), 0, 0,
                     *cfg.mat_bc)
        so = SolveOpts(SOLVE_MODES[solve], rtol, inner_rtol, max_iters, 50, check_every,
                       PRECONDS[precond], reuse)
        d = QdDesc(g, int(cfg.remote_loading_R), cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0,
                   cfg.root_tol, 10000, cfg.state_law, jac_mode, so)
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
                      for x in (cfg.mu, cfg.y, cfg.z, cfg.a, cfg.b, cfg.Dc, cfg.sN, cfg.eta,
                                cfg.locked, cfg.prestress, cfg.bcRShift)]
        h = C.c_void_p()
        self._create(d, h)
        self.h = h
        self.Nz, self.N = cfg.Nz, cfg.Ny * cfg.Nz

    def _create(self, d, h):
        check(load().sc_qd_create(C.byref(d), *[dptr(x) for x in self._keep], C.byref(h)))

    @classmethod
    def from_file(cls, in_path):
        """The mediator a reference .in deck describes (sc_qd_create_from_file:
        the deck's keys, binary Vec overrides from its inputDir, solveSS when
        it sets guessSteadyStateICs = 1)."""
        self = cls.__new__(cls)
        p = str(in_path).encode()
        buf = C.create_string_buffer(64)
        dims = []
        for key in (b"Ny", b"Nz"):
            check(load().sc_in_get(p, key, buf, 64))
            dims.append(int(float(buf.value)))
        self.cfg = None
        self.Nz, self.N = dims[1], dims[0] * dims[1]
        h = C.c_void_p()
        check(load().sc_qd_create_from_file(p, C.byref(h)))
        self.h = h
        return self

    def file_run(self):
        """The deck's time-integration settings (sc_qd_file_run) as a QdRun."""
        r = QdRun()
        check(load().sc_qd_file_run(self.h, C.byref(r)))
        return r

    def solve_ss(self, psi, have_state_vals=False):
        """StrikeSlip_LinearElastic_qd::solveSS: returns (psi, slip) and leaves
        the mediator at the steady-state start (V = vL)."""
        p = np.ascontiguousarray(psi, dtype=np.float64).copy()
        slip = np.zeros(self.Nz)
        check(load().sc_qd_solve_ss(self.h, dptr(p), int(bool(have_state_vals)), dptr(slip)))
        return p, slip

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().sc_qd_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, psi, slip, V=None):
        p = np.ascontiguousarray(psi, dtype=np.float64)
        s = np.ascontiguousarray(slip, dtype=np.float64)
        v = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        check(load().sc_qd_set_state(self.h, dptr(p), dptr(s), dptr(v)))

    def get_state(self):
        psi, slip, V, tau = (np.zeros(self.Nz) for _ in range(4))
        check(load().sc_qd_get_state(self.h, dptr(psi), dptr(slip), dptr(V), dptr(tau)))
        return dict(psi=psi, slip=slip, V=V, tau=tau)

    def get_u(self):
        u = np.zeros(self.N)
        check(load().sc_qd_get_u(self.h, dptr(u)))
        return u

    def d_dt(self, t):
        dpsi, dslip = np.zeros(self.Nz), np.zeros(self.Nz)
        check(load().sc_qd_ddt(self.h, float(t), dptr(dpsi), dptr(dslip)))
        return dpsi, dslip

    def integrate(self, max_steps, integrator="RK32", tol=1e-8, init_dt=1e-3, min_dt=0.0,
                  max_dt=1e10, init_time=0.0, max_time=1e15, control="PID",
                  norm="L2_absolute", replay_dt=None, defer="auto", curr_err=0.0, prev_err=0.0,
                  step_control="auto"):
        """defer: "auto", "never" or "always" (sc_qd_run.deferChecks);
        curr_err, prev_err: the PID error history at the start (a restart's
        chkpt_currErr / chkpt_prevErr, see load_checkpoint); step_control:
        "auto" (on the device where it can) or "host" (sc_qd_run.stepControl)."""
        return self.integrate_run(self._run(max_steps, integrator, tol, init_dt, min_dt, max_dt, init_time,
                                            max_time, control, norm, replay_dt, defer, curr_err, prev_err,
                                            step_control))

    def _run(self, max_steps, integrator="RK32", tol=1e-8, init_dt=1e-3, min_dt=0.0, max_dt=1e10,
             init_time=0.0, max_time=1e15, control="PID", norm="L2_absolute", replay_dt=None,
             defer="auto", curr_err=0.0, prev_err=0.0, step_control="auto"):
        rep = None if replay_dt is None else np.ascontiguousarray(replay_dt, dtype=np.float64)
        self._replay_keep = rep
        return QdRun(0 if integrator == "RK32" else 1, 0 if control == "P" else 1,
                     0 if norm == "L2_absolute" else 1, int(max_steps), init_time, max_time,
                     init_dt, min_dt, max_dt, tol, 0 if rep is None else len(rep), dptr(rep),
                     {"auto": 0, "never": 1, "always": 2}[defer], curr_err, prev_err,
                     {"auto": 0, "host": 1}[step_control])

    def integrate_timed(self, max_steps, **kw):
        """integrate() with the device loop's stage launches timed one by one
        with HIP events (sc_qd_integrate_timed); returns (tEnd, steps, dts,
        mean in-step ms of the plain stage kernel, of the attempt-end kernel,
        attempts timed)."""
        r = self._run(max_steps, **kw)
        tEnd, steps = C.c_double(), C.c_int()
        dts = np.zeros(max(int(r.maxSteps), 1))
        ms_s, ms_e, att = C.c_double(), C.c_double(), C.c_longlong()
        check(load().sc_qd_integrate_timed(self.h, C.byref(r), C.byref(tEnd), C.byref(steps), dptr(dts),
                                           len(dts), C.byref(ms_s), C.byref(ms_e), C.byref(att)))
        return tEnd.value, steps.value, dts[:steps.value].copy(), ms_s.value, ms_e.value, att.value

    def integrate_run(self, r):
        """sc_qd_integrate with a QdRun (e.g. file_run()); returns (tEnd,
        steps, accepted step sizes)."""
        tEnd = C.c_double()
        steps = C.c_int()
        dts = np.zeros(max(int(r.maxSteps), 1))
        check(load().sc_qd_integrate(self.h, C.byref(r), C.byref(tEnd), C.byref(steps),
                                     dptr(dts), len(dts)))
        return tEnd.value, steps.value, dts[:steps.value].copy()

    # ------------------------------------------- output and checkpoints --
    def set_output(self, output_dir, stride1D=1, stride2D=0, checkpoint=0, interval=10000):
        """Output series and checkpoints written by integrate (timeMonitor of
        strikeSlip_linearElastic_qd.cpp:417-463); output_dir is the prefix of
        every file name.  None switches output off."""
        if output_dir is None:
            check(load().sc_qd_set_output(self.h, None))
            return
        self._out_dir = str(output_dir).encode()
        o = QdOutput(self._out_dir, int(stride1D), int(stride2D), int(checkpoint), int(interval))
        check(load().sc_qd_set_output(self.h, C.byref(o)))

    def load_checkpoint(self):
        """Restore the state of the last checkpoint under the output prefix;
        returns None when there is none, else the integrate() arguments it
        sets: init_time, init_dt, curr_err, prev_err (and ckpt_number)."""
        r = QdRun()
        n = C.c_int()
        check(load().sc_qd_load_checkpoint(self.h, C.byref(r), C.byref(n)))
        if n.value == 0:
            return None
        return dict(init_time=r.initTime, init_dt=r.initDt, curr_err=r.currErr, prev_err=r.prevErr,
                    ckpt_number=n.value)

    # -------------------------------------------------- discrete adjoint --
    def set_friction(self, a=None, b=None, Dc=None):
        """Replace the fault's a, b, Dc (Nz each; None keeps one)."""
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, Dc)]
        check(load().sc_qd_set_friction(self.h, *[dptr(x) for x in arrs]))

    def _adj_desc(self, dts, t0, stations, data, ckpt_every, solve, rtol, precond, reuse, green="auto"):
        from .sbp_ops import PRECONDS
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        js = np.ascontiguousarray(stations, dtype=np.int32)
        dd = None if data is None else np.ascontiguousarray(data, dtype=np.float64)
        so = SolveOpts(SOLVE_MODES[solve], rtol, 1e-5, 400000, 50, 64, PRECONDS[precond], reuse)
        d = AdjDesc(len(dts), dts.ctypes.data, float(t0), int(ckpt_every), len(js), js.ctypes.data,
                    None if dd is None else dd.ctypes.data, so, GREEN_MODES[green])
        return d, (dts, js, dd)

    def forward_misfit(self, dts, stations, data=None, t0=0.0, ckpt_every=50):
        """J = 1/2 sum_s (u(j_s, z=0, t_N) - d_s)^2 after len(dts) RK32 steps with
        these step sizes from the current state (left unchanged); returns
        (J, station displacements)."""
        d, keep = self._adj_desc(dts, t0, stations, data, ckpt_every, "cg64", 1e-12, "jacobi", 0)
        J = C.c_double()
        us = np.zeros(max(len(keep[1]), 1))
        check(load().sc_qd_forward_misfit(self.h, C.byref(d), C.byref(J), dptr(us)))
        return J.value, us[:len(keep[1])]

    def adjoint(self, dts, stations, data=None, t0=0.0, ckpt_every=50, solve="cg64",
                rtol=1e-12, precond="fdiag64", reuse=4, green="auto"):
        """dJ/da, dJ/db, dJ/dDc per fault node by the discrete adjoint; the
        adjoint's elasticity solves use (solve, rtol, precond, reuse).  green:
        "always" applies the adjoint solves of the reverse pass through the
        precomputed Nz x Nz boundary operator, "never" solves per stage,
        "auto" chooses (sc_adj_desc.greenMode).  Returns (ga, gb, gDc, stats)."""
        d, keep = self._adj_desc(dts, t0, stations, data, ckpt_every, solve, rtol, precond, reuse, green)
        g = [np.zeros(self.Nz) for _ in range(3)]
        st = AdjStats()
        check(load().sc_qd_adjoint(self.h, C.byref(d), *[dptr(x) for x in g], C.byref(st)))
        return g[0], g[1], g[2], {k: getattr(st, k) for k, _ in AdjStats._fields_}

    def counters(self):
        c = QdCounters()
        check(load().sc_qd_get_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in QdCounters._fields_}


ERR_FIELDS = {"psi": 1, "slip": 2, "gVxy": 4, "gVxz": 8}


class PowerLawMediator(QdMediator):
    """StrikeSlip_PowerLaw_qd (strikeSlip_powerLaw_qd.cpp) with dislocation
    creep: the integrated variables are psi, slip and the viscous strains
    gVxy, gVxz.  cfg is a scycle_amd.configs.PowerLawConfig; err_inds are the
    RK error fields (timeIntInds)."""

    def __init__(self, cfg, err_inds=("psi", "slip"), **kw):
        self._pl = [np.ascontiguousarray(x, dtype=np.float64) for x in (cfg.T, cfg.A, cfg.B, cfg.n)]
        self._mask = sum(ERR_FIELDS[k] for k in err_inds)
        super().__init__(cfg, **kw)

    def _create(self, d, h):
        check(load().sc_qd_create_powerlaw(C.byref(d), *[dptr(x) for x in self._keep],
                                           *[dptr(x) for x in self._pl], self.cfg.visc_cap,
                                           self._mask, C.byref(h)))

    def set_visc(self, gVxy=None, gVxz=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (gVxy, gVxz)]
        check(load().sc_qd_set_visc(self.h, dptr(a[0]), dptr(a[1])))

    def get_visc(self):
        out = {k: np.zeros(self.N) for k in ("gVxy", "gVxz", "dgVxy", "dgVxz", "effVisc", "sxy")}
        check(load().sc_qd_get_visc(self.h, *[dptr(out[k]) for k in
                                             ("gVxy", "gVxz", "dgVxy", "dgVxz", "effVisc", "sxy")]))
        return out
