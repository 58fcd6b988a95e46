"""Python mirror of strikeSlip_linearElastic_qd_fd (strikeSlip_linearElastic_qd_fd.cpp),
the mediator that switches between quasi-dynamic and fully dynamic phases,
driven through the C ABI (sc_qdfd_*)."""
import ctypes as C

import numpy as np

from ._lib import (GridDesc, QdDesc, QdFdDesc, QdFdPhase, QdFdRun, SolveOpts, check, dptr, load)
from .fd import BULK
from .qd import SOLVE_MODES

REGIMES = {0: "qd", 1: "fd", 2: "qd2fd_step"}


class QdFdMediator:
    def __init__(self, cfg, rho, fd_bc=(1, 1, 1, 1), out=(1, 0, 1), ay_compat=True, bulk="fp64",
                 deltaT_fd=-1.0, cfl=-1.0, trigger_qd2fd=1e-3, trigger_fd2qd=1e-3, limit_qd=None,
                 limit_fd=1e-1, max_num_cycles=1000, fd_stop=False, solve="cg64", rtol=1e-12,
                 precond="jacobi", reuse=0, jac_mode=1):
        """cfg: a scycle_amd.configs.QdConfig (grid, QD-phase SAT types,
        friction, eta); rho: density (scalar or N values), cs = sqrt(mu/rho);
        fd_bc: material SAT types of the FD phase (R, T, L, B; reference:
        Neumann); out: outgoing-characteristics flags (R, T, B) of the FD
        phase (reference: R and B); ay_compat: the reference's a_y of this
        mediator; fd_stop: FD phases end on trigger_fd2qd (the reference runs
        them to the end); limit_qd defaults to 10 vL as the reference."""
        from .sbp_ops import PRECONDS
        self.cfg = cfg
        N = cfg.Ny * cfg.Nz
        g = GridDesc(cfg.Ny, cfg.Nz, cfg.Ly, cfg.Lz, cfg.order, int(cfg.var_grid), 0, 0, *cfg.mat_bc)
        so = SolveOpts(SOLVE_MODES[solve], rtol, 1e-5, 400000, 50, 64, PRECONDS[precond], reuse)
        qd = QdDesc(g, int(cfg.remote_loading_R), cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0,
                    cfg.root_tol, 10000, cfg.state_law, jac_mode, so)
        d = QdFdDesc(qd, *[int(x) for x in fd_bc], *[int(x) for x in out], 1, int(ay_compat),
                     BULK[bulk], deltaT_fd, cfl, trigger_qd2fd, trigger_fd2qd,
                     10 * cfg.vL if limit_qd is None else limit_qd, limit_fd, int(max_num_cycles),
                     int(fd_stop))
        rho_a = np.broadcast_to(np.asarray(rho, dtype=np.float64), (N,)).copy()
        cs = np.sqrt(np.asarray(cfg.mu, dtype=np.float64) / rho_a)
        self.rho, self.cs = rho_a, cs
        self._keep = [None if x is None else np.ascontiguousarray(x, dtype=np.float64)
                      for x in (cfg.mu, rho_a, cs, cfg.y, cfg.z, cfg.a, cfg.b, cfg.Dc, cfg.sN,
                                cfg.eta, cfg.locked, cfg.prestress, cfg.bcRShift)]
        h = C.c_void_p()
        check(load().sc_qdfd_create(C.byref(d), *[dptr(x) for x in self._keep], C.byref(h)))
        self.h = h
        self.Nz, self.N = cfg.Nz, N

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().sc_qdfd_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def deltaT_fd(self):
        dt = C.c_double()
        check(load().sc_qdfd_time_step(self.h, C.byref(dt)))
        return dt.value

    def set_state(self, psi, slip, V=None):
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (psi, slip, V)]
        check(load().sc_qdfd_set_state(self.h, *[dptr(x) for x in arrs]))

    def get_state(self):
        out = {k: np.zeros(self.Nz) for k in ("psi", "slip", "V")}
        u = np.zeros(self.N)
        reg = C.c_int()
        check(load().sc_qdfd_get_state(self.h, dptr(out["psi"]), dptr(out["slip"]), dptr(out["V"]),
                                       dptr(u), C.byref(reg)))
        out["u"] = u
        out["regime"] = REGIMES[reg.value]
        return out

    def integrate(self, max_step_count, max_time=1e15, init_time=0.0, integrator="RK32", tol=1e-8,
                  min_dt=1e-3, max_dt=1e10, control="PID", norm="L2_absolute", defer="auto",
                  log_cap=4096):
        """strikeSlip_linearElastic_qd_fd::integrate; returns (tEnd, steps,
        cycles, phases) with phases a list of dicts (regime, steps, times,
        switching ratio at the phase's last step)."""
        r = QdFdRun(0 if integrator == "RK32" else 1, 0 if control == "P" else 1,
                    0 if norm == "L2_absolute" else 1, int(max_step_count), init_time, max_time,
                    min_dt, max_dt, tol, {"auto": 0, "never": 1, "always": 2}[defer])
        log = (QdFdPhase * log_cap)()
        tEnd, steps, cycles, n = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        check(load().sc_qdfd_integrate(self.h, C.byref(r), C.byref(tEnd), C.byref(steps),
                                       C.byref(cycles), log, log_cap, C.byref(n)))
        phases = [dict(regime=REGIMES[p.regime], step_start=p.stepStart, step_end=p.stepEnd,
                       t_start=p.tStart, t_end=p.tEnd, R=p.R) for p in log[:min(n.value, log_cap)]]
        return tEnd.value, steps.value, cycles.value, phases
