"""Python mirror of StrikeSlip_LinearElastic_fd (strikeSlip_linearElastic_fd.cpp)
driven through the C ABI: the wave step (propagateWaves) and Fault_fd::d_dt
run in libscycle_hip.so, stepped by the constant-dt loop of
OdeSolver_WaveEq::integrate (odeSolver_WaveEq.cpp:147-186)."""
import ctypes as C

import numpy as np

from ._lib import FdDesc, GridDesc, QdOutput, check, dptr, load

BULK = {"fp64": 0, "fp32": 1}


class FdMediator:
    def __init__(self, cfg, bulk="fp64"):
        """cfg: a scycle_amd.configs.FdConfig; bulk "fp64" or "fp32" (body
        arrays and update in FP32, fault state and root solves in FP64)."""
        self.cfg = cfg
        g = GridDesc(cfg.Ny, cfg.Nz, cfg.Ly, cfg.Lz, cfg.order, int(cfg.var_grid), 0, 0,
                     1, 1, 1, 1)
        d = FdDesc(g, int(cfg.outR), int(cfg.outT), int(cfg.outB), int(cfg.faultL),
                   cfg.fault_type_scale, cfg.vL, cfg.f0, cfg.v0, cfg.root_tol, 10000,
                   cfg.state_law, BULK[bulk])
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
                      for x in (cfg.mu, cfg.rho, cfg.cs, cfg.y, cfg.z, cfg.a, cfg.b, cfg.Dc,
                                cfg.sNEff, cfg.tau0, cfg.prestress, cfg.slip0, cfg.locked)]
        h = C.c_void_p()
        check(load().sc_fd_create(C.byref(d), *[dptr(x) for x in self._keep], C.byref(h)))
        self.h = h
        self.Nz, self.N = cfg.Nz, cfg.Ny * cfg.Nz

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().sc_fd_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, u_prev, u, psi_prev, psi, slip_prev, slip, V=None):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
                (u_prev, u, psi_prev, psi, slip_prev, slip)]
        Vc = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        check(load().sc_fd_set_state(self.h, *[dptr(x) for x in arrs], dptr(Vc)))

    def steps(self, n, dt, t0=0.0):
        check(load().sc_fd_steps(self.h, t0, dt, int(n)))

    def set_output(self, output_dir, stride1D=1, stride2D=0):
        """Output series written by integrate (timeMonitor of
        strikeSlip_linearElastic_fd.cpp:218-251); output_dir is the prefix of
        every file name.  None switches output off."""
        if output_dir is None:
            check(load().sc_fd_set_output(self.h, None))
            return
        self._out_dir = str(output_dir).encode()
        o = QdOutput(self._out_dir, int(stride1D), int(stride2D), 0, 1)
        check(load().sc_fd_set_output(self.h, C.byref(o)))

    def integrate(self, max_steps, dt, init_time=0.0, max_time=1e15):
        """OdeSolver_WaveEq::integrate from the current state; returns
        (tEnd, steps)."""
        tEnd, n = C.c_double(), C.c_int()
        check(load().sc_fd_integrate(self.h, float(init_time), float(dt), int(max_steps), float(max_time),
                                     C.byref(tEnd), C.byref(n)))
        return tEnd.value, n.value

    def get_state(self):
        u = np.empty(self.N)
        out = {k: np.empty(self.Nz) for k in ("psi", "slip", "V", "tau", "strength")}
        check(load().sc_fd_get_state(self.h, dptr(u), dptr(out["psi"]), dptr(out["slip"]),
                                     dptr(out["V"]), dptr(out["tau"]), dptr(out["strength"])))
        out["u"] = u
        return out
