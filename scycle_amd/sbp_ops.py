"""Python mirror of the reference's SbpOps interface (sbpOps.hpp:42-104) over
the matrix-free HIP operator.  Methods take and return host numpy arrays and
move data through device buffers; the computation runs in libscycle_hip.so.
"""
import ctypes as C

import numpy as np

from ._lib import GridDesc, SolveOpts, SolveStats, check, dptr, load
from .device import DeviceArray

DIRICHLET, NEUMANN = 0, 1
_LAPLACE = {"yz": 0, "y": 1, "z": 2}
PRECONDS = {"jacobi": 0, "fdiag64": 1, "fdiag32": 2, "fdiag_bf16": 3}


class SbpOps:
    def __init__(self, order, Ny, Nz, Ly, Lz, mu, y=None, z=None, var_grid=True,
                 bc=(DIRICHLET, NEUMANN, DIRICHLET, NEUMANN), compat="fullyCompatible",
                 laplace="yz"):
        """bc is (R, T, L, B) like SbpOps::setBCTypes."""
        self.Ny, self.Nz, self.N = Ny, Nz, Ny * Nz
        d = GridDesc(Ny, Nz, float(Ly), float(Lz), order, int(var_grid),
                     0 if compat == "fullyCompatible" else 1, _LAPLACE[laplace], *bc)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        zz = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
        h = C.c_void_p()
        check(load().sc_ctx_create(C.byref(d), dptr(yy), dptr(zz), dptr(mu), C.byref(h)))
        self.h = h
        self._a = DeviceArray(self.N)
        self._b = DeviceArray(self.N)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().sc_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(load().sc_ctx_sync(self.h))

    # ---------------------------------------------------------------- ops
    def apply(self, u, prec="fp64"):
        """A*u (getA + MatMult)."""
        if prec == "fp64":
            self._a.upload(u)
            check(load().sc_sbp_apply(self.h, 0, self._a.ptr, self._b.ptr))
            self.sync()
            return self._b.to_host()
        a = DeviceArray(self.N, np.float32, u)
        b = DeviceArray(self.N, np.float32)
        check(load().sc_sbp_apply(self.h, 1, a.ptr, b.ptr))
        self.sync()
        return b.to_host()

    def _unary(self, fn, u):
        self._a.upload(u)
        check(fn(self.h, self._a.ptr, self._b.ptr))
        self.sync()
        return self._b.to_host()

    def Dy(self, u):
        return self._unary(load().sc_sbp_dy, u)

    def muxDy(self, u):
        return self._unary(load().sc_sbp_mux_dy, u)

    def Dz(self, u):
        return self._unary(load().sc_sbp_dz, u)

    def muxDz(self, u):
        return self._unary(load().sc_sbp_mux_dz, u)

    def Dyxmu(self, u):
        return self._unary(load().sc_sbp_dyxmu, u)

    def Dzxmu(self, u):
        return self._unary(load().sc_sbp_dzxmu, u)

    def _face(self, fn, face):
        """fn on Nz values of a face (Hyinvxe0y, HyinvxeNy): a grid field."""
        a = DeviceArray(self.Nz, host=np.ascontiguousarray(face, dtype=np.float64))
        check(fn(self.h, a.ptr, self._b.ptr))
        self.sync()
        return self._b.to_host()

    def Hyinvxe0y(self, face):
        return self._face(load().sc_sbp_hyinvx_e0y, face)

    def HyinvxeNy(self, face):
        return self._face(load().sc_sbp_hyinvx_eNy, face)

    def HyinvxE0y(self, u):
        return self._unary(load().sc_sbp_hyinvx_E0y, u)

    def HyinvxENy(self, u):
        return self._unary(load().sc_sbp_hyinvx_ENy, u)

    def HzinvxE0z(self, u):
        return self._unary(load().sc_sbp_hzinvx_E0z, u)

    def HzinvxENz(self, u):
        return self._unary(load().sc_sbp_hzinvx_ENz, u)

    def H(self, u):
        return self._unary(load().sc_sbp_h, u)

    def Hinv(self, u):
        return self._unary(load().sc_sbp_hinv, u)

    def diag(self):
        check(load().sc_sbp_diag(self.h, self._b.ptr))
        self.sync()
        return self._b.to_host()

    def setRhs(self, bcL, bcR, bcT, bcB):
        arrs = [DeviceArray(len(x), host=x) for x in (bcL, bcR, bcT, bcB)]
        check(load().sc_sbp_set_rhs(self.h, self._b.ptr, *[a.ptr for a in arrs]))
        self.sync()
        return self._b.to_host()

    def changeBCTypes(self, bcR, bcT, bcL, bcB):
        check(load().sc_change_bc_types(self.h, bcR, bcT, bcL, bcB))

    def updateVarCoeff(self, mu):
        check(load().sc_update_var_coeff(self.h, dptr(np.ascontiguousarray(mu, dtype=np.float64))))

    def geth11(self):
        a, b = C.c_double(), C.c_double()
        check(load().sc_geth11(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def metric(self):
        yq, zr = np.zeros(self.Ny), np.zeros(self.Nz)
        check(load().sc_get_metric(self.h, dptr(yq), dptr(zr)))
        return yq, zr

    def pc_apply(self, r, precond="fdiag64"):
        """PCApply: z = M^-1 r with one of the solve's preconditioners."""
        self._a.upload(r)
        check(load().sc_precond_apply(self.h, PRECONDS[precond], self._a.ptr, self._b.ptr))
        self.sync()
        return self._b.to_host()

    def solve(self, rhs, u0=None, mode="cg64", rtol=1e-12, inner_rtol=1e-5, max_iters=200000,
              check_every=64, precond="jacobi", reuse=0):
        """KSPSolve(A, rhs, u) with u0 as the initial guess; returns (u, stats).
        precond: "jacobi", "fdiag64", "fdiag32" or "fdiag_bf16"; reuse: number
        of previous solutions of this operator kept to project the initial
        guess (0: start from u0)."""
        b = DeviceArray(self.N, host=rhs)
        u = DeviceArray(self.N, host=np.zeros(self.N) if u0 is None else u0)
        o = SolveOpts(0 if mode == "cg64" else 1, rtol, inner_rtol, max_iters, 50, check_every,
                      PRECONDS[precond], reuse)
        st = SolveStats()
        check(load().sc_solve(self.h, b.ptr, u.ptr, C.byref(o), C.byref(st)))
        self.sync()
        return u.to_host(), dict(iters=st.iters, outer=st.outer, converged=st.converged,
                                 relres=st.relres)
