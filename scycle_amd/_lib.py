"""ctypes binding of libscycle_hip.so (the C ABI declared in include/scycle_hip.h).

The product path is the HIP library: if it is not built, or no GPU is visible
when a compute entry point is called, this module raises instead of falling
back to anything else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscycle_hip.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class ScError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("Ny", C.c_int), ("Nz", C.c_int), ("Ly", C.c_double), ("Lz", C.c_double),
                ("order", C.c_int), ("var_grid", C.c_int), ("compat", C.c_int),
                ("laplace", C.c_int), ("bcR", C.c_int), ("bcT", C.c_int), ("bcL", C.c_int),
                ("bcB", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("rtol", C.c_double), ("inner_rtol", C.c_double),
                ("max_iters", C.c_int), ("max_outer", C.c_int), ("check_every", C.c_int),
                ("precond", C.c_int), ("reuse", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer", C.c_int), ("converged", C.c_int),
                ("relres", C.c_double)]


class FaultQdArgs(C.Structure):
    _fields_ = [("N", C.c_int), ("v0", C.c_double), ("vL", C.c_double), ("f0", C.c_double),
                ("rootTol", C.c_double), ("maxNumIts", C.c_int), ("stateLaw", C.c_int),
                ("jac_mode", C.c_int),
                ("eta_d", _vp), ("sN_d", _vp), ("a_d", _vp), ("b_d", _vp), ("Dc_d", _vp),
                ("locked_d", _vp), ("prestress_d", _vp), ("psi_d", _vp), ("tauQS_d", _vp),
                ("V_d", _vp), ("dpsi_d", _vp), ("tau_d", _vp), ("strength_d", _vp),
                ("status_d", _vp)]


class FaultFdArgs(C.Structure):
    _fields_ = [("N", C.c_int), ("dt", C.c_double), ("v0", C.c_double), ("vL", C.c_double),
                ("f0", C.c_double), ("rootTol", C.c_double), ("faultTypeScale", C.c_double),
                ("maxNumIts", C.c_int), ("stateLaw", C.c_int),
                ("a_d", _vp), ("b_d", _vp), ("Dc_d", _vp), ("sNEff_d", _vp), ("rho_d", _vp),
                ("alphay_d", _vp), ("tau0_d", _vp), ("slip0_d", _vp), ("locked_d", _vp),
                ("d2u_d", _vp), ("psi_d", _vp), ("psiPrev_d", _vp), ("slip_d", _vp),
                ("slipPrev_d", _vp), ("V_d", _vp), ("psiNext_d", _vp), ("slipNext_d", _vp),
                ("strength_d", _vp), ("uFault_d", _vp), ("status_d", _vp)]


class FdDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("outR", C.c_int), ("outT", C.c_int), ("outB", C.c_int),
                ("faultL", C.c_int), ("faultTypeScale", C.c_double), ("vL", C.c_double),
                ("f0", C.c_double), ("v0", C.c_double), ("rootTol", C.c_double),
                ("maxNumIts", C.c_int), ("stateLaw", C.c_int), ("bulkPrec", C.c_int)]


class PowerLawArgs(C.Structure):
    _fields_ = [("N", C.c_int), ("effViscCap", C.c_double), ("sxy_d", _vp), ("sxz_d", _vp),
                ("T_d", _vp), ("A_d", _vp), ("B_d", _vp), ("n_d", _vp), ("mu_d", _vp),
                ("sat_d", _vp), ("effVisc_d", _vp), ("gxy_t_d", _vp), ("gxz_t_d", _vp)]


class QdDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("remoteLoadingR", C.c_int), ("faultTypeScale", C.c_double),
                ("vL", C.c_double), ("f0", C.c_double), ("v0", C.c_double),
                ("rootTol", C.c_double), ("maxNumIts", C.c_int), ("stateLaw", C.c_int),
                ("jac_mode", C.c_int), ("solve", SolveOpts)]


class QdRun(C.Structure):
    _fields_ = [("integrator", C.c_int), ("control", C.c_int), ("normType", C.c_int),
                ("maxSteps", C.c_int), ("initTime", C.c_double), ("maxTime", C.c_double),
                ("initDt", C.c_double), ("minDt", C.c_double), ("maxDt", C.c_double),
                ("tol", C.c_double), ("nReplay", C.c_int), ("replayDt_h", _dp),
                ("deferChecks", C.c_int), ("currErr", C.c_double), ("prevErr", C.c_double),
                ("stepControl", C.c_int)]


class QdFdDesc(C.Structure):
    _fields_ = [("qd", QdDesc), ("fdBcR", C.c_int), ("fdBcT", C.c_int), ("fdBcL", C.c_int),
                ("fdBcB", C.c_int), ("outR", C.c_int), ("outT", C.c_int), ("outB", C.c_int),
                ("faultL", C.c_int), ("ayCompat", C.c_int), ("bulkPrec", C.c_int),
                ("deltaT_fd", C.c_double), ("CFL", C.c_double), ("trigger_qd2fd", C.c_double),
                ("trigger_fd2qd", C.c_double), ("limit_qd", C.c_double), ("limit_fd", C.c_double),
                ("maxNumCycles", C.c_int), ("fdStop", C.c_int)]


class QdFdRun(C.Structure):
    _fields_ = [("integrator", C.c_int), ("control", C.c_int), ("normType", C.c_int),
                ("maxStepCount", C.c_int), ("initTime", C.c_double), ("maxTime", C.c_double),
                ("minDt", C.c_double), ("maxDt", C.c_double), ("tol", C.c_double),
                ("deferChecks", C.c_int)]


class QdFdPhase(C.Structure):
    _fields_ = [("regime", C.c_int), ("stepStart", C.c_int), ("stepEnd", C.c_int),
                ("tStart", C.c_double), ("tEnd", C.c_double), ("R", C.c_double)]


class QdOutput(C.Structure):
    _fields_ = [("outputDir", C.c_char_p), ("stride1D", C.c_int), ("stride2D", C.c_int),
                ("enableCheckpointing", C.c_int), ("interval", C.c_int)]


class QdCounters(C.Structure):
    _fields_ = [("d_dt_calls", C.c_longlong), ("solves", C.c_longlong),
                ("cg_iters", C.c_longlong), ("outer_sweeps", C.c_longlong),
                ("rejected_steps", C.c_longlong), ("solve_seconds", C.c_double),
                ("total_seconds", C.c_double), ("deferred_checks", C.c_longlong),
                ("missed_guesses", C.c_longlong), ("device_steps", C.c_longlong),
                ("graph_builds", C.c_longlong)]


class AdjDesc(C.Structure):
    _fields_ = [("nSteps", C.c_int), ("dts", C.c_void_p), ("t0", C.c_double),
                ("ckptEvery", C.c_int), ("nStations", C.c_int), ("stationJ", C.c_void_p),
                ("data", C.c_void_p), ("adjSolve", SolveOpts), ("greenMode", C.c_int)]


class AdjStats(C.Structure):
    _fields_ = [("J", C.c_double), ("forward_seconds", C.c_double),
                ("reverse_seconds", C.c_double), ("fwd_solves", C.c_longlong),
                ("adj_solves", C.c_longlong), ("adj_cg_iters", C.c_longlong),
                ("recomputed_steps", C.c_longlong), ("green_solves", C.c_longlong),
                ("green_seconds", C.c_double)]


# Every symbol include/scycle_hip.h declares, with (restype, argtypes).
SIGNATURES = {
    "sc_last_error": (C.c_char_p, []),
    "sc_version": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sc_malloc": (C.c_int, [C.POINTER(_vp), C.c_ulonglong]),
    "sc_free": (C.c_int, [_vp]),
    "sc_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_ulonglong]),
    "sc_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_ulonglong]),
    "sc_ctx_create": (C.c_int, [C.POINTER(GridDesc), _dp, _dp, _dp, C.POINTER(_vp)]),
    "sc_ctx_destroy": (C.c_int, [_vp]),
    "sc_ctx_sync": (C.c_int, [_vp]),
    "sc_ctx_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sc_update_var_coeff": (C.c_int, [_vp, _dp]),
    "sc_change_bc_types": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "sc_geth11": (C.c_int, [_vp, _dp, _dp]),
    "sc_get_metric": (C.c_int, [_vp, _dp, _dp]),
    "sc_sbp_apply": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "sc_sbp_diag": (C.c_int, [_vp, _vp]),
    "sc_sbp_set_rhs": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sc_sbp_dy": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_mux_dy": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_dz": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_mux_dz": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_h": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hinv": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_dyxmu": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_dzxmu": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hyinvx_e0y": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hyinvx_eNy": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hyinvx_E0y": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hyinvx_ENy": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hzinvx_E0z": (C.c_int, [_vp, _vp, _vp]),
    "sc_sbp_hzinvx_ENz": (C.c_int, [_vp, _vp, _vp]),
    "sc_bench_apply": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "sc_qd_bench_stage": (C.c_int, [_vp, C.c_double, C.c_int, _dp, _dp]),
    "sc_qd_guess_residual": (C.c_int, [_vp, C.c_double, _dp]),
    "sc_qd_integrate_timed": (C.c_int, [_vp, C.POINTER(QdRun), _dp, C.POINTER(C.c_int), _dp, C.c_int,
                                        _dp, _dp, C.POINTER(C.c_longlong)]),
    "sc_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "sc_solve_noalloc": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "sc_precond_apply": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "sc_fault_qd_rates": (C.c_int, [_vp, C.POINTER(FaultQdArgs)]),
    "sc_fault_fd_step": (C.c_int, [_vp, C.POINTER(FaultFdArgs)]),
    "sc_wave_step": (C.c_int, [_vp, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sc_powerlaw_rates": (C.c_int, [_vp, C.c_int, C.POINTER(PowerLawArgs)]),
    "sc_qd_create_powerlaw": (C.c_int, [C.POINTER(QdDesc)] + [_dp] * 15 + [C.c_double, C.c_int,
                                                                        C.POINTER(_vp)]),
    "sc_qd_set_visc": (C.c_int, [_vp, _dp, _dp]),
    "sc_qd_set_friction": (C.c_int, [_vp, _dp, _dp, _dp]),
    "sc_qd_forward_misfit": (C.c_int, [_vp, C.POINTER(AdjDesc), _dp, _dp]),
    "sc_qd_adjoint": (C.c_int, [_vp, C.POINTER(AdjDesc), _dp, _dp, _dp, C.POINTER(AdjStats)]),
    "sc_qd_get_visc": (C.c_int, [_vp] + [_dp] * 6),
    "sc_fd_create": (C.c_int, [C.POINTER(FdDesc)] + [_dp] * 13 + [C.POINTER(_vp)]),
    "sc_fd_destroy": (C.c_int, [_vp]),
    "sc_fd_ctx": (_vp, [_vp]),
    "sc_fd_set_state": (C.c_int, [_vp] + [_dp] * 7),
    "sc_fd_steps": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int]),
    "sc_fd_get_state": (C.c_int, [_vp] + [_dp] * 6),
    "sc_qd_create": (C.c_int, [C.POINTER(QdDesc)] + [_dp] * 11 + [C.POINTER(_vp)]),
    "sc_qd_create_from_file": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "sc_qd_destroy": (C.c_int, [_vp]),
    "sc_qd_ctx": (_vp, [_vp]),
    "sc_qd_set_state": (C.c_int, [_vp, _dp, _dp, _dp]),
    "sc_qd_get_state": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "sc_qd_get_u": (C.c_int, [_vp, _dp]),
    "sc_qd_ddt": (C.c_int, [_vp, C.c_double, _dp, _dp]),
    "sc_qd_integrate": (C.c_int, [_vp, C.POINTER(QdRun), _dp, C.POINTER(C.c_int), _dp, C.c_int]),
    "sc_qd_get_counters": (C.c_int, [_vp, C.POINTER(QdCounters)]),
    "sc_qd_file_run": (C.c_int, [_vp, C.POINTER(QdRun)]),
    "sc_qd_solve_ss": (C.c_int, [_vp, _dp, C.c_int, _dp]),
    "sc_qd_set_output": (C.c_int, [_vp, C.POINTER(QdOutput)]),
    "sc_fd_set_output": (C.c_int, [_vp, C.POINTER(QdOutput)]),
    "sc_fd_integrate": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_double, _dp, C.POINTER(C.c_int)]),
    "sc_qdfd_create": (C.c_int, [C.POINTER(QdFdDesc)] + [_dp] * 13 + [C.POINTER(_vp)]),
    "sc_qdfd_destroy": (C.c_int, [_vp]),
    "sc_qdfd_time_step": (C.c_int, [_vp, _dp]),
    "sc_qdfd_set_state": (C.c_int, [_vp, _dp, _dp, _dp]),
    "sc_qdfd_get_state": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]),
    "sc_qdfd_integrate": (C.c_int, [_vp, C.POINTER(QdFdRun), _dp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(QdFdPhase), C.c_int,
                                    C.POINTER(C.c_int)]),
    "sc_qd_load_checkpoint": (C.c_int, [_vp, C.POINTER(QdRun), C.POINTER(C.c_int)]),
    "sc_in_get": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
    "sc_in_get_vector": (C.c_int, [C.c_char_p, C.c_char_p, _dp, C.c_int, C.POINTER(C.c_int)]),
    "sc_vec_count": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
    "sc_vec_read": (C.c_int, [C.c_char_p, C.c_int, _dp, C.c_int, C.POINTER(C.c_int)]),
    "sc_vec_write": (C.c_int, [C.c_char_p, _dp, C.c_int, C.c_int]),
}

_lib = None


def load():
    """Load libscycle_hip.so; raises ScError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ScError(f"libscycle_hip.so is not built ({LIB_PATH}); run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != 0:
        msg = load().sc_last_error()
        raise ScError(f"scycle_hip error {code}: {msg.decode() if msg else ''}")


def dptr(a):
    """double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    import numpy as np
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)
