"""sc_solve_noalloc (the allocation-free FP64 Jacobi-PCG) against sc_solve,
the entry point it mirrors: on the same inputs it runs the same kernels, so
the solution, iteration count and residual are the same bit for bit.  Two
calls in a row, the second on a smaller grid (the shared scratch keeps the
first call's contents); options it does not serve and grids larger than its
scratch are refused before any work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(Ny, Nz, Ly, Lz, bz):
    q = np.repeat(np.arange(Ny) / (Ny - 1.0), Nz)
    r = np.tile(np.arange(Nz) / (Nz - 1.0), Ny)
    return Ly * q, Lz * np.sinh(bz * r) / np.sinh(bz)


def _case(Ny, Nz, seed):
    from scycle_amd import SbpOps
    rng = np.random.default_rng(seed)
    y, z = _grid(Ny, Nz, 30.0, 30.0, 4.0)
    mu = 30 + 6 * rng.random(Ny * Nz)
    g = SbpOps(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=(0, 1, 0, 1))
    rhs = g.setRhs(rng.standard_normal(Nz), np.full(Nz, 0.3), np.zeros(Ny), np.zeros(Ny))
    return g, rhs


def _noalloc(g, rhs, rtol=1e-12, mode=0, precond=0, reuse=0):
    from scycle_amd import DeviceArray
    from scycle_amd._lib import SolveOpts, SolveStats, load
    b = DeviceArray(g.N, host=rhs)
    u = DeviceArray(g.N, host=np.zeros(g.N))
    o = SolveOpts(mode, rtol, 1e-5, 200000, 50, 64, precond, reuse)
    st = SolveStats()
    rc = load().sc_solve_noalloc(g.h, b.ptr, u.ptr, C.byref(o), C.byref(st))
    g.sync()
    return rc, u.to_host(), dict(iters=st.iters, converged=st.converged, relres=st.relres)


def test_noalloc_solve_matches_sc_solve():
    for Ny, Nz, seed in ((121, 117, 1), (45, 41, 2)):
        print("noalloc case", Ny, Nz, flush=True)
        g, rhs = _case(Ny, Nz, seed)
        uref, sref = g.solve(rhs, mode="cg64", rtol=1e-12, precond="jacobi")
        rc, u, st = _noalloc(g, rhs)
        print(Ny, Nz, "sc_solve", sref, "sc_solve_noalloc", st, flush=True)
        assert rc == 0 and st["converged"] == 1
        assert st["iters"] == sref["iters"] and st["relres"] == sref["relres"]
        assert np.array_equal(u, uref)
        # the context's own solve still works after the allocation-free one
        u2, s2 = g.solve(rhs, mode="cg64", rtol=1e-12, precond="jacobi")
        assert np.array_equal(u2, uref) and s2["iters"] == sref["iters"]


def test_noalloc_solve_refuses_what_it_cannot_serve():
    from scycle_amd import SbpOps
    from scycle_amd._lib import load
    g, rhs = _case(33, 29, 3)
    for kw in (dict(mode=1), dict(precond=1), dict(reuse=2), dict(rtol=0.0)):
        rc, _, _ = _noalloc(g, rhs, **kw)
        assert rc == 1, kw
        print(kw, load().sc_last_error().decode(), flush=True)
    # 1025^2 = 1050625 nodes > 2^20: refused before any kernel runs
    n = 1025
    big = SbpOps(2, n, n, 30.0, 30.0, np.full(n * n, 36.0), var_grid=False, bc=(0, 1, 0, 1))
    rc, _, _ = _noalloc(big, np.zeros(n * n))
    assert rc == 1
    assert b"more than the allocation-free scratch" in load().sc_last_error()
