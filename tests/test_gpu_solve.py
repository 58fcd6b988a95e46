"""GPU linear solve (Jacobi-PCG FP64, FP32-inner iterative refinement) against
the oracle's banded Cholesky (the stand-in for the reference's MUMPS
Cholesky, linearElastic.cpp:251-259).  Tolerance: relative residual 1e-12 is
requested; the solution then agrees with the direct solve to 1e-9 relative
(condition number of the 4th-order operator at these sizes is < 1e4)."""
import numpy as np
import pytest

from oracle_lib import OracleSbp

pytestmark = pytest.mark.gpu


def _grid(Ny, Nz, Ly, Lz, bz):
    q = np.repeat(np.arange(Ny) / (Ny - 1.0), Nz)
    r = np.tile(np.arange(Nz) / (Nz - 1.0), Ny)
    return Ly * q, Lz * np.sinh(bz * r) / np.sinh(bz)


@pytest.mark.parametrize("mode", ["cg64", "ir32"])
@pytest.mark.parametrize("order,var_grid,Ny,Nz", [(2, False, 41, 41), (4, True, 51, 47),
                                                  (4, True, 101, 101)])
def test_solve_matches_direct(mode, order, var_grid, Ny, Nz):
    from scycle_amd import SbpOps
    rng = np.random.default_rng(Ny + order)
    y, z = _grid(Ny, Nz, 30.0, 30.0, 5.0) if var_grid else (None, None)
    mu = 30 + 6 * rng.random(Ny * Nz)
    bc = (0, 1, 0, 1)
    o = OracleSbp(order, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=var_grid, bc=bc)
    g = SbpOps(order, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=var_grid, bc=bc)
    bcL = rng.standard_normal(Nz)
    bcR = np.full(Nz, 0.3)
    rhs = o.set_rhs(bcL, bcR, np.zeros(Ny), np.zeros(Ny))
    uref = o.chol_solve(rhs)
    u, st = g.solve(rhs, mode=mode, rtol=1e-12)
    print(mode, order, Ny, Nz, st, flush=True)
    assert st["converged"] and st["relres"] <= 1e-12
    assert np.abs(u - uref).max() <= 1e-9 * np.abs(uref).max()
    # warm start from the solution converges immediately
    u2, st2 = g.solve(rhs, u0=u, mode=mode, rtol=1e-12)
    assert st2["iters"] == 0


@pytest.mark.parametrize("precond", ["fdiag64", "fdiag32", "fdiag_bf16"])
@pytest.mark.parametrize("layered", [True, False])
def test_fdiag_preconditioned_solve(precond, layered):
    """CG preconditioned with the fast-diagonalisation inverse of the separable
    operator: same solution as the direct solve, in a few iterations when mu
    depends on depth only (layered) and still converging for general mu."""
    from scycle_amd import SbpOps
    Ny, Nz = 81, 73
    y, z = _grid(Ny, Nz, 30.0, 30.0, 5.0)
    rng = np.random.default_rng(3)
    if layered:
        t = np.clip(z / 4.0, 0, 1)
        mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    else:
        mu = 30 + 6 * rng.random(Ny * Nz)
    bc = (0, 1, 0, 1)
    o = OracleSbp(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    g = SbpOps(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    rhs = o.set_rhs(rng.standard_normal(Nz), np.full(Nz, 0.3), np.zeros(Ny), np.zeros(Ny))
    uref = o.chol_solve(rhs)
    u, st = g.solve(rhs, rtol=1e-12, precond=precond, max_iters=2000)
    print(precond, layered, st, flush=True)
    assert st["converged"] and st["relres"] <= 1e-12
    assert np.abs(u - uref).max() <= 1e-9 * np.abs(uref).max()
    if layered:
        assert st["iters"] <= (12 if precond != "fdiag_bf16" else 30)


def test_solve_full_size_residual():
    """1001^2 C2 operator: the FP64 residual of both solvers meets 1e-12 and
    the two solutions agree to the conditioning-limited 1e-8."""
    from scycle_amd import SbpOps
    N = 1001
    y, z = _grid(N, N, 30.0, 30.0, 5.0)
    t = np.clip(z / 4.0, 0, 1)
    mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    g = SbpOps(4, N, N, 30.0, 30.0, mu, y, z, var_grid=True, bc=(0, 1, 0, 1))
    rng = np.random.default_rng(7)
    rhs = g.setRhs(1e-3 * rng.random(N), np.full(N, 5e-4), np.zeros(N), np.zeros(N))
    u1, s1 = g.solve(rhs, mode="cg64", rtol=1e-10)
    u2, s2 = g.solve(rhs, mode="ir32", rtol=1e-10)
    print("cg64", s1, "ir32", s2, flush=True)
    assert s1["converged"] and s2["converged"]
    assert np.abs(u1 - u2).max() <= 1e-7 * np.abs(u1).max()


@pytest.mark.parametrize("precond", ["fdiag32", "fdiag_bf16"])
@pytest.mark.parametrize("layered", [True, False])
def test_ir32_fdiag_solve(precond, layered):
    """FP32 flexible PCG with the fast-diagonalisation preconditioner (FP32 or
    bf16 GEMMs) inside FP64 iterative refinement: the FP64 residual meets
    1e-12 and the solution matches the direct solve to 1e-9."""
    from scycle_amd import SbpOps
    Ny, Nz = 81, 73
    y, z = _grid(Ny, Nz, 30.0, 30.0, 5.0)
    rng = np.random.default_rng(5)
    if layered:
        t = np.clip(z / 4.0, 0, 1)
        mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    else:
        mu = 30 + 6 * rng.random(Ny * Nz)
    bc = (0, 1, 0, 1)
    o = OracleSbp(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    g = SbpOps(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    rhs = o.set_rhs(rng.standard_normal(Nz), np.full(Nz, 0.3), np.zeros(Ny), np.zeros(Ny))
    uref = o.chol_solve(rhs)
    u, st = g.solve(rhs, mode="ir32", rtol=1e-12, inner_rtol=1e-5, precond=precond, max_iters=4000)
    print("ir32", precond, layered, st, flush=True)
    assert st["converged"] and st["relres"] <= 1e-12
    assert np.abs(u - uref).max() <= 1e-9 * np.abs(uref).max()
    if layered:
        assert st["outer"] <= 4 and st["iters"] <= (16 if precond == "fdiag32" else 40)
    # warm start from the solution: no sweep
    u2, st2 = g.solve(rhs, u0=u, mode="ir32", rtol=1e-12, precond=precond)
    assert st2["iters"] == 0 and st2["outer"] == 0


def test_ir32_fdiag_full_size():
    """1001^2 C2 operator: IR32 with the FP32 and bf16 fast-diagonalisation
    inner solves meets the FP64 residual 1e-12 and agrees with FP64 CG with
    the FP64 preconditioner to the conditioning-limited 1e-8."""
    from scycle_amd import SbpOps
    N = 1001
    y, z = _grid(N, N, 30.0, 30.0, 5.0)
    t = np.clip(z / 4.0, 0, 1)
    mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    g = SbpOps(4, N, N, 30.0, 30.0, mu, y, z, var_grid=True, bc=(0, 1, 0, 1))
    rng = np.random.default_rng(9)
    rhs = g.setRhs(1e-3 * rng.random(N), np.full(N, 5e-4), np.zeros(N), np.zeros(N))
    u0, s0 = g.solve(rhs, mode="cg64", rtol=1e-12, precond="fdiag64")
    for precond in ("fdiag32", "fdiag_bf16"):
        u, st = g.solve(rhs, mode="ir32", rtol=1e-12, precond=precond)
        print("1001^2 ir32", precond, st, "cg64 fdiag64", s0, flush=True)
        assert st["converged"] and st["relres"] <= 1e-12
        assert np.abs(u - u0).max() <= 1e-8 * np.abs(u0).max()


@pytest.mark.parametrize("mode,precond", [("cg64", "jacobi"), ("cg64", "fdiag64"), ("cg64", "fdiag32"),
                                          ("ir32", "jacobi"), ("ir32", "fdiag32"), ("ir32", "fdiag_bf16")])
def test_reuse_projection_guess(mode, precond):
    """Initial guess projected on the previous solutions of the same operator
    (Fischer's method): every solution still matches the direct solve, a
    right-hand side in the span of earlier ones needs no iteration, and a BC
    type change resets the basis."""
    from scycle_amd import SbpOps
    Ny, Nz = 61, 57
    y, z = _grid(Ny, Nz, 30.0, 30.0, 5.0)
    rng = np.random.default_rng(11)
    mu = 30 + 6 * rng.random(Ny * Nz)
    bc = (0, 1, 0, 1)
    o = OracleSbp(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    g = SbpOps(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=bc)
    rhss = [o.set_rhs(rng.standard_normal(Nz), rng.standard_normal(Nz), np.zeros(Ny), np.zeros(Ny))
            for _ in range(3)]
    rhss.append(0.7 * rhss[0] - 1.3 * rhss[1] + 0.2 * rhss[2])
    its = []
    # the basis solutions to 1e-13, so that the combination of them meets 1e-10
    # before any iteration
    for b, rtol in zip(rhss, (1e-13, 1e-13, 1e-13, 1e-10)):
        uref = o.chol_solve(b)
        u, st = g.solve(b, mode=mode, rtol=rtol, precond=precond, reuse=3, max_iters=20000)
        assert st["converged"] and st["relres"] <= rtol
        assert np.abs(u - uref).max() <= 1e-8 * np.abs(uref).max()
        its.append(st["iters"])
    print(mode, precond, "iterations", its, flush=True)
    assert its[3] == 0
    # two more solutions replace the two oldest basis vectors (capacity 3);
    # the guesses from the mixed basis still lead to the direct solution
    for _ in range(2):
        rhss.append(o.set_rhs(rng.standard_normal(Nz), rng.standard_normal(Nz), np.zeros(Ny), np.zeros(Ny)))
        u, st = g.solve(rhss[-1], mode=mode, rtol=1e-13, precond=precond, reuse=3, max_iters=20000)
        assert st["converged"] and st["iters"] > 0
    for b in (0.4 * rhss[2] + 1.1 * rhss[4] - 0.6 * rhss[5], 0.5 * rhss[0] + 0.3 * rhss[5], rhss[5]):
        u, st = g.solve(b, mode=mode, rtol=1e-10, precond=precond, reuse=3, max_iters=20000)
        uref = o.chol_solve(b)
        assert st["converged"] and st["relres"] <= 1e-10
        assert np.abs(u - uref).max() <= 1e-8 * np.abs(uref).max()
        print(mode, precond, "after eviction, iterations", st["iters"], flush=True)
    # the basis belongs to the operator: after a BC type change it starts afresh
    g.changeBCTypes(0, 1, 1, 1)
    o2 = OracleSbp(4, Ny, Nz, 30.0, 30.0, mu, y, z, var_grid=True, bc=(0, 1, 1, 1))
    b = o2.set_rhs(rng.standard_normal(Nz), rng.standard_normal(Nz), np.zeros(Ny), np.zeros(Ny))
    u, st = g.solve(b, mode=mode, rtol=1e-11, precond=precond, reuse=3, max_iters=20000)
    uref = o2.chol_solve(b)
    assert np.abs(u - uref).max() <= 1e-8 * np.abs(uref).max()


@pytest.mark.parametrize("Ny,Nz,var_grid,bc", [(37, 53, True, (0, 1, 0, 1)), (64, 128, False, (0, 1, 0, 1)),
                                               (130, 70, True, (1, 1, 0, 0)), (21, 21, False, (0, 0, 0, 0))])
def test_fdiag_apply_is_the_separable_inverse(Ny, Nz, var_grid, bc):
    """With a constant shear modulus the operator is separable (the two-point
    mu average is mu itself), so the fast-diagonalisation preconditioner is
    A^-1: z = M^-1 r must match the oracle's direct solve of A z = r.  Sizes
    off and on the GEMM tiles' 64 (padding, several tiles, Ny != Nz), both
    grids and BC sets.  Tolerances: FP64 GEMMs 1e-10; FP32 and bf16 GEMMs by
    their unit roundoff (6e-8, 4e-3) times the operator's condition number
    (< 1e4 at these sizes, the error entering through the inverse eigenvalue
    scaling)."""
    from scycle_amd import SbpOps
    y, z = _grid(Ny, Nz, 30.0, 24.0, 5.0) if var_grid else (None, None)
    mu = np.full(Ny * Nz, 32.0)
    o = OracleSbp(4, Ny, Nz, 30.0, 24.0, mu, y, z, var_grid=var_grid, bc=bc)
    g = SbpOps(4, Ny, Nz, 30.0, 24.0, mu, y, z, var_grid=var_grid, bc=bc)
    rng = np.random.default_rng(Ny * 1000 + Nz)
    r = rng.standard_normal(Ny * Nz)
    zref = o.chol_solve(r)
    tol = {"fdiag64": 1e-10, "fdiag32": 5e-4, "fdiag_bf16": 0.5}
    for prec in ("fdiag64", "fdiag32", "fdiag_bf16"):
        zz = g.pc_apply(r, prec)
        err = np.abs(zz - zref).max() / np.abs(zref).max()
        print(Ny, Nz, var_grid, bc, prec, f"rel err vs direct solve {err:.3e}", flush=True)
        assert err <= tol[prec], (prec, err)
    # Jacobi: r / diag(A)
    zj = g.pc_apply(r, "jacobi")
    assert np.allclose(zj, r / o.diag(), rtol=1e-13, atol=0)
