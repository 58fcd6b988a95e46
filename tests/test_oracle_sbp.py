"""Pins the CPU oracle's SBP operators to what the mathematics fixes.

The reference ships no numeric golden vectors for the operators, so each
check below is chosen so that a plausible transcription mistake in the oracle
(a wrong closure coefficient, sign, transposed factor, index) fails it:
SBP identities in exact form, polynomial exactness, symmetry and
definiteness of the assembled 2-D operator, its Neumann null space, the
compact closed form of the 2nd-order interior stencil (the same formula as
sfsbp/FDP.cpp:800-803), const-grid == var-grid on a uniform grid, and
convergence rates on the reference's manufactured solution
(linearElastic.cpp:1051-1128).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle_lib import DIRICHLET, NEUMANN, OracleSbp, sbp1d


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("N", [9, 21, 40])
def test_sbp_identity_and_exactness(order, N):
    if order == 4 and N <= 8:
        pytest.skip()
    h = 1.0 / (N - 1)
    m = sbp1d(order, N, 1.0 / h)
    H, D1 = m["H"], m["D1"]
    Q = H @ D1
    B = np.zeros((N, N))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    # Q + Q^T = B (summation by parts)
    assert np.abs(Q + Q.T - B).max() < 1e-13
    # H is a positive diagonal quadrature, exact for low-degree polynomials
    assert np.all(np.diag(H) > 0) and np.count_nonzero(H - np.diag(np.diag(H))) == 0
    x = np.linspace(0.0, 1.0, N)
    for p in range(order):  # 2p-th order quadrature; check degrees < order
        assert abs(np.diag(H) @ x ** p - 1.0 / (p + 1)) < 1e-13
    # D1 exactness: boundary closure order order/2, interior order
    bnd = order // 2
    for p in range(order + 1):
        ex = p * x ** (p - 1) if p > 0 else 0 * x
        err = np.abs(D1 @ x ** p - ex)
        if p <= bnd:
            assert err.max() < 1e-10
        inner = err[order + 1:-(order + 1)]
        if p <= order and inner.size:
            assert inner.max() < 1e-9
    # Hinv * H = I
    assert np.abs(m["Hinv"] @ H - np.eye(N)).max() < 1e-13
    # fully compatible boundary derivative: BS rows are the outward derivative
    BS = m["BS"]
    assert np.abs(BS[0] + D1[0]).max() < 1e-12 and np.abs(BS[-1] - D1[-1]).max() < 1e-12


def test_sbp4_remainder_structure():
    N = 21
    m = sbp1d(4, N, float(N - 1))
    D3, D4, C3, C4 = m["D3"], m["D4"], m["C3"], m["C4"]
    # D3 annihilates quadratics, D4 annihilates cubics (difference operators)
    x = np.arange(N, dtype=float)
    for p in range(3):
        assert np.abs(D3 @ x ** p).max() < 1e-7
    for p in range(4):
        assert np.abs(D4 @ x ** p).max() < 1e-6
    # interior D3 row is the 3rd difference, D4 the 4th
    assert np.allclose(D3[8, 7:11], [-1, 3, -3, 1])
    assert np.allclose(D4[8, 6:11], [1, -4, 6, -4, 1])
    # C weights vanish on the outermost rows and row N-4 of D3 is empty
    assert C3[0, 0] == C3[1, 1] == C3[-1, -1] == C3[-2, -2] == C3[N - 4, N - 4] == 0
    assert C4[0, 0] == C4[1, 1] == C4[-1, -1] == C4[-2, -2] == 0
    assert np.count_nonzero(D3[N - 4]) == 0


def _grid(Ny, Nz, Ly, Lz, by=0.0, bz=0.0):
    q = np.repeat(np.linspace(0, 1, Ny), Nz)
    r = np.tile(np.linspace(0, 1, Nz), Ny)
    y = Ly * (np.sinh(by * q) / np.sinh(by) if by else q)
    z = Lz * (np.sinh(bz * r) / np.sinh(bz) if bz else r)
    return y, z


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("var_grid", [False, True])
@pytest.mark.parametrize("bc", [(0, 1, 0, 1), (0, 0, 0, 0), (1, 0, 0, 1)])
def test_operator_symmetric_definite(order, var_grid, bc):
    Ny, Nz = 19, 23
    rng = np.random.default_rng(1)
    mu = 30 + 6 * rng.random(Ny * Nz)
    y, z = _grid(Ny, Nz, 30, 30, 0, 2.0) if var_grid else (None, None)
    o = OracleSbp(order, Ny, Nz, 30, 30, mu, y, z, var_grid=var_grid, bc=bc)
    A = o.csr()
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()
    lam = sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0]
    assert lam < 0  # negative definite with at least one Dirichlet side


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("var_grid", [False, True])
def test_neumann_nullspace(order, var_grid):
    Ny, Nz = 17, 21
    rng = np.random.default_rng(2)
    mu = 30 + 6 * rng.random(Ny * Nz)
    y, z = _grid(Ny, Nz, 30, 20, 1.5, 3.0) if var_grid else (None, None)
    o = OracleSbp(order, Ny, Nz, 30, 20, mu, y, z, var_grid=var_grid, bc=(1, 1, 1, 1))
    Au = o.apply(np.ones(Ny * Nz))
    assert np.abs(Au).max() < 1e-12 * np.abs(o.csr()).max()


def test_order2_interior_closed_form():
    """Interior rows of H^-1 A (order 2, const grid) equal the compact
    variable-coefficient stencil 1/(2h^2)[(mu_j+mu_{j-1})u_{j-1}
    - (mu_{j-1}+2mu_j+mu_{j+1})u_j + (mu_j+mu_{j+1})u_{j+1}] in each direction."""
    Ny, Nz, Ly, Lz = 13, 15, 3.0, 4.0
    rng = np.random.default_rng(3)
    mu = 1 + rng.random(Ny * Nz)
    u = rng.standard_normal(Ny * Nz)
    o = OracleSbp(2, Ny, Nz, Ly, Lz, mu, var_grid=False)
    D2u = o.Hinv(o.apply(u)).reshape(Ny, Nz)
    M, U = mu.reshape(Ny, Nz), u.reshape(Ny, Nz)
    hy, hz = Ly / (Ny - 1), Lz / (Nz - 1)
    for j in range(2, Ny - 2):
        for k in range(2, Nz - 2):
            fy = ((M[j, k] + M[j - 1, k]) * U[j - 1, k] - (M[j - 1, k] + 2 * M[j, k] + M[j + 1, k]) * U[j, k]
                  + (M[j, k] + M[j + 1, k]) * U[j + 1, k]) / (2 * hy ** 2)
            fz = ((M[j, k] + M[j, k - 1]) * U[j, k - 1] - (M[j, k - 1] + 2 * M[j, k] + M[j, k + 1]) * U[j, k]
                  + (M[j, k] + M[j, k + 1]) * U[j, k + 1]) / (2 * hz ** 2)
            assert abs(D2u[j, k] - (fy + fz)) < 1e-10 * (abs(fy) + abs(fz) + 1)


@pytest.mark.parametrize("order", [2, 4])
def test_uniform_var_grid_equals_const_grid(order):
    Ny, Nz = 17, 19
    rng = np.random.default_rng(4)
    mu = np.repeat(30 + 6 * rng.random(Ny), Nz)  # varies in y only
    u = rng.standard_normal(Ny * Nz)
    y, z = _grid(Ny, Nz, 30.0, 25.0)
    oc = OracleSbp(order, Ny, Nz, 30.0, 25.0, mu, var_grid=False)
    ov = OracleSbp(order, Ny, Nz, 30.0, 25.0, mu, y, z, var_grid=True)
    a, b = oc.apply(u), ov.apply(u)
    assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()


def _mms(order, var_grid, Ny, bc):
    """Solve div(mu grad u) = s with the reference's manufactured solution."""
    Nz = Ny
    Ly = Lz = 2.0
    y, z = _grid(Ny, Nz, Ly, Lz, 1.0 if var_grid else 0, 2.0 if var_grid else 0)
    mu = np.sin(y) * np.sin(z) + 30
    uex = np.cos(y) * np.sin(z)
    mu_y, mu_z = np.cos(y) * np.sin(z), np.sin(y) * np.cos(z)
    u_y, u_z = -np.sin(y) * np.sin(z), np.cos(y) * np.cos(z)
    s = mu * (-2 * np.cos(y) * np.sin(z)) + mu_y * u_y + mu_z * u_z
    o = OracleSbp(order, Ny, Nz, Ly, Lz, mu, y if var_grid else None, z if var_grid else None,
                  var_grid=var_grid, bc=bc)
    Y, Z = y.reshape(Ny, Nz), z.reshape(Ny, Nz)
    UE, MU = uex.reshape(Ny, Nz), mu.reshape(Ny, Nz)
    UY, UZ = u_y.reshape(Ny, Nz), u_z.reshape(Ny, Nz)
    g = {}
    g["R"] = UE[-1, :] if bc[0] == DIRICHLET else (MU * UY)[-1, :]
    g["T"] = UE[:, 0] if bc[1] == DIRICHLET else (MU * UZ)[:, 0]
    g["L"] = UE[0, :] if bc[2] == DIRICHLET else (MU * UY)[0, :]
    g["B"] = UE[:, -1] if bc[3] == DIRICHLET else (MU * UZ)[:, -1]
    rhs = o.set_rhs(g["L"], g["R"], g["T"], g["B"])
    if var_grid:
        yq, zr = o.metric()
        Hs = o.H(s * yq * zr)  # H acts on logical coordinates; J = yq*zr
    else:
        Hs = o.H(s)
    u = o.chol_solve(rhs + Hs)
    e = u - uex
    w = o.H(np.ones(Ny * Nz)) * (o.metric()[0] * o.metric()[1] if var_grid else 1)
    return np.sqrt(np.sum(w * e * e))


@pytest.mark.parametrize("order,var_grid,rate", [(2, False, 1.8), (4, False, 2.7), (2, True, 1.8),
                                                 (4, True, 2.7)])
def test_mms_convergence(order, var_grid, rate):
    bc = (DIRICHLET, NEUMANN, DIRICHLET, NEUMANN)
    e1 = _mms(order, var_grid, 21, bc)
    e2 = _mms(order, var_grid, 41, bc)
    assert np.log2(e1 / e2) > rate, (e1, e2)


def test_mu3_quirk_uses_global_neighbour():
    """The 4th-order y-remainder averages mu with global index I+1 (the
    z-neighbour), sbpOps_m_varGrid.cpp:1428-1448: a mu that varies only in z
    changes the y-part; the corrected y-average would not depend on it."""
    Ny, Nz = 17, 17
    rng = np.random.default_rng(5)
    u = rng.standard_normal(Ny * Nz)
    mu_a = np.full(Ny * Nz, 30.0)
    mu_b = np.tile(30 + np.arange(Nz, dtype=float), Ny)
    # laplace "y" only, all Neumann in y to isolate the interior y-operator
    oa = OracleSbp(4, Ny, Nz, 1, 1, mu_a, var_grid=False, bc=(1, 1, 1, 1), laplace=1)
    ob = OracleSbp(4, Ny, Nz, 1, 1, mu_b, var_grid=False, bc=(1, 1, 1, 1), laplace=1)
    ra = oa.apply(u).reshape(Ny, Nz) / 30.0
    rb = ob.apply(u).reshape(Ny, Nz) / mu_b.reshape(Ny, Nz)
    # with mu(z) the y-operator on column k is mu_k*Y1 + ((mu_k+mu_{k+1})/2)*Y3; dividing by mu_k
    # leaves a column-dependent remainder exactly when the quirk is present
    assert np.abs(ra[:, 3] - rb[:, 3]).max() > 1e-6
