"""Synthetic inputs of the benchmark configurations (quasi-dynamic strike-slip
cycle on a 2-D antiplane domain), built the way the reference builds them from
a .in deck: depth profiles are piecewise-linear through (vals, depths) knots
with linear extrapolation past the last knot (genFuncs.cpp:574-617 setVec).

Units: km, GPa, g/cm^3, MPa, m, m/s, s.
"""
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n uniforms in [0,1) from SplitMix64 (53-bit mantissa)."""
    out = np.empty(n)
    x = seed & MASK64
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def set_vec(coord, vals, depths):
    """Reference setVec: piecewise-linear interpolation, last segment extended."""
    coord = np.asarray(coord, dtype=np.float64)
    out = np.full(coord.shape, float(vals[0]))
    if coord.size == 1:
        return out
    n = len(depths)
    for ind in range(n - 1):
        z0, z1, v0, v1 = depths[ind], depths[ind + 1], vals[ind], vals[ind + 1]
        inside = (coord >= z0) & (coord <= z1)
        beyond = (coord > z1) & (ind == n - 2)
        m = inside | beyond
        out[m] = (v1 - v0) / (z1 - z0) * (coord[m] - z0) + v0
    return out


@dataclass
class QdConfig:
    name: str
    Ny: int
    Nz: int
    Ly: float
    Lz: float
    order: int
    var_grid: bool
    mu: np.ndarray
    y: np.ndarray
    z: np.ndarray
    a: np.ndarray
    b: np.ndarray
    Dc: np.ndarray
    sN: np.ndarray
    eta: np.ndarray
    psi0: np.ndarray
    slip0: np.ndarray
    locked: np.ndarray = None
    prestress: np.ndarray = None
    bcRShift: np.ndarray = None
    mat_bc: tuple = (0, 1, 0, 1)          # R, T, L, B: remoteLoading, freeSurface, symmFault, freeSurface
    remote_loading_R: bool = True
    fault_type_scale: float = 2.0
    vL: float = 1e-9
    f0: float = 0.6
    v0: float = 1e-6
    root_tol: float = 1e-12
    state_law: int = 0                      # 0 aging, 1 slip
    steps: int = 2000
    integrator: str = "RK32"
    tol: float = 1e-8
    min_dt: float = 0.0
    init_dt: float = 1e-3
    extra: dict = field(default_factory=dict)


def _grid(Ny, Nz, Ly, Lz, zstretch=None):
    q = np.repeat(np.arange(Ny) * (1.0 / (Ny - 1.0)), Nz)
    r = np.tile(np.arange(Nz) * (1.0 / (Nz - 1.0)), Ny)
    y = Ly * q
    if zstretch:
        z = Lz * np.sinh(zstretch * r) / np.sinh(zstretch)
    else:
        z = Lz * r
    return y, z


def strike_slip(Ny, Nz, order=2, var_grid=False, zstretch=None, layered_mu=False,
                Ly=30.0, Lz=30.0, seed=0x5C7C1E, name="custom", steps=2000):
    """The C1/C2 physics: aging law, f0=0.6, v0=1e-6, b=0.02, Dc=8mm, sN=50 MPa,
    a = 0.015 above 12 km ramping to 0.025 at 17 km and below; vL = 1e-9;
    psi0 = psi_ss(vL)*(1 + eps), eps uniform in [-1%, 1%] from SplitMix64."""
    y, z = _grid(Ny, Nz, Ly, Lz, zstretch)
    zf = z[:Nz]
    if layered_mu:
        t = np.clip(z / 4.0, 0.0, 1.0)
        mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    else:
        mu = np.full(Ny * Nz, 36.0)
    rho = np.full(Nz, 3.0)
    mu_f = mu[:Nz]
    a = set_vec(zf, [0.015, 0.015, 0.025, 0.025], [0, 12, 17, 30])
    b = np.full(Nz, 0.02)
    Dc = np.full(Nz, 8e-3)
    sN = np.full(Nz, 50.0)
    fts = 2.0
    eta = np.sqrt(mu_f * rho) / fts
    f0, v0, vL = 0.6, 1e-6, 1e-9
    psi_ss = f0 - b * np.log(vL / v0)
    eps = 0.02 * splitmix64(seed, Nz) - 0.01
    psi0 = psi_ss * (1.0 + eps)
    return QdConfig(name=name, Ny=Ny, Nz=Nz, Ly=Ly, Lz=Lz, order=order, var_grid=var_grid,
                    mu=mu, y=y, z=z, a=a, b=b, Dc=Dc, sN=sN, eta=eta, psi0=psi0,
                    slip0=np.zeros(Nz), steps=steps)


def config_c1(Ny=101, Nz=101):
    """101^2, order 2, constant grid, constant mu = 36 GPa."""
    return strike_slip(Ny, Nz, order=2, var_grid=False, name=f"C1-{Ny}x{Nz}", steps=2000)


def config_c2(Ny=1001, Nz=1001):
    """1001^2, order 4, sinh-stretched z (b = 5), layered mu 30 -> 36 GPa over 4 km."""
    return strike_slip(Ny, Nz, order=4, var_grid=True, zstretch=5.0, layered_mu=True,
                       name=f"C2-{Ny}x{Nz}", steps=500)


@dataclass
class FdConfig:
    name: str
    Ny: int
    Nz: int
    Ly: float
    Lz: float
    order: int
    var_grid: bool
    mu: np.ndarray
    rho: np.ndarray
    cs: np.ndarray
    y: np.ndarray
    z: np.ndarray
    a: np.ndarray
    b: np.ndarray
    Dc: np.ndarray
    sNEff: np.ndarray
    tau0: np.ndarray
    prestress: np.ndarray
    slip0: np.ndarray
    psi0: np.ndarray
    dt: float
    locked: np.ndarray = None
    outR: bool = True                       # outGoingCharacteristics (strikeSlip_linearElastic_fd.cpp:21)
    outT: bool = False                      # freeSurface
    outB: bool = True
    faultL: bool = True                     # symmFault
    fault_type_scale: float = 2.0
    vL: float = 1e-9
    f0: float = 0.6
    v0: float = 1e-6
    root_tol: float = 1e-12
    state_law: int = 1                      # slip law
    steps: int = 2000


# Time step of the FD runs: dt = CFL * g_cfl * min(dy, dz) / cs with
# g_cfl = 0.7071/sqrt(1.4498) for the 4th-order operator.  The reference's
# computeTimeStep (strikeSlip_linearElastic_fd.cpp:521-575) multiplies g_cfl in
# twice (max_deltaT already holds it); this keeps it once.
G_CFL4 = 0.7071 / np.sqrt(1.4498)


def fd_rupture(Ny, Nz, order=4, Ly=30.0, Lz=30.0, cfl=0.5, name="custom", steps=2000):
    """The C4 physics: constant grid, mu = 36 GPa, rho = 3 (cs = 3.4641 km/s),
    slip law with f0 = 0.6, v0 = 1e-6, b = 0.02, Dc = 8 mm, sNEff = 50 MPa;
    a = 0.015 on the nucleation patch z in [10, 14] km (a - b = -0.005) and
    0.025 elsewhere; tau0 = 30 + 10 exp(-(z - 12)^2 / 2) MPa (the Heaviside
    form of Fault_fd::updatePrestress, fault.cpp:1418-1421); the fault starts
    locked-in at steady state for V = 1e-9 m/s with u = slip = 0."""
    y, z = _grid(Ny, Nz, Ly, Lz)
    zf = z[:Nz]
    mu = np.full(Ny * Nz, 36.0)
    rho = np.full(Ny * Nz, 3.0)
    cs = np.sqrt(mu / rho)
    a = np.where((zf >= 10.0) & (zf <= 14.0), 0.015, 0.025)
    b = np.full(Nz, 0.02)
    Dc = np.full(Nz, 8e-3)
    sNEff = np.full(Nz, 50.0)
    tau0 = 30.0 + 10.0 * np.exp(-((zf - 12.0) ** 2) / 2.0)
    f0, v0 = 0.6, 1e-6
    psi0 = f0 - b * np.log(1e-9 / v0)
    dy, dz = Ly / (Ny - 1.0), Lz / (Nz - 1.0)
    dt = cfl * G_CFL4 * min(dy, dz) / cs.max()
    return FdConfig(name=name, Ny=Ny, Nz=Nz, Ly=Ly, Lz=Lz, order=order, var_grid=False, mu=mu,
                    rho=rho, cs=cs, y=None, z=None, a=a, b=b, Dc=Dc, sNEff=sNEff, tau0=tau0,
                    prestress=np.zeros(Nz), slip0=np.zeros(Nz), psi0=psi0, dt=dt, steps=steps)


def config_c4(Ny=2001, Nz=2001):
    """2001^2 fully dynamic rupture, order 4, constant grid, CFL 0.5, 2000 steps."""
    return fd_rupture(Ny, Nz, name=f"C4-{Ny}x{Nz}", steps=2000)


@dataclass
class PowerLawConfig(QdConfig):
    T: np.ndarray = None        # temperature (K), N
    A: np.ndarray = None        # dislocation-creep prefactor, N
    B: np.ndarray = None        # Q/R (K), N
    n: np.ndarray = None        # stress exponent, N
    visc_cap: float = 1e30      # PowerLaw _effViscCap default (powerLaw.cpp:757)


def power_law(Ny, Nz, order=4, Ly=30.0, Lz=30.0, name="custom", steps=300):
    """The C3 physics: the C1 friction on a constant grid with dislocation
    creep, A = 1e-3, n = 3.5, B = Q/R = 130 kJ/mol / 8.314462618 J/(mol K)
    (disl_AVals/nVals/BVals), geotherm T = 293 + 25 z K (TVals = [293 1043],
    TDepths = [0 30]); viscous strains start at zero."""
    base = strike_slip(Ny, Nz, order=order, var_grid=False, Ly=Ly, Lz=Lz, name=name, steps=steps)
    z = base.z
    N = Ny * Nz
    T = set_vec(z, [293.0, 1043.0], [0.0, 30.0])
    kw = {f: getattr(base, f) for f in QdConfig.__dataclass_fields__}
    return PowerLawConfig(**kw, T=T, A=np.full(N, 1e-3), B=np.full(N, 130000.0 / 8.314462618),
                          n=np.full(N, 3.5))


def config_c3(Ny=801, Nz=801):
    """801^2 power-law viscoelastic QD, order 4, constant grid, 300 steps."""
    return power_law(Ny, Nz, name=f"C3-{Ny}x{Nz}", steps=300)


def steady_prestress(cfg):
    """Set the prestress to the steady sliding stress at vL for the initial
    state, tau0 = a sN asinh(vL/(2 v0) e^{psi0/a}) + eta vL, so that the
    whole fault creeps at about vL from t = 0 instead of being loaded for
    centuries first."""
    cfg.prestress = (cfg.a * cfg.sN * np.arcsinh(cfg.vL / (2 * cfg.v0) * np.exp(cfg.psi0 / cfg.a))
                     + cfg.eta * cfg.vL)
    return cfg


def config_c5(Ny=401, Nz=401):
    """401^2, the C2 material and grid (order 4, sinh-stretched z, layered mu)
    with the C1/C2 friction and a steady prestress; 500 RK32 steps for the
    adjoint."""
    return steady_prestress(strike_slip(Ny, Nz, order=4, var_grid=True, zstretch=5.0,
                                        layered_mu=True, name=f"C5-{Ny}x{Nz}", steps=500))


def hat_matrix(z, n_nodes, L):
    """P[k, i] = hat function i of n_nodes equispaced nodes on [0, L] at z[k]:
    a nodal field f = P theta is the piecewise-linear profile through theta."""
    nodes = np.linspace(0.0, L, n_nodes)
    P = np.zeros((len(z), n_nodes))
    h = nodes[1] - nodes[0]
    for i, zi in enumerate(nodes):
        P[:, i] = np.clip(1.0 - np.abs(z - zi) / h, 0.0, None)
    return P, nodes


def splitmix_choice(seed, n, lo, hi):
    """n distinct integers from [lo, hi) by a partial Fisher-Yates shuffle
    driven by SplitMix64 uniforms."""
    pool = list(range(lo, hi))
    u = splitmix64(seed, n)
    out = []
    for i in range(n):
        j = i + int(u[i] * (len(pool) - i))
        pool[i], pool[j] = pool[j], pool[i]
        out.append(pool[i])
    return np.array(out, dtype=np.int32)


@dataclass
class AdjointProblem:
    P: np.ndarray            # Nz x n_nodes hat functions
    nodes: np.ndarray        # node depths (km)
    theta: dict              # nodal profiles "a", "b", "Dc" (n_nodes each)
    stations: np.ndarray     # station y-indices (surface nodes j * Nz)


def adjoint_problem(cfg, n_nodes=64, n_stations=50, seed=0xAD70):
    """C5's inverse problem on a QD config: a, b, Dc as n_nodes-node
    piecewise-linear depth profiles (through the config's values at the
    nodes) and n_stations surface stations drawn without replacement from
    [1, Ny-1)."""
    zf = cfg.z[:cfg.Nz]
    P, nodes = hat_matrix(zf, n_nodes, cfg.Lz)
    theta = {k: np.interp(nodes, zf, getattr(cfg, k)) for k in ("a", "b", "Dc")}
    st = splitmix_choice(seed, n_stations, 1, cfg.Ny - 1)
    return AdjointProblem(P=P, nodes=nodes, theta=theta, stations=st)
