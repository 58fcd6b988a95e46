"""GPU parity of the steady-state start and of the .in deck drop-in.

* sc_qd_solve_ss (StrikeSlip_LinearElastic_qd::solveSS + solveSSb,
  strikeSlip_linearElastic_qd.cpp:895-1006, Fault::guessSS fault.cpp:633-709)
  against the oracle's solveSS (direct Cholesky solves): psi, slip, the
  shifted displacement u and bcRShift to 1e-9 of their magnitudes, then the
  rates of the first d_dt, for remote loading on the right and for a
  free-surface right side (bcR keeps bcRShift as its datum afterwards, as
  solveSSb leaves it; the bottom is then the Dirichlet side).
* sc_qd_create_from_file on the committed decks tests/golden/deck_c1 (21x21,
  order 2, constant grid) and deck_c2 (25x25, order 4, sinh z from a file,
  layered mu from momBal_mu): the mediator's state after its solveSS, its
  d_dt, the deck's run settings (RK32, PID, maxStepCount, the minimum step
  of computeMinTimeStep, strikeSlip_linearElastic_qd.cpp:262-322) and 20
  RK32 steps against the oracle built from the same deck.  The test turns
  the deck into arrays itself (the reference's grid, setVec interpolation,
  Q2's y-interpolated bulk mu, file overrides), independently of the
  library's deck reader.
Tolerances: solves end at relative residual 1e-12 on the GPU (Jacobi-PCG,
the deck's solver) against direct solves in the oracle, so fields agree to
~1e-10; the parity bar is 1e-9 (1e-8 for integrated slip in replay mode).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import OracleQd, OracleSbp, lib as olib, ptr as optr, qd_solve_ss

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _dpsi_err(dpsi, ref, cfg, psi, V):
    """Error of the aging-law state rate relative to the magnitude of its two
    terms, (b v0/Dc)(e^((f0-psi)/b) + V/v0): near steady state the rate is a
    difference of nearly equal terms, so its own size is no scale."""
    scale = cfg.b * cfg.v0 / cfg.Dc * (np.exp((cfg.f0 - psi) / cfg.b) + np.abs(V) / cfg.v0)
    return (np.abs(dpsi - ref) / scale).max()


def _cfg(order, var_grid, N, mat_bc=(0, 1, 0, 1), remote=True):
    from scycle_amd.configs import strike_slip
    cfg = strike_slip(N, N, order=order, var_grid=var_grid, zstretch=3.0 if var_grid else None,
                      layered_mu=var_grid)
    cfg.mat_bc = mat_bc
    cfg.remote_loading_R = remote
    return cfg


@pytest.mark.parametrize("order,var_grid,N,mat_bc,remote", [
    (2, False, 21, (0, 1, 0, 1), True),
    (4, True, 25, (0, 1, 0, 1), True),
    (4, True, 25, (1, 1, 0, 0), False),
])
def test_solve_ss_parity(order, var_grid, N, mat_bc, remote):
    from scycle_amd import QdMediator
    print("solve_ss", order, var_grid, N, mat_bc, remote, flush=True)
    cfg = _cfg(order, var_grid, N, mat_bc, remote)
    g = QdMediator(cfg, solve="cg64", rtol=1e-13, jac_mode=0)
    o = OracleQd(cfg, jac_mode=0)
    for keep in (False, True):
        psi, slip = g.solve_ss(cfg.psi0, have_state_vals=keep)
        ref = qd_solve_ss(o.h, N, N * N, cfg.psi0, have_state_vals=keep)
        u = g.get_u()
        assert np.abs(psi - ref["psi"]).max() < 1e-15
        e = dict(u=_rel(u, ref["u"]), slip=np.abs(slip - ref["slip"]).max() / np.abs(ref["u"]).max(),
                 shift=_rel(u[(N - 1) * N:], ref["bcRShift"]))
        print(keep, e, flush=True)
        assert max(e.values()) < 1e-9, e
        st = g.get_state()
        assert np.all(st["V"] == 1e-9)
    # the first d_dt from the steady-state start (the oracle's bcR and bcL as
    # solveSSb leaves them)
    for t in (0.0, 2.5e7):
        dpsi, dslip = g.d_dt(t)
        Vo = np.full(N, 1e-9)
        dp_o, ds_o = np.zeros(N), np.zeros(N)
        assert olib().orc_qd_ddt(o.h, t, optr(ref["psi"]), optr(ref["slip"]), optr(Vo), optr(dp_o),
                                 optr(ds_o), None, None) == 0
        ep = _dpsi_err(dpsi, dp_o, cfg, ref["psi"], ds_o)
        print("d_dt", t, _rel(dslip, ds_o), ep, flush=True)
        assert _rel(dslip, ds_o) < 1e-9 and ep < 1e-9


# ------------------------------------------------------------- decks ----
def _read_in(path):
    """key -> value of a .in deck (the value cut at the first space, the last
    occurrence winning; vectors "[a b ...]")."""
    out = {}
    for line in open(path):
        line = line.split("#")[0]
        if " = " not in line:
            continue
        k, v = line.split(" = ", 1)
        v = v.strip()
        if v.startswith("["):
            items = v[1:v.index("]")].split()
            try:
                out[k.strip()] = [float(x) for x in items]
            except ValueError:
                out[k.strip()] = items
        else:
            out[k.strip()] = v.split(" ")[0]
    return out


def _read_vec(path):
    raw = open(path, "rb").read()
    assert int.from_bytes(raw[:4], "big") == 1211214
    n = int.from_bytes(raw[4:8], "big")
    return np.frombuffer(raw[8:8 + 8 * n], dtype=">f8").astype(np.float64)


def _set_vec(coord, vals, depths):
    """setVec (genFuncs.cpp:573-617): linear between the depth pairs, the
    last segment extended beyond the last depth, vals[0] elsewhere."""
    out = np.full(len(coord), vals[0])
    for i in range(len(depths) - 1):
        z0, z1, v0, v1 = depths[i], depths[i + 1], vals[i], vals[i + 1]
        m = (coord >= z0) & (coord <= z1) | ((coord > z1) & (i == len(depths) - 2))
        out[m] = (v1 - v0) / (z1 - z0) * (coord[m] - z0) + v0
    return out


def _deck_config(deck_path):
    from scycle_amd.configs import QdConfig
    d = _read_in(deck_path)
    Ny, Nz, Ly, Lz = int(d["Ny"]), int(d["Nz"]), float(d["Ly"]), float(d["Lz"])
    inp = d["inputDir"]
    q = np.repeat(np.arange(Ny) / (Ny - 1.0), Nz)
    r = np.tile(np.arange(Nz) / (Nz - 1.0), Ny)
    y, z = Ly * q, Lz * r
    if os.path.exists(inp + "z"):
        z = _read_vec(inp + "z")
    mu = _set_vec(y, d["muVals"], d["muDepths"])          # Q2: bulk mu along y
    if os.path.exists(inp + "momBal_mu"):
        mu = _read_vec(inp + "momBal_mu")
    zf = z[:Nz]
    fts = 2.0
    muf = _set_vec(zf, d["muVals"], d["muDepths"])
    rhof = _set_vec(zf, d["rhoVals"], d["rhoDepths"])
    psi = _set_vec(zf, d["stateVals"], d["stateDepths"])
    if os.path.exists(inp + "psi"):
        psi = _read_vec(inp + "psi")
    cfg = QdConfig(name=os.path.basename(deck_path), Ny=Ny, Nz=Nz, Ly=Ly, Lz=Lz,
                   order=int(d["order"]), var_grid=d["gridSpacingType"] == "variableGridSpacing",
                   mu=mu, y=y, z=z, a=_set_vec(zf, d["aVals"], d["aDepths"]),
                   b=_set_vec(zf, d["bVals"], d["bDepths"]),
                   Dc=_set_vec(zf, d["DcVals"], d["DcDepths"]),
                   sN=_set_vec(zf, d["sNVals"], d["sNDepths"]),
                   eta=np.sqrt(muf * rhof) / fts, psi0=psi, slip0=np.zeros(Nz))
    rho = _set_vec(z, d["rhoVals"], d["rhoDepths"])
    cs = np.sqrt(mu / rho)
    if os.path.exists(inp + "momBal_cs"):
        cs = _read_vec(inp + "momBal_cs")
    return cfg, d, cs


@pytest.mark.parametrize("deck", ["deck_c1/c1.in", "deck_c2/c2.in"])
def test_deck_mediator_matches_oracle(deck, monkeypatch):
    from scycle_amd import QdMediator
    monkeypatch.chdir(ROOT)  # the deck's inputDir is relative to the working directory
    path = os.path.join("tests", "golden", deck)
    print("deck", path, flush=True)
    cfg, d, cs = _deck_config(path)
    Ny, Nz, N = cfg.Ny, cfg.Nz, cfg.Ny * cfg.Nz
    g = QdMediator.from_file(path)
    o = OracleQd(cfg, jac_mode=0)
    ref = qd_solve_ss(o.h, Nz, N, cfg.psi0, have_state_vals=True)  # stateVals given
    st = g.get_state()
    e = dict(psi=_rel(st["psi"], ref["psi"]), u=_rel(g.get_u(), ref["u"]),
             slip=np.abs(st["slip"] - ref["slip"]).max() / np.abs(ref["u"]).max())
    print("state after solveSS", e, flush=True)
    assert max(e.values()) < 1e-9, e
    assert np.all(st["V"] == 1e-9)
    # the deck's run settings
    run = g.file_run()
    if cfg.var_grid:  # the metric's spacing: yq/(Ny-1), zr/(Nz-1), yq and zr SBP-D1 of y, z
        yq, zr = OracleSbp(cfg.order, Ny, Nz, cfg.Ly, cfg.Lz, cfg.mu, cfg.y, cfg.z).metric()
        dy, dz = yq / (Ny - 1), zr / (Nz - 1)
    else:
        dy, dz = np.full(N, cfg.Ly / (Ny - 1.0)), np.full(N, cfg.Lz / (Nz - 1.0))
    ts = min((dy / cs).min(), (dz / cs).min())
    assert abs(run.minDt - ts) <= 1e-14 * ts, (run.minDt, ts)
    assert (run.integrator, run.control, run.normType, run.maxSteps) == (0, 1, 0, 20)
    assert run.tol == 1e-8 and run.initDt == 1e-3 and run.maxTime == 1e15
    # d_dt at the start
    dpsi, dslip = g.d_dt(0.0)
    Vo = np.full(Nz, 1e-9)
    dp_o, ds_o = np.zeros(Nz), np.zeros(Nz)
    assert olib().orc_qd_ddt(o.h, 0.0, optr(ref["psi"]), optr(ref["slip"]), optr(Vo), optr(dp_o),
                             optr(ds_o), None, None) == 0
    ep = _dpsi_err(dpsi, dp_o, cfg, ref["psi"], ds_o)
    print("d_dt", _rel(dslip, ds_o), ep, flush=True)
    assert _rel(dslip, ds_o) < 1e-9 and ep < 1e-9
    # 20 RK32 steps of the deck's run from the mediator's own start, the
    # oracle replaying the GPU's steps from its solveSS state
    tEnd, steps, dts = g.integrate_run(run)
    # checkInput (strikeSlip_linearElastic_qd.cpp:230-231) compares initDeltaT
    # with minDeltaT before computeMinTimeStep sets it, so the first step is
    # initDeltaT even below the minimum; the controller clamps the later ones
    assert steps == 20 and dts[0] == 1e-3 and np.all(dts[1:] >= run.minDt)
    o2 = OracleQd(cfg, jac_mode=0)
    qd_solve_ss(o2.h, Nz, N, cfg.psi0, have_state_vals=True)
    rp = o2.replay(dts, ref["psi"], ref["slip"], V=np.full(Nz, 1e-9))
    gs = g.get_state()
    e = dict(slip=_rel(gs["slip"], rp["slip"]), psi=_rel(gs["psi"], rp["psi"]), V=_rel(gs["V"], rp["V"]),
             u=_rel(g.get_u(), rp["u"]), t=abs(tEnd / rp["t"] - 1))
    print("20 RK32 steps", e, flush=True)
    assert max(e.values()) < 1e-8, e
