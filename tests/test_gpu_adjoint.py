"""Discrete adjoint of the QD step (sc_qd_adjoint) against finite
differences of the GPU forward and of the FP64 oracle's forward (direct
Cholesky solves), with the step sizes held fixed.

Parameters are piecewise-linear depth profiles of a, b and Dc (hat
functions); the misfit is J = 1/2 sum_s (u(j_s, 0, t_N) - d_s)^2 with d from
a perturbed model.  Central differences with h = 1e-6 |theta_i|: their
truncation error is O(h^2) and the forward's solve/root tolerances
(1e-13 relative, 1e-12 MPa) put a noise of about 1e-7 relative on them; the
bound used is 1e-5 relative.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _problem(N=33, steps=40, n_nodes=8, n_st=6, reuse=4, precond="fdiag64"):
    from scycle_amd import QdMediator
    from scycle_amd.configs import adjoint_problem, steady_prestress, strike_slip
    cfg = steady_prestress(strike_slip(N, N, order=4, var_grid=True, zstretch=3.0, layered_mu=True))
    prob = adjoint_problem(cfg, n_nodes, n_st)
    th = prob.theta
    m = QdMediator(cfg, solve="cg64", rtol=1e-13, precond=precond, reuse=reuse, jac_mode=1)
    m.set_friction(prob.P @ th["a"], prob.P @ th["b"], prob.P @ th["Dc"])
    m.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    _, n, dts = m.integrate(steps, tol=1e-8, init_dt=1e-3, max_dt=1e10)
    assert n == steps
    print("window", dts.sum() / 3.15e7, "yr, dt", dts.min(), "..", dts.max(), flush=True)
    m.set_state(cfg.psi0, cfg.slip0, np.zeros(N))
    # data: a model with a 10% weaker velocity-weakening patch at mid depth
    bump = 1.0 - 0.1 * np.exp(-((prob.nodes - 0.45 * cfg.Lz) / (0.15 * cfg.Lz)) ** 2)
    m.set_friction(prob.P @ (th["a"] * bump), None, None)
    _, d = m.forward_misfit(dts, prob.stations)
    m.set_friction(prob.P @ th["a"], None, None)
    return cfg, prob, m, dts, d


def _J(m, prob, dts, d, th):
    m.set_friction(prob.P @ th["a"], prob.P @ th["b"], prob.P @ th["Dc"])
    return m.forward_misfit(dts, prob.stations, d)[0]


def _theta_grad(prob, g):
    return {k: prob.P.T @ gk for k, gk in zip(("a", "b", "Dc"), g)}


def test_adjoint_matches_gpu_finite_differences():
    cfg, prob, m, dts, d = _problem()
    th = prob.theta
    ga, gb, gdc, st = m.adjoint(dts, prob.stations, d, ckpt_every=10)
    G = _theta_grad(prob, (ga, gb, gdc))
    J0 = _J(m, prob, dts, d, th)
    print("J", st["J"], "forward J", J0, "stats", st, flush=True)
    assert abs(st["J"] - J0) <= 1e-10 * J0
    # a random direction over all three profiles
    rng = np.random.default_rng(3)
    dirn = {k: np.abs(th[k]) * rng.standard_normal(len(th[k])) for k in th}
    eps = 1e-6
    Jp = _J(m, prob, dts, d, {k: th[k] + eps * dirn[k] for k in th})
    Jm = _J(m, prob, dts, d, {k: th[k] - eps * dirn[k] for k in th})
    fd = (Jp - Jm) / (2 * eps)
    ad = sum(float(G[k] @ dirn[k]) for k in th)
    print("directional: adjoint", ad, "fd", fd, "rel", abs(ad - fd) / abs(fd), flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    # single profile nodes: the largest component of each family
    for k in th:
        i = int(np.argmax(np.abs(G[k])))
        h = 1e-6 * abs(th[k][i])
        tp = {q: th[q].copy() for q in th}
        tm = {q: th[q].copy() for q in th}
        tp[k][i] += h
        tm[k][i] -= h
        fdi = (_J(m, prob, dts, d, tp) - _J(m, prob, dts, d, tm)) / (2 * h)
        print(f"dJ/d{k}[{i}]: adjoint {G[k][i]:.10e} fd {fdi:.10e}", flush=True)
        assert abs(G[k][i] - fdi) <= 1e-5 * abs(fdi)
    m.close()


def test_adjoint_matches_oracle_finite_differences():
    """The FP64 reference: central differences of the oracle's forward."""
    cfg, prob, m, dts, d = _problem(N=25, steps=30, n_nodes=6, n_st=5)
    th = prob.theta
    ga, gb, gdc, st = m.adjoint(dts, prob.stations, d, ckpt_every=8)
    G = _theta_grad(prob, (ga, gb, gdc))
    rng = np.random.default_rng(5)
    dirn = {k: np.abs(th[k]) * rng.standard_normal(len(th[k])) for k in th}
    eps = 1e-6

    def J_oracle(t):
        o = ol.OracleQd(cfg, a=prob.P @ t["a"], b=prob.P @ t["b"], Dc=prob.P @ t["Dc"])
        r = o.replay(dts, cfg.psi0, cfg.slip0)
        return 0.5 * float(np.sum((r["u"][prob.stations * cfg.Nz] - d) ** 2))

    J0 = J_oracle(th)
    fd = (J_oracle({k: th[k] + eps * dirn[k] for k in th}) -
          J_oracle({k: th[k] - eps * dirn[k] for k in th})) / (2 * eps)
    ad = sum(float(G[k] @ dirn[k]) for k in th)
    print("oracle J", J0, "GPU J", st["J"], "directional: adjoint", ad, "oracle fd", fd, flush=True)
    assert abs(st["J"] - J0) <= 1e-8 * J0
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    m.close()


def test_adjoint_checkpoint_interval_and_precision():
    """The gradient does not depend on the checkpoint interval (the
    recomputed segments reproduce the forward), and the mixed-precision
    preconditioner of the adjoint solves changes it only at the solve
    tolerance."""
    cfg, prob, m, dts, d = _problem(N=33, steps=24)
    g1 = m.adjoint(dts, prob.stations, d, ckpt_every=5)
    g2 = m.adjoint(dts, prob.stations, d, ckpt_every=100)
    g3 = m.adjoint(dts, prob.stations, d, ckpt_every=7, precond="fdiag32")
    g4 = m.adjoint(dts, prob.stations, d, ckpt_every=7, reuse=0, rtol=1e-12)
    assert g1[3]["recomputed_steps"] == 24 - 4 and g2[3]["recomputed_steps"] == 0
    for other, tol in ((g2, 1e-9), (g3, 1e-8), (g4, 1e-8)):
        for x, y in zip(g1[:3], other[:3]):
            err = np.abs(x - y).max() / np.abs(x).max()
            print("gradient difference", err, flush=True)
            assert err <= tol
    m.close()


def test_adjoint_torch_function():
    """J through torch.autograd: the backward is P^T times the adjoint."""
    import torch
    from qd_adjoint_fn import MisfitContext, QdMisfit
    cfg, prob, m, dts, d = _problem(N=25, steps=20, n_nodes=6, n_st=5)
    th = {k: torch.tensor(prob.theta[k], requires_grad=True) for k in prob.theta}
    mc = MisfitContext(m, prob.P, dts, prob.stations, d, ckpt_every=6)
    J = QdMisfit.apply(th["a"], th["b"], th["Dc"], mc)
    J.backward()
    ga, gb, gdc, _ = m.adjoint(dts, prob.stations, d, ckpt_every=6)
    for k, g in zip(("a", "b", "Dc"), (ga, gb, gdc)):
        assert np.allclose(th[k].grad.numpy(), prob.P.T @ g, rtol=1e-10, atol=0)
    # one descent step along -grad lowers J
    with torch.no_grad():
        step = 1e-3 / max(float(th[k].grad.abs().max() / th[k].abs().max()) for k in th)
        new = {k: th[k] - step * th[k].grad for k in th}
    J2 = QdMisfit.apply(new["a"], new["b"], new["Dc"], mc)
    print("J", float(J), "after a descent step", float(J2), flush=True)
    assert float(J2) < float(J)
    m.close()


def test_adjoint_boundary_operator_matches_per_stage_solves():
    """The reverse pass through the precomputed boundary operator
    B = (1/fts) R_L^T A^-1 S^T gives the gradient of the pass that solves
    per stage (both to the solve tolerance, 1e-12 relative residual); B is
    formed once (Nz solves), reused by the next gradient and after the
    friction parameters change (A does not depend on them), and formed
    again for other adjoint solve options."""
    cfg, prob, m, dts, d = _problem(N=33, steps=24)
    gs = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="never")
    gb1 = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="always")
    gb2 = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="always")
    print("per-stage solves", gs[3], "\nboundary operator", gb1[3], "\nreused", gb2[3], flush=True)
    assert gs[3]["green_solves"] == 0 and gs[3]["adj_solves"] == 3 * 24 + 1
    assert gb1[3]["green_solves"] == cfg.Nz and gb1[3]["adj_solves"] == 1
    assert gb2[3]["green_solves"] == 0
    for x, y in zip(gs[:3], gb1[:3]):
        err = np.abs(x - y).max() / np.abs(x).max()
        print("per-stage vs operator", err, flush=True)
        assert err <= 1e-9
    for x, y in zip(gb1[:3], gb2[:3]):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(x).max()
    m.set_friction(prob.P @ (1.01 * prob.theta["a"]), None, None)
    gb3 = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="always")
    gs3 = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="never")
    assert gb3[3]["green_solves"] == 0
    for x, y in zip(gs3[:3], gb3[:3]):
        assert np.abs(x - y).max() <= 1e-9 * np.abs(x).max()
    gb4 = m.adjoint(dts, prob.stations, d, ckpt_every=7, green="always", precond="fdiag32", rtol=1e-11)
    assert gb4[3]["green_solves"] == cfg.Nz
    for x, y in zip(gb3[:3], gb4[:3]):
        assert np.abs(x - y).max() <= 1e-8 * np.abs(x).max()
    m.close()
