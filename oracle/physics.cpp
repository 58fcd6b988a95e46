// Oracle restatement of the fault, integrator, wave-step and power-law parts
// of the SCycle step (test infrastructure only; see oracle.hpp).
#include "physics.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>

namespace orc {

// fault.cpp:2246-2250
double strength_psi(double sN, double psi, double V, double a, double v0) {
  return a * sN * std::asinh((double)(0.5 * V / v0) * std::exp(psi / a));
}

// fault.cpp:1984-1995
double agingLaw_psi(double psi, double V, double b, double f0, double v0, double Dc) {
  double A = std::exp((double)(f0 - psi) / b);
  double dstate = 0.;
  if (!std::isinf(A) && b > 1e-3) dstate = (double)(b * v0 / Dc) * (A - V / v0);
  return dstate;
}

// fault.cpp:2072-2088
double slipLaw_psi(double psi, double V, double a, double b, double f0, double v0, double Dc) {
  if (V == 0) return 0.0;
  double absV = std::fabs(V);
  double fss = f0 + (a - b) * std::log(absV / v0);
  double f = a * std::asinh((double)(absV / 2. / v0) * std::exp(psi / a));
  return -absV / Dc * (f - fss);
}

// rootFinder.cpp:174-257 (the first residual check at the uninitialised
// member _x, rootFinder.cpp:182, is not reproduced).
int bracketed_newton(const std::function<void(double, double*, double*)>& resid, double left,
                     double right, double x0, double atol, int maxIts, double* out, int* its) {
  double f, fp, fLeft, fRight, dummy;
  resid(left, &fLeft, &dummy);
  resid(right, &fRight, &dummy);
  *its = 0;
  if (std::sqrt(fLeft * fLeft) <= atol) { *out = left; return 0; }
  if (std::sqrt(fRight * fRight) <= atol) { *out = right; return 0; }
  if (fLeft > 0) {
    std::swap(left, right);
    std::swap(fLeft, fRight);
  }
  int numIts = 0;
  double x = x0;
  double dxOld = std::fabs(right - left);
  double dx = dxOld;
  resid(x, &f, &fp);
  while ((numIts <= maxIts) && (std::fabs(f) >= atol)) {
    if (((x - right) * fp - f) * ((x - left) * fp - f) > 0.0 ||
        std::fabs(2.0 * f) > std::fabs(dxOld * fp)) {
      dxOld = dx;
      dx = 0.5 * (right - left);
      x = left + dx;
    } else {
      dxOld = dx;
      dx = f / fp;
      x -= dx;
    }
    resid(x, &f, &fp);
    if (f < 0.0) left = x; else right = x;
    numIts++;
  }
  *out = x;
  *its = numIts;
  return std::fabs(f) > atol ? 1 : 0;
}

// fault.cpp:954-1059
int computeVel_qd(int N, const double* eta, const double* tauQS, const double* sN,
                  const double* psi, const double* a, const double* locked, double v0, double vL,
                  double rootTol, int maxIts, int jac_mode, double* V, long* rootIts) {
  int bad = 0;
  for (int j = 0; j < N; ++j) {
    if (locked && locked[j] > 0.5) { V[j] = 0.; continue; }
    if (locked && locked[j] < -0.5) { V[j] = vL; continue; }
    double left = 0., right = tauQS[j] / eta[j];
    if (std::isnan(left) || std::isnan(right)) return -1;
    double out = V[j];
    if (std::fabs(left - right) < 1e-14) {
      out = left;
    } else {
      auto resid = [&](double vel, double* f, double* J) {
        double strength = strength_psi(sN[j], psi[j], vel, a[j], v0);
        double stress = tauQS[j] - eta[j] * vel;
        *f = strength - stress;
        double A = a[j] * sN[j];
        double B = std::exp(psi[j] / a[j]) / (2. * v0);
        if (jac_mode == 0) *J = A * vel / std::sqrt(B * B * vel * vel + 1.) + eta[j];
        else *J = A * B / std::sqrt(B * B * vel * vel + 1.) + eta[j];
      };
      int its = 0;
      bad |= bracketed_newton(resid, left, right, V[j], rootTol, maxIts, &out, &its);
      if (rootIts) *rootIts += its;
    }
    V[j] = out;
  }
  return bad;
}

// ------------------------------------------------------------------ RK ----
static void waxpy(std::vector<double>& w, double a, const std::vector<double>& x,
                  const std::vector<double>& y) {
  w.resize(x.size());
  for (size_t i = 0; i < x.size(); ++i) w[i] = a * x[i] + y[i];
}
static void vaxpy(std::vector<double>& y, double a, const std::vector<double>& x) {
  for (size_t i = 0; i < x.size(); ++i) y[i] = y[i] + a * x[i];
}

// odeSolver.cpp:395-445 / 842-898
double RkSolver::computeStepSize(double totErr) {
  double stepRatio;
  if (controlType == "P") {
    stepRatio = kappa * std::pow(totTol / totErr, 1. / (1. + ord));
  } else {
    double alpha = 0.49 / ord, beta = 0.34 / ord, gamma = 0.1 / ord;
    if (stepCount - pidStep0 < 4) stepRatio = kappa * std::pow(totTol / totErr, 1. / (1. + ord));
    else
      stepRatio = kappa * std::pow(totTol / totErr, alpha) * std::pow(errA[0] / totTol, beta) *
                  std::pow(totTol / errA[1], gamma);
  }
  double dt = stepRatio * deltaT;
  dt = std::min(deltaT * 5.0, dt);
  dt = std::min(maxDeltaT, dt);
  dt = std::max(minDeltaT, dt);
  return dt;
}

// odeSolver.cpp:449-501 (RK32, lo=y2 hi=y3) / 902-969 (RK43, lo=y3 hi=y4)
double RkSolver::computeError(const VarMap& lo, const VarMap& hi) {
  double tot = 0;
  for (size_t i = 0; i < errInds.size(); ++i) {
    const auto& l = lo.at(errInds[i]);
    const auto& h = hi.at(errInds[i]);
    if (normType == "L2_absolute") {
      double s = 0;
      for (size_t k = 0; k < l.size(); ++k) { double e = -1.0 * h[k] + l[k]; s += e * e; }
      tot += std::sqrt(s) / (std::sqrt((double)h.size()) * scaleV[i]);
    } else if (ord == 3) {
      double s = 0, n = 0;
      for (size_t k = 0; k < l.size(); ++k) { double e = -1.0 * h[k] + l[k]; s += e * e; n += h[k] * h[k]; }
      tot += std::sqrt(s) / (std::sqrt(n) * scaleV[i]);
    } else {
      double m = -INFINITY;
      for (size_t k = 0; k < l.size(); ++k) m = std::max(m, std::fabs(-1.0 * h[k] + l[k]) / h[k]);
      tot += m / scaleV[i];
    }
  }
  return tot;
}

int RkSolver::integrate(OdeCtx* obj) {
  if (errInds.empty())
    for (auto& kv : var) errInds.push_back(kv.first);
  if (scaleV.empty()) scaleV.assign(errInds.size(), 1.0);
  if (finalT == initT) return 0;
  if (deltaT == 0) deltaT = (finalT - initT) / maxNumSteps;
  if (maxNumSteps == 0) return 0;
  for (auto& kv : var) dvar[kv.first].assign(kv.second.size(), 0.0);
  int ierr = obj->d_dt(currT, var, dvar);
  if (ierr) return ierr;
  int stop = 0;
  double totErr = 0;
  while (stepCount < maxNumSteps && currT < finalT) {
    stepCount++;
    int attempt = 0;
    VarMap y2, y3;
    while (attempt < 100) {
      attempt++;
      if (force_dt) deltaT = replayDt[stepCount - 1];
      if (currT + deltaT > finalT) deltaT = finalT - currT;
      if (ord == 3) {
        // odeSolver.cpp:554-597
        VarMap k1, f1, k2, f2;
        for (auto& kv : var) waxpy(k1[kv.first], 0.5 * deltaT, dvar[kv.first], kv.second);
        for (auto& kv : var) f1[kv.first].assign(kv.second.size(), 0.0);
        if ((ierr = obj->d_dt(currT + 0.5 * deltaT, k1, f1))) return ierr;
        for (auto& kv : var) {
          waxpy(k2[kv.first], -deltaT, dvar[kv.first], kv.second);
          vaxpy(k2[kv.first], 2 * deltaT, f1[kv.first]);
        }
        for (auto& kv : var) f2[kv.first].assign(kv.second.size(), 0.0);
        if ((ierr = obj->d_dt(currT + deltaT, k2, f2))) return ierr;
        for (auto& kv : var) {
          const auto& k = kv.first;
          waxpy(y2[k], 0.5 * deltaT, dvar[k], kv.second);
          vaxpy(y2[k], 0.5 * deltaT, f2[k]);
          waxpy(y3[k], deltaT / 6.0, dvar[k], kv.second);
          vaxpy(y3[k], 2 * deltaT / 3.0, f1[k]);
          vaxpy(y3[k], deltaT / 6.0, f2[k]);
        }
        totErr = computeError(y2, y3);
        if (force_dt) break;
        if (totErr <= totTol) break;
      } else {
        // odeSolver.cpp:986-1148 (Kennedy-Carpenter ARK4(3)6L[2]SA, explicit part)
        const double c2 = 1. / 2., c3 = 83. / 250., c4 = 31. / 50., c5 = 17. / 20., c6 = 1.;
        const double b1 = 82889. / 524892., b3 = 15625. / 83664., b4 = 69875. / 102672.,
                     b5 = -2260. / 8211., b6 = 1. / 4.;
        const double hb1 = 4586570599. / 29645900160., hb3 = 178811875. / 945068544.,
                     hb4 = 814220225. / 1159782912., hb5 = -3700637. / 11593932.,
                     hb6 = 61727. / 225920.;
        const double a21 = 1. / 2., a31 = 13861. / 62500., a32 = 6889. / 62500.;
        const double a41 = -116923316275. / 2393684061468., a42 = -2731218467317. / 15368042101831.,
                     a43 = 9408046702089. / 11113171139209.;
        const double a51 = -451086348788. / 2902428689909., a52 = -2682348792572. / 7519795681897.,
                     a53 = 12662868775082. / 11960479115383., a54 = 3355817975965. / 11060851509271.;
        const double a61 = 647845179188. / 3216320057751., a62 = 73281519250. / 8382639484533.,
                     a63 = 552539513391. / 3454668386233., a64 = 3354512671639. / 8306763924573.,
                     a65 = 4040. / 17871.;
        VarMap k2, k3, k4, k5, k6, f2, f3, f4, f5, f6;
        const VarMap& f1 = dvar;
        auto zero = [&](VarMap& m) { for (auto& kv : var) m[kv.first].assign(kv.second.size(), 0.0); };
        for (auto& kv : var) waxpy(k2[kv.first], a21 * deltaT, f1.at(kv.first), kv.second);
        zero(f2);
        if ((ierr = obj->d_dt(currT + c2 * deltaT, k2, f2))) return ierr;
        for (auto& kv : var) {
          waxpy(k3[kv.first], a31 * deltaT, f1.at(kv.first), kv.second);
          vaxpy(k3[kv.first], a32 * deltaT, f2[kv.first]);
        }
        zero(f3);
        if ((ierr = obj->d_dt(currT + c3 * deltaT, k3, f3))) return ierr;
        for (auto& kv : var) {
          const auto& k = kv.first;
          waxpy(k4[k], a41 * deltaT, f1.at(k), kv.second);
          vaxpy(k4[k], a42 * deltaT, f2[k]);
          vaxpy(k4[k], a43 * deltaT, f3[k]);
        }
        zero(f4);
        if ((ierr = obj->d_dt(currT + c4 * deltaT, k4, f4))) return ierr;
        for (auto& kv : var) {
          const auto& k = kv.first;
          waxpy(k5[k], a51 * deltaT, f1.at(k), kv.second);
          vaxpy(k5[k], a52 * deltaT, f2[k]);
          vaxpy(k5[k], a53 * deltaT, f3[k]);
          vaxpy(k5[k], a54 * deltaT, f4[k]);
        }
        zero(f5);
        if ((ierr = obj->d_dt(currT + c5 * deltaT, k5, f5))) return ierr;
        for (auto& kv : var) {
          const auto& k = kv.first;
          waxpy(k6[k], a61 * deltaT, f1.at(k), kv.second);
          vaxpy(k6[k], a62 * deltaT, f2[k]);
          vaxpy(k6[k], a63 * deltaT, f3[k]);
          vaxpy(k6[k], a64 * deltaT, f4[k]);
          vaxpy(k6[k], a65 * deltaT, f5[k]);
        }
        zero(f6);
        if ((ierr = obj->d_dt(currT + c6 * deltaT, k6, f6))) return ierr;
        for (auto& kv : var) {
          const auto& k = kv.first;
          waxpy(y2[k], hb1 * deltaT, f1.at(k), kv.second);  // y2 holds the 3rd-order y3
          vaxpy(y2[k], hb3 * deltaT, f3[k]);
          vaxpy(y2[k], hb4 * deltaT, f4[k]);
          vaxpy(y2[k], hb5 * deltaT, f5[k]);
          vaxpy(y2[k], hb6 * deltaT, f6[k]);
          waxpy(y3[k], b1 * deltaT, f1.at(k), kv.second);   // y3 holds the 4th-order y4
          vaxpy(y3[k], b3 * deltaT, f3[k]);
          vaxpy(y3[k], b4 * deltaT, f4[k]);
          vaxpy(y3[k], b5 * deltaT, f5[k]);
          vaxpy(y3[k], b6 * deltaT, f6[k]);
        }
        totErr = computeError(y2, y3);
        if (force_dt) break;
        if (totErr < totTol) break;
      }
      deltaT = computeStepSize(totErr);
      if (minDeltaT == deltaT) break;
      numRejected++;
    }
    currT = currT + deltaT;
    for (auto& kv : var) {
      kv.second = y3[kv.first];
      std::fill(dvar[kv.first].begin(), dvar[kv.first].end(), 0.0);
    }
    if ((ierr = obj->d_dt(currT, var, dvar))) return ierr;
    if (totErr != 0.0) newDeltaT = computeStepSize(totErr);
    errA[1] = errA[0];
    errA[0] = totErr;
    dtHistory.push_back(deltaT);
    errHistory.push_back(totErr);
    newDtHistory.push_back(newDeltaT);
    obj->timeMonitor(currT, deltaT, stepCount, stop);
    if (stop > 0) break;  // odeSolver.cpp:614, 1161
    deltaT = newDeltaT;
  }
  return 0;
}

// ---------------------------------------------------------------- QD -----
int QdModel::init(const QdConfig& cfg, const double* mu, const double* y, const double* z) {
  c = cfg;
  int e = sbp2d_build(sbp, cfg.sbp, mu, y, z);
  if (e) return e;
  Ny = cfg.sbp.Ny;
  Nz = cfg.sbp.Nz;
  N = Ny * Nz;
  bcL.assign(Nz, 0.0); bcR.assign(Nz, 0.0); bcRShift.assign(Nz, 0.0);
  bcT.assign(Ny, 0.0); bcB.assign(Ny, 0.0);
  rhs.assign(N, 0.0); u.assign(N, 0.0); sxy.assign(N, 0.0);
  tauQS.assign(Nz, 0.0); tau.assign(Nz, 0.0); V.assign(Nz, 0.0); strength.assign(Nz, 0.0);
  psi.assign(Nz, 0.0); slip.assign(Nz, 0.0);
  if (locked.empty()) locked.assign(Nz, 0.0);
  if (prestress.empty()) prestress.assign(Nz, 0.0);
  if (c.solver == 0) {
    if (chol.factor(sbp.A)) return 10;
  }
  return 0;
}

// solveMomentumBalance (strikeSlip_linearElastic_qd.cpp:874-891)
int QdModel::solveMomentumBalance(double /*t*/) {
  sbp2d_set_rhs(sbp, rhs.data(), bcL.data(), bcR.data(), bcT.data(), bcB.data());
  if (c.solver == 0) {
    chol.solve(rhs.data(), u.data());
  } else {
    PcgStats st = pcg(sbp.A, rhs.data(), u.data(), c.cg_rtol, c.cg_maxit);
    cgIters += st.iters;
    if (st.relres > c.cg_rtol) return 20;
  }
  nSolves++;
  sbp2d_muxDy(sbp, u.data(), sxy.data());
  return 0;
}

// strikeSlip_linearElastic_qd.cpp:756-798 + Fault_qd::d_dt fault.cpp:861-918
int QdModel::d_dt(double t, const VarMap& var, VarMap& dvar) {
  const auto& vslip = var.at("slip");
  for (int k = 0; k < Nz; ++k) bcL[k] = vslip[k] * (1.0 / c.faultTypeScale);
  if (c.remoteLoadingR)
    for (int k = 0; k < Nz; ++k) bcR[k] = c.vL * t / c.faultTypeScale + 1.0 * bcRShift[k];
  psi = var.at("psi");
  slip = vslip;
  int e = solveMomentumBalance(t);
  if (e) return e;
  for (int k = 0; k < Nz; ++k) tauQS[k] = sxy[k];
  for (int k = 0; k < Nz; ++k) tauQS[k] = tauQS[k] + 1.0 * prestress[k];
  if (computeVel_qd(Nz, eta.data(), tauQS.data(), sN.data(), psi.data(), a.data(), locked.data(),
                    c.v0, c.vL, c.rootTol, c.maxNumIts, c.jac_mode, V.data(), &rootIts))
    return 30;
  dvar["slip"] = V;
  auto& dpsi = dvar["psi"];
  dpsi.resize(Nz);
  for (int k = 0; k < Nz; ++k)
    dpsi[k] = c.stateLaw == 0 ? agingLaw_psi(psi[k], V[k], b[k], c.f0, c.v0, Dc[k])
                              : slipLaw_psi(psi[k], V[k], a[k], b[k], c.f0, c.v0, Dc[k]);
  for (int k = 0; k < Nz; ++k) tau[k] = -1.0 * (eta[k] * V[k]) + tauQS[k];
  for (int k = 0; k < Nz; ++k) strength[k] = strength_psi(sN[k], psi[k], V[k], a[k], c.v0);
  return 0;
}

int QdModel::solveSS(bool haveStateVals) {
  // Fault::guessSS (fault.cpp:633-668): slip velocity vL, the steady state
  // psi at vL (computePsiSS :672-709) unless the deck gave stateVals, and
  // the steady-state friction as shear stress
  V.assign(Nz, c.vL);
  if (!haveStateVals)
    for (int k = 0; k < Nz; ++k) psi[k] = c.f0 - b[k] * std::log(std::fabs(c.vL) / c.v0);
  std::vector<double> tauP(Nz);
  for (int k = 0; k < Nz; ++k)
    tauP[k] = sN[k] * a[k] * std::asinh((double)0.5 * c.vL * std::exp(psi[k] / a[k]) / c.v0);
  // the mediator builds the momentum balance with a Neumann fault when it
  // guesses steady-state ICs (strikeSlip_linearElastic_qd.cpp:60-62); the
  // traction is the fault datum bcL (:912-915), bcR as set up (0)
  const int bR = sbp.d.bcR, bT = sbp.d.bcT, bL = sbp.d.bcL, bB = sbp.d.bcB;
  if (sbp2d_change_bc(sbp, bR, bT, NEUMANN, bB)) return 40;
  if (c.solver == 0 && chol.factor(sbp.A)) return 41;
  bcL = tauP;
  std::fill(u.begin(), u.end(), 0.0);
  int e = solveMomentumBalance(0.0);
  if (e) return e;
  // solveSSb (:954-1006): shift u so that it has no negative values, the
  // remote boundary's displacement becomes bcRShift and bcR, the fault's
  // the slip (times faultTypeScale) and bcL
  double mn = u[0];
  for (int i = 1; i < N; ++i) mn = std::min(mn, u[i]);
  if (mn < 0) {
    mn = std::fabs(mn);
    for (int i = 0; i < N; ++i) u[i] = u[i] + 1.0 * mn;
  }
  for (int k = 0; k < Nz; ++k) bcRShift[k] = u[(size_t)(Ny - 1) * Nz + k];
  bcR = bcRShift;
  for (int k = 0; k < Nz; ++k) {
    slip[k] = u[k] * c.faultTypeScale;
    bcL[k] = u[k];
  }
  // changeBCTypes back to the mediator's material types (:930)
  if (sbp2d_change_bc(sbp, bR, bT, bL, bB)) return 42;
  if (c.solver == 0 && chol.factor(sbp.A)) return 43;
  return 0;
}

// ---------------------------------------------------------------- FD -----
void fd_propagate(const Sbp2D& s, double dt, const double* u, const double* uPrev,
                  const double* rho, const double* ay, double* uNext, double* d2u) {
  const int N = s.N();
  std::vector<double> temp(N), t2(N);
  spmv(s.A, u, temp.data());
  spmv(s.Hinv, temp.data(), d2u);
  if (s.d.var_grid) {
    spmv(s.Jinv, d2u, t2.data());
    std::copy(t2.begin(), t2.end(), d2u);
  }
  for (int i = 0; i < N; ++i) {
    double c1 = dt * dt / rho[i];
    double c2 = dt * ay[i] - 1.0;
    double c3 = dt * ay[i] + 1.0;
    uNext[i] = (c1 * d2u[i] + 2. * u[i] + c2 * uPrev[i]) / c3;
  }
}

void fd_ay(const Sbp2D& s, const double* cs, int outR, int outT, int outB, int faultL, double* ay) {
  const int Ny = s.d.Ny, Nz = s.d.Nz, N = Ny * Nz;
  for (int i = 0; i < N; ++i) {
    double v = 0;
    if ((i / Nz == Ny - 1) && outR) v += 0.5 / s.h11y;
    if ((i % Nz == 0) && outT) v += 0.5 / s.h11z;
    if (((i + 1) % Nz == 0) && outB) v += 0.5 / s.h11z;
    if ((i / Nz == 0) && faultL) v += 0.5 / s.h11y;
    ay[i] = v * cs[i];
  }
}

// Fault_fd::d_dt (fault.cpp:1436-1523), setPhi (:1527-1575), ComputeVel_fd
// (:1588-1685), ComputeAging_fd / ComputeSlipLaw_fd (:1697-1879)
int FdFault::step(double dt, const double* psiCur, const double* psiPrev, const double* slipCur,
                  const double* slipPrev, double* psiNext, double* slipNext) {
  int bad = 0;
  for (int j = 0; j < N; ++j) {
    psi[j] = psiCur[j];
    slip[j] = slipCur[j];
    uPrev[j] = (-1.0 * slip0[j] + slipPrev[j]) * (1.0 / faultTypeScale);
    u[j] = (-1.0 * slip0[j] + slipCur[j]) * (1.0 / faultTypeScale);
  }
  for (int j = 0; j < N; ++j) {
    an[j] = d2u[j] + tau0[j] / alphay[j];
    Phi[j] = 2.0 / dt * (u[j] - uPrev[j]) + dt * an[j] / rho[j];
    fricPen[j] = dt / alphay[j] / rho[j];
  }
  for (int j = 0; j < N; ++j) {
    if (locked.size() && locked[j] > 0.5) { V[j] = 0.; continue; }
    if (locked.size() && locked[j] < -0.5) { V[j] = vL; continue; }
    double left = 0., right = std::fabs(Phi[j]);
    double out;
    if (std::fabs(left - right) < 1e-14) out = left;
    else {
      auto resid = [&](double vel, double* f, double* J) {
        double cons = fricPen[j] * strength_psi(sNEff[j], psi[j], vel, a[j], v0);
        double P = Phi[j] < 0 ? -Phi[j] : Phi[j];
        *f = cons - (P - vel);
        double A = a[j] * sNEff[j];
        double B = std::exp(psi[j] / a[j]) / (2. * v0);
        *J = 1 + fricPen[j] * A * B / std::sqrt(1. + B * B * vel * vel);
      };
      int its;
      bad |= bracketed_newton(resid, left, right, std::fabs(V[j]), rootTol, maxNumIts, &out, &its);
      rootIts += its;
    }
    V[j] = out;
  }
  for (int j = 0; j < N; ++j) {
    if (V[j] < 1e-14) continue;
    double fric = strength_psi(sNEff[j], psi[j], V[j], a[j], v0);
    double alpha = 1.0 / (rho[j] * alphay[j]) * fric / V[j];
    double A = 1.0 + alpha * dt;
    V[j] = Phi[j] / (1. + dt * alpha);
    u[j] = (2. * u[j] + (an[j] * dt * dt / rho[j]) + (dt * alpha - 1.) * uPrev[j]) / A;
  }
  for (int j = 0; j < N; ++j) {
    auto resid = [&](double state, double* f, double* J) {
      double mid = (psiPrev[j] + state) / 2.0;
      if (stateLaw == 0) {
        double G = agingLaw_psi(mid, V[j], b[j], f0, v0, Dc[j]);
        *f = -2 * dt * G + (state - psiPrev[j]);
        *J = 1 + dt * v0 / Dc[j] * std::exp((f0 - mid) / b[j]);
      } else {
        double G = slipLaw_psi(mid, V[j], a[j], b[j], f0, v0, Dc[j]);
        *f = -2 * dt * G + (state - psiPrev[j]);
        double A = std::fabs(V[j]) / 2. / v0 * std::exp((psiPrev[j] + state) / 2.0 / a[j]);
        *J = 1 + dt * std::fabs(V[j]) / Dc[j] * A / std::sqrt(1 + A * A);
      }
    };
    int its;
    double out;
    bad |= bracketed_newton(resid, -2., 2., psiCur[j], rootTol, maxNumIts, &out, &its);
    rootIts += its;
    psiNext[j] = out;
    psi[j] = out;
  }
  for (int j = 0; j < N; ++j) {
    slip[j] = faultTypeScale * u[j] + slip0[j];
    slipNext[j] = slip[j];
    strength[j] = strength_psi(sNEff[j], psi[j], V[j], a[j], v0);
  }
  return bad;
}

// ----------------------------------------------------- QD <-> FD -----
static double csr_diag(const Csr& m, int i) {
  for (int64_t p = m.rp[i]; p < m.rp[i + 1]; ++p)
    if (m.ci[p] == i) return m.v[p];
  return 0.0;
}

// constructor + computePenaltyVectors (:446-509) + computeTimeStep (:356-443)
int QdFdModel::init(QdModel* q, const QdFdConfig& cfg, const double* mu, const double* y,
                    const double* z, const double* rho_, const double* cs_) {
  qd = q;
  c = cfg;
  Ny = q->Ny;
  Nz = q->Nz;
  N = q->N;
  SbpDesc d = q->c.sbp;
  d.bcR = c.fdBcR;
  d.bcT = c.fdBcT;
  d.bcL = c.fdBcL;
  d.bcB = c.fdBcB;
  int e = sbp2d_build(fs, d, mu, y, z);
  if (e) return e;
  rho.assign(rho_, rho_ + N);
  cs.assign(cs_, cs_ + N);
  ay.assign(N, 0.0);
  if (c.ayCompat) {
    // the loop as it runs: Jj advances only on the fault row (the if-statement
    // of the fault side has Jj++ as its body)
    int Jj = 0;
    for (int Ii = 0; Ii < N; ++Ii) {
      ay[Jj] = 0;
      if ((Ii / Nz == Ny - 1) && c.outR) ay[Jj] += 0.5 / fs.h11y;
      if ((Ii % Nz == 0) && c.outT) ay[Jj] += 0.5 / fs.h11z;
      if (((Ii + 1) % Nz == 0) && c.outB) ay[Jj] += 0.5 / fs.h11z;
      if (Ii / Nz == 0) Jj++;
    }
    for (int i = 0; i < N; ++i) ay[i] = ay[i] * cs[i];
  } else {
    fd_ay(fs, cs.data(), c.outR, c.outT, c.outB, 1, ay.data());
  }
  // FD fault: the QD fault's parameters; alphay = h11y (times yq on a
  // variable grid) on the fault; rho of the fault column
  ff.N = Nz;
  ff.v0 = q->c.v0; ff.f0 = q->c.f0; ff.vL = q->c.vL; ff.rootTol = q->c.rootTol;
  ff.faultTypeScale = q->c.faultTypeScale; ff.maxNumIts = q->c.maxNumIts; ff.stateLaw = q->c.stateLaw;
  ff.a = q->a; ff.b = q->b; ff.Dc = q->Dc; ff.sNEff = q->sN; ff.locked = q->locked;
  ff.rho.assign(rho.begin(), rho.begin() + Nz);
  ff.alphay.assign(Nz, fs.h11y);
  if (d.var_grid)
    for (int k = 0; k < Nz; ++k) ff.alphay[k] = csr_diag(fs.yq, k) * fs.h11y;
  ff.tau0.assign(Nz, 0.0); ff.slip0.assign(Nz, 0.0);
  for (auto* v : {&ff.psi, &ff.slip, &ff.V, &ff.u, &ff.uPrev, &ff.d2u, &ff.an, &ff.Phi, &ff.fricPen,
                  &ff.strength})
    v->assign(Nz, 0.0);
  for (int l = 0; l < 3; ++l) { psiF[l].assign(Nz, 0.0); slipF[l].assign(Nz, 0.0); }
  u0.assign(N, 0.0); du.assign(N, 0.0); duPrev.assign(N, 0.0); duNext.assign(N, 0.0); d2u.assign(N, 0.0);
  // computeTimeStep
  const double gcfl = d.order == 4 ? 0.7071 / std::sqrt(1.4498) : 0.7071;
  double minY = INFINITY, minZ = INFINITY;
  for (int i = 0; i < N; ++i) {
    const double dy = d.var_grid ? csr_diag(fs.yq, i) / (Ny - 1) : d.Ly / (Ny - 1.0);
    const double dz = d.var_grid ? csr_diag(fs.zr, i) / (Nz - 1) : d.Lz / (Nz - 1.0);
    minY = std::min(minY, dy / cs[i]);
    minZ = std::min(minZ, dz / cs[i]);
  }
  const double maxDt = gcfl * std::min(std::fabs(minY), std::fabs(minZ));
  const double cflDt = c.CFL * gcfl * maxDt;
  const double req = c.deltaT_fd;
  dtFd = maxDt;
  if (req <= 0. && cflDt <= 0.) dtFd = maxDt;
  else if (req > 0. && cflDt <= 0.) dtFd = req;
  else if (req <= 0. && cflDt > 0.) dtFd = cflDt;
  else dtFd = req;
  // checkInput (:282-293)
  if (c.limit_fd < c.trigger_qd2fd) c.limit_fd = 10 * c.trigger_qd2fd;
  if (c.limit_qd > c.trigger_fd2qd) c.limit_qd = c.trigger_qd2fd / 10.0;
  qvar["psi"].assign(Nz, 0.0);
  qvar["slip"].assign(Nz, 0.0);
  return 0;
}

// checkSwitchRegime (:585-640)
bool QdFdModel::checkSwitch(double R) {
  bool must = false;
  if (inDynamic && !allowed && R > c.limit_fd) allowed = true;
  if (!inDynamic && !allowed && R < c.limit_qd) allowed = true;
  if (inDynamic && allowed && R < c.trigger_fd2qd) must = true;
  if (!inDynamic && allowed && R > c.trigger_qd2fd) must = true;
  return must;
}

// R = eta V / tauQSP, VecMax
double QdFdModel::qdRatio() const {
  double m = -INFINITY;
  for (int k = 0; k < Nz; ++k) m = std::fmax(m, (qd->eta[k] * qd->V[k]) / qd->tauQS[k]);
  return m;
}

// the FD phase's tauQSP = sxy(u + u0) on the fault + prestress (:1696-1708)
double QdFdModel::fdRatio() const {
  std::vector<double> tot(N), sxy(N);
  for (int i = 0; i < N; ++i) tot[i] = du[i] + 1.0 * u0[i];
  sbp2d_muxDy(fs, tot.data(), sxy.data());
  double m = -INFINITY;
  for (int k = 0; k < Nz; ++k) {
    const double tq = sxy[k] + 1.0 * qd->prestress[k];
    m = std::fmax(m, (qd->eta[k] * ff.V[k]) / tq);
  }
  return m;
}

namespace {
struct QdPhaseCtx : OdeCtx {
  QdFdModel* m = nullptr;
  int d_dt(double t, const VarMap& var, VarMap& dvar) override { return m->qd->d_dt(t, var, dvar); }
  int timeMonitor(double, double, int, int& stop) override {
    m->lastR = m->qdRatio();
    if (m->checkSwitch(m->lastR)) stop = 1;
    return 0;
  }
};
}  // namespace

// integrate_qd (:1436-1509) / integrate_singleQDTimeStep (:891-957)
int QdFdModel::qdPhase(const QdFdRunCfg& r, int regime, int maxSteps, double finalT, double minDt,
                       double maxDt) {
  QdPhaseCtx cx;
  cx.m = this;
  RkSolver rk(r.integrator == 0 ? 3 : 4, maxSteps, finalT, dtFd, r.control == 0 ? "P" : "PID");
  rk.initT = rk.currT = currT;
  rk.totTol = r.tol;
  rk.minDeltaT = minDt;
  rk.maxDeltaT = maxDt;
  rk.normType = r.normType == 0 ? "L2_absolute" : "L2_relative";
  rk.errInds = {"psi", "slip"};
  rk.stepCount = stepCount;             // setInitialStepCount
  rk.maxNumSteps = stepCount + maxSteps;
  if (c.fdStop) rk.pidStep0 = stepCount; // a fresh controller for each QD phase
  rk.var = qvar;
  const int s0 = stepCount;
  const double t0 = currT;
  int e = rk.integrate(&cx);
  if (e) return e;
  qvar = rk.var;
  currT = rk.currT;
  stepCount = rk.stepCount;
  log.push_back(QdFdPhase{regime, s0, stepCount, t0, currT, lastR});
  return 0;
}

// prepare_fd2qd (:747-805)
int QdFdModel::prepareFd2Qd() {
  if (!started) { started = true; return 0; }
  qvar["psi"] = psiF[1];
  qvar["slip"] = slipF[1];
  qd->V = ff.V;
  for (int i = 0; i < N; ++i) qd->u[i] = du[i] + 1.0 * u0[i];
  return 0;
}

// prepare_qd2fd (:807-889)
int QdFdModel::prepareQd2Fd(const QdFdRunCfg& r) {
  const std::vector<double> uStart = qd->u;
  inDynamic = false;
  int e = qdPhase(r, 2, 1, currT + dtFd, dtFd, dtFd);
  inDynamic = true;
  if (e) return e;
  for (int i = 0; i < N; ++i) {
    du[i] = qd->u[i] + -1.0 * uStart[i];
    duPrev[i] = du[i];
  }
  u0 = uStart;
  for (int l : {0, 1}) {
    psiF[l] = qvar["psi"];
    slipF[l] = qvar["slip"];
  }
  ff.V = qd->V;
  ff.slip0 = qvar["slip"];
  for (int k = 0; k < Nz; ++k) ff.tau0[k] = qd->strength[k] + 1.0 * qd->prestress[k];
  return 0;
}

// d_dt of the FD phase (:1676-1759): propagateWaves, Fault_fd::d_dt, the
// fault's u into the body; levels 0 prev, 1 cur, 2 next
int QdFdModel::fdStep() {
  fd_propagate(fs, dtFd, du.data(), duPrev.data(), rho.data(), ay.data(), duNext.data(), d2u.data());
  std::copy(d2u.begin(), d2u.begin() + Nz, ff.d2u.begin());
  int e = ff.step(dtFd, psiF[1].data(), psiF[0].data(), slipF[1].data(), slipF[0].data(), psiF[2].data(),
                  slipF[2].data());
  if (e) return 40;
  std::copy(ff.u.begin(), ff.u.end(), duNext.begin());
  duPrev.swap(du);
  du.swap(duNext);
  psiF[0].swap(psiF[1]);
  psiF[1].swap(psiF[2]);
  slipF[0].swap(slipF[1]);
  slipF[1].swap(slipF[2]);
  return 0;
}

// integrate_fd (:1511-1546), OdeSolver_WaveEq::integrate (odeSolver_WaveEq.cpp:147-186)
int QdFdModel::fdPhase(const QdFdRunCfg& r) {
  const double finalT = r.maxTime;
  const int maxNum = stepCount + r.maxStepCount;
  const int s0 = stepCount;
  const double t0 = currT;
  if (finalT != currT) {
    while (stepCount < maxNum && currT < finalT) {
      currT = currT + dtFd;
      if (currT > finalT) currT = finalT;
      stepCount++;
      int e = fdStep();
      if (e) return e;
      if (c.fdStop) {
        lastR = fdRatio();
        if (checkSwitch(lastR)) break;
      }
    }
    if (!c.fdStop) lastR = fdRatio();
  }
  log.push_back(QdFdPhase{1, s0, stepCount, t0, currT, lastR});
  return 0;
}

// integrate (:511-580)
int QdFdModel::integrate(const QdFdRunCfg& r) {
  log.clear();
  currT = r.initTime;
  stepCount = 0;
  cycleCount = 0;
  started = false;
  auto qdp = [&]() {
    allowed = false;
    inDynamic = false;
    prepareFd2Qd();
    return qdPhase(r, 0, r.maxStepCount, r.maxTime, r.minDt, r.maxDt);
  };
  auto fdp = [&]() {
    allowed = false;
    inDynamic = true;
    int e = prepareQd2Fd(r);
    return e ? e : fdPhase(r);
  };
  int e = qdp();
  if (e) return e;
  if (currT >= r.maxTime || stepCount >= r.maxStepCount) return 0;
  if ((e = fdp())) return e;
  if (currT >= r.maxTime || stepCount >= r.maxStepCount || c.maxNumCycles <= 1) return 0;
  cycleCount++;
  while (cycleCount < c.maxNumCycles && stepCount <= r.maxStepCount && currT <= r.maxTime) {
    if ((e = qdp())) return e;
    if ((e = fdp())) return e;
    cycleCount++;
  }
  return 0;
}

// ---------------------------------------------------- power-law QD -----
int PowerLawModel::initPL(const double* T_, const double* A_, const double* B_, const double* n_,
                          double cap) {
  T.assign(T_, T_ + N);
  Ad.assign(A_, A_ + N);
  Bd.assign(B_, B_ + N);
  nd.assign(n_, n_ + N);
  effViscCap = cap;
  visc.assign(N, 0.0);
  gTxy.assign(N, 0.0); gTxz.assign(N, 0.0); sxz.assign(N, 0.0);
  effVisc.assign(N, 0.0); sat.assign(N, 0.0);
  const Sbp2D& s = sbp;
  // helpful factors yq*zr*H, yq*Hy, zr*Hz (powerLaw.cpp:1256-1268)
  Csr yqxzrxH, yqxHy, zrxHz;
  if (s.d.var_grid) {
    yqxzrxH = matmul3(s.yq, s.zr, s.H);
    yqxHy = matmul(s.yq, s.Hy_Iz);
    zrxHz = matmul(s.zr, s.Iy_Hz);
  } else {
    yqxHy = matmul(s.H, s.Iy_Hzinv);
    zrxHz = matmul(s.H, s.Hyinv_Iz);
    yqxzrxH = s.H;
  }
  // B = yq*zr*H * Dy * mu (+ Neumann SATs on L/R), C likewise with Dz (T/B)
  Bm = matmul3(yqxzrxH, s.Dy_Iz, s.muM);
  if (s.d.bcL == 1) Bm = axpy(Bm, 1.0, matmul3(zrxHz, s.muM, s.E0y_Iz));
  if (s.d.bcR == 1) Bm = axpy(Bm, -1.0, matmul3(zrxHz, s.muM, s.ENy_Iz));
  Cm = matmul3(yqxzrxH, s.Iy_Dz, s.muM);
  if (s.d.bcT == 1) Cm = axpy(Cm, 1.0, matmul3(yqxHy, s.muM, s.Iy_E0z));
  if (s.d.bcB == 1) Cm = axpy(Cm, -1.0, matmul3(yqxHy, s.muM, s.Iy_ENz));
  return 0;
}

void powerlaw_source(const PowerLawModel& m, const double* gxy, const double* gxz, double* out) {
  spmv(m.Bm, gxy, out);
  spmv_add(m.Cm, gxz, out);
}

void powerlaw_sat(const Sbp2D& s, int bcLDirichlet, int bcRDirichlet, const double* u,
                  const double* gL, const double* gR, double* out) {
  const int N = s.N();
  std::vector<double> temp1(N), G(N);
  std::fill(out, out + N, 0.0);
  if (bcLDirichlet) {
    // out += Hyinv E0y u - Hyinv e0y gL
    std::vector<double> t(N);
    spmv(s.E0y_Iz, u, t.data());
    spmv(s.Hyinv_Iz, t.data(), temp1.data());
    spmv(s.e0y_Iz, gL, t.data());
    spmv(s.Hyinv_Iz, t.data(), G.data());
    for (int i = 0; i < N; ++i) out[i] = out[i] + 1.0 * temp1[i];
    for (int i = 0; i < N; ++i) out[i] = out[i] + (-1.0) * G[i];
  }
  if (bcRDirichlet) {
    std::vector<double> t(N);
    spmv(s.ENy_Iz, u, t.data());
    spmv(s.Hyinv_Iz, t.data(), temp1.data());
    spmv(s.eNy_Iz, gR, t.data());
    spmv(s.Hyinv_Iz, t.data(), G.data());
    for (int i = 0; i < N; ++i) out[i] = out[i] + (-1.0) * temp1[i];
    for (int i = 0; i < N; ++i) out[i] = out[i] + 1.0 * G[i];
  }
  if (s.d.var_grid) {
    std::vector<double> t(out, out + N);
    spmv(s.qy, t.data(), out);
  }
}

// strikeSlip_powerLaw_qd.cpp:633-686 and solveMomentumBalance :787-823
int PowerLawModel::d_dt(double t, const VarMap& var, VarMap& dvar) {
  const auto& vslip = var.at("slip");
  for (int k = 0; k < Nz; ++k) bcL[k] = vslip[k] * (1.0 / c.faultTypeScale);
  if (c.remoteLoadingR)
    for (int k = 0; k < Nz; ++k) bcR[k] = c.vL * t / c.faultTypeScale + 1.0 * bcRShift[k];
  psi = var.at("psi");
  slip = vslip;
  const auto& gVxy = var.at("gVxy");
  const auto& gVxz = var.at("gVxz");
  powerlaw_source(*this, gVxy.data(), gVxz.data(), visc.data());
  sbp2d_set_rhs(sbp, rhs.data(), bcL.data(), bcR.data(), bcT.data(), bcB.data());
  for (int i = 0; i < N; ++i) rhs[i] = rhs[i] + 1.0 * visc[i];
  if (c.solver == 0) {
    chol.solve(rhs.data(), u.data());
  } else {
    PcgStats st = pcg(sbp.A, rhs.data(), u.data(), c.cg_rtol, c.cg_maxit);
    cgIters += st.iters;
    if (st.relres > c.cg_rtol) return 20;
  }
  nSolves++;
  // total strains, stresses, viscosity, strain rates (powerLaw.cpp:1707-1887)
  sbp2d_Dy(sbp, u.data(), gTxy.data());
  sbp2d_Dz(sbp, u.data(), gTxz.data());
  for (int i = 0; i < N; ++i) sxy[i] = (gTxy[i] + (-1.0) * gVxy[i]) * sbp.mu[i];
  for (int i = 0; i < N; ++i) sxz[i] = (gTxz[i] + (-1.0) * gVxz[i]) * sbp.mu[i];
  powerlaw_sat(sbp, sbp.d.bcL == 0, sbp.d.bcR == 0, u.data(), bcL.data(), bcR.data(), sat.data());
  auto& dgxy = dvar["gVxy"];
  auto& dgxz = dvar["gVxz"];
  dgxy.resize(N);
  dgxz.resize(N);
  powerlaw_rates(N, sxy.data(), sxz.data(), T.data(), Ad.data(), Bd.data(), nd.data(), sbp.mu.data(),
                 sat.data(), effViscCap, effVisc.data(), dgxy.data(), dgxz.data());
  // fault (Fault_qd::d_dt) on sxy at the fault
  for (int k = 0; k < Nz; ++k) tauQS[k] = sxy[k];
  for (int k = 0; k < Nz; ++k) tauQS[k] = tauQS[k] + 1.0 * prestress[k];
  if (computeVel_qd(Nz, eta.data(), tauQS.data(), sN.data(), psi.data(), a.data(), locked.data(),
                    c.v0, c.vL, c.rootTol, c.maxNumIts, c.jac_mode, V.data(), &rootIts))
    return 30;
  dvar["slip"] = V;
  auto& dpsi = dvar["psi"];
  dpsi.resize(Nz);
  for (int k = 0; k < Nz; ++k)
    dpsi[k] = c.stateLaw == 0 ? agingLaw_psi(psi[k], V[k], b[k], c.f0, c.v0, Dc[k])
                              : slipLaw_psi(psi[k], V[k], a[k], b[k], c.f0, c.v0, Dc[k]);
  for (int k = 0; k < Nz; ++k) tau[k] = -1.0 * (eta[k] * V[k]) + tauQS[k];
  for (int k = 0; k < Nz; ++k) strength[k] = strength_psi(sN[k], psi[k], V[k], a[k], c.v0);
  return 0;
}

// ---------------------------------------------------------- power law ----
void powerlaw_rates(int N, const double* sxy, const double* sxz, const double* T,
                    const double* A, const double* B, const double* n, const double* mu,
                    const double* sat, double effViscCap, double* effVisc, double* gxy_t,
                    double* gxz_t) {
  for (int i = 0; i < N; ++i) {
    double sdev = sxy[i] * sxy[i];
    if (sxz) sdev += sxz[i] * sxz[i];
    sdev = std::sqrt(std::fabs(sdev));
    double inv = 1e3 * A[i] * std::pow(sdev, n[i] - 1.0) * std::exp(-B[i] / T[i]);
    double ev = 1.0 / effViscCap;
    ev = ev + 1.0 * inv;
    ev = 1.0 / ev;
    effVisc[i] = ev;
    gxy_t[i] = (mu[i] * (sat ? sat[i] : 0.0) + 1.0 * sxy[i]) / ev;
    if (sxz && gxz_t) gxz_t[i] = sxz[i] / ev;
  }
}

}  // namespace orc
