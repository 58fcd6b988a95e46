// Device-side rate-and-state functions (FP64).
#pragma once
#include <hip/hip_runtime.h>

#include "scycle_hip.h"

namespace sc {

// a*sN*asinh(V/(2 v0) * exp(psi/a))  (fault.cpp:2246-2250)
__device__ __forceinline__ double strength_psi(double sN, double psi, double V, double a, double v0) {
  return a * sN * asinh((0.5 * V / v0) * exp(psi / a));
}

// dpsi/dt = (b v0/Dc)(exp((f0-psi)/b) - V/v0), 0 when b <= 1e-3 or the
// exponential overflows  (fault.cpp:1984-1995)
__device__ __forceinline__ double aging_law_psi(double psi, double V, double b, double f0, double v0,
                                                double Dc) {
  const double A = exp((f0 - psi) / b);
  double d = 0.;
  if (!isinf(A) && b > 1e-3) d = (b * v0 / Dc) * (A - V / v0);
  return d;
}

// dpsi/dt = -|V|/Dc (f - fss), f regularised, fss = f0 + (a-b) ln(|V|/v0)
// (fault.cpp:2072-2088)
__device__ __forceinline__ double slip_law_psi(double psi, double V, double a, double b, double f0,
                                               double v0, double Dc) {
  if (V == 0) return 0.0;
  const double absV = fabs(V);
  const double fss = f0 + (a - b) * log(absV / v0);
  const double f = a * asinh((absV / 2. / v0) * exp(psi / a));
  return -absV / Dc * (f - fss);
}

// Newton with bisection safeguard on a bracket (rootFinder.cpp:174-257);
// R(x, &f, &J) evaluates the residual and slope.  Returns 0 when |f| < atol.
template <class R>
__device__ int bracketed_newton(const R& resid, double left, double right, double x0, double atol,
                                int maxIts, double* out, int* its) {
  double f, fp, fLeft, fRight, dummy;
  resid(left, &fLeft, &dummy);
  resid(right, &fRight, &dummy);
  *its = 0;
  if (sqrt(fLeft * fLeft) <= atol) { *out = left; return 0; }
  if (sqrt(fRight * fRight) <= atol) { *out = right; return 0; }
  if (fLeft > 0) {
    double t = left; left = right; right = t;
    t = fLeft; fLeft = fRight; fRight = t;
  }
  int n = 0;
  double x = x0, dxOld = fabs(right - left), dx = dxOld;
  resid(x, &f, &fp);
  while (n <= maxIts && fabs(f) >= atol) {
    if (((x - right) * fp - f) * ((x - left) * fp - f) > 0.0 || fabs(2.0 * f) > fabs(dxOld * fp)) {
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
    ++n;
  }
  *out = x;
  *its = n;
  return fabs(f) > atol ? 1 : 0;
}

// Quasi-dynamic residual F(V) = strength(V) - (tauQS - eta V)
// (ComputeVel_qd::getResid, fault.cpp:1038-1059).  jac_mode 0 keeps the
// reference's slope expression A*V/sqrt(B^2V^2+1)+eta; 1 uses A*B/sqrt(..)+eta.
struct QdResid {
  double psi, a, sN, eta, tauQS, v0;
  int jac_mode;
  double ex;  // exp(psi / a), fixed over the root solve
  __device__ QdResid(double psi_, double a_, double sN_, double eta_, double tauQS_, double v0_, int jac)
      : psi(psi_), a(a_), sN(sN_), eta(eta_), tauQS(tauQS_), v0(v0_), jac_mode(jac), ex(exp(psi_ / a_)) {}
  __device__ void operator()(double V, double* f, double* J) const {
    const double strength = a * sN * asinh((0.5 * V / v0) * ex);  // strength_psi
    const double stress = tauQS - eta * V;
    *f = strength - stress;
    const double A = a * sN;
    const double B = ex / (2. * v0);
    const double den = sqrt(B * B * V * V + 1.);
    *J = (jac_mode == 0 ? A * V / den : A * B / den) + eta;
  }
};

// Fault_qd::d_dt at fault node k (fault.cpp:861-918) from the quasi-static
// shear stress tq0 (before the prestress): slip velocity from the root solve,
// state rate, shear stress and strength; returns V.
__device__ __forceinline__ double fault_qd_node(const sc_fault_qd_args& a, int k, double tq0) {
  const double tq = tq0 + 1.0 * (a.prestress_d ? a.prestress_d[k] : 0.0);
  a.tauQS_d[k] = tq;
  const double psi = a.psi_d[k], aa = a.a_d[k], sN = a.sN_d[k], eta = a.eta_d[k];
  const double lk = a.locked_d ? a.locked_d[k] : 0.0;
  double V = a.V_d[k];
  int bad = 0;
  if (lk > 0.5) {
    V = 0.;
  } else if (lk < -0.5) {
    V = a.vL;
  } else {
    const double left = 0., right = tq / eta;
    if (isnan(right)) {
      bad = SC_ERR_NONFINITE;
    } else if (fabs(left - right) < 1e-14) {
      V = left;
    } else {
      QdResid r{psi, aa, sN, eta, tq, a.v0, a.jac_mode};
      double out;
      int its;
      // the previous V starts the search; a non-finite one (left by rates of
      // a discarded attempt) would stall Newton, so such a start moves to the
      // middle of the bracket
      if (!isfinite(V)) V = 0.5 * (left + right);
      if (bracketed_newton(r, left, right, V, a.rootTol, a.maxNumIts, &out, &its)) bad = SC_ERR_NOCONV;
      V = out;
    }
  }
  a.V_d[k] = V;
  double dpsi = a.stateLaw == SC_AGING_LAW ? aging_law_psi(psi, V, a.b_d[k], a.f0, a.v0, a.Dc_d[k])
                                           : slip_law_psi(psi, V, aa, a.b_d[k], a.f0, a.v0, a.Dc_d[k]);
  if (!isfinite(dpsi)) bad = SC_ERR_NONFINITE;
  a.dpsi_d[k] = dpsi;
  a.tau_d[k] = -1.0 * (eta * V) + tq;
  a.strength_d[k] = strength_psi(sN, psi, V, aa, a.v0);
  if (bad && a.status_d) atomicOr(a.status_d, bad);
  return V;
}

}  // namespace sc
