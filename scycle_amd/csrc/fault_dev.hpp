// Device-side rate-and-state functions (FP64).
#pragma once
#include <hip/hip_runtime.h>

#include "scycle_hip.h"

namespace sc {

// 2 atanh(s) for |s| <= 0.23 by its Taylor series 2 (s + s^3/3 + s^5/5 + ...)
// through s^25 (the first omitted term is below 2^-60 of the sum there).
__device__ __forceinline__ double two_atanh_small(double s) {
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  p = fma(p, s2, 1.0 / 23.0);
  p = fma(p, s2, 1.0 / 21.0);
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double t = 2.0 * s;
  return fma(t * s2, p, t);
}

// log x for finite x > 0: x = 2^k m with m in [sqrt(1/2), sqrt(2)), log m =
// 2 atanh((m - 1) / (m + 1)) with |(m - 1) / (m + 1)| <= 0.172 (m - 1 is exact),
// and k ln 2 from a two-part ln 2 (the high part has 32 significant bits, so
// that k times it is exact).
__device__ __forceinline__ double log_pos(double x) {
  int k;
  double m = frexp(x, &k);  // m in [0.5, 1)
  const bool lo = m < 0.70710678118654752440;
  m = lo ? 2.0 * m : m;
  k = lo ? k - 1 : k;
  const double r = two_atanh_small((m - 1.0) / (m + 1.0));
  constexpr double kLn2Hi = 0x1.62e42ffp-1;             // ln 2 rounded to 32 bits
  constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;      // ln 2 - kLn2Hi
  const double kd = (double)k;
  return fma(kd, kLn2Hi, fma(kd, kLn2Lo, r));
}

// asinh z from asinh z = log(z + sqrt(z^2 + 1)): for |z| >= 1/2 with log_pos
// (log(2|z|) once z^2 + 1 rounds to z^2, from 2^28 on), below 1/2 as log1p(t),
// t = |z| + z^2 / (1 + sqrt(1 + z^2)), = 2 atanh(t / (2 + t)) with
// t / (2 + t) <= 0.23, so that small arguments keep their relative accuracy.
// Odd in z.  About a fifth of the instructions of the library's asinh, to a
// few units in the last place.
__device__ __forceinline__ double asinh_qd(double z) {
  const double az = fabs(z);
  double r;
  if (az < 0.5) {
    const double q = sqrt(fma(az, az, 1.0));
    const double t = az + (az * az) / (1.0 + q);
    r = two_atanh_small(t / (2.0 + t));
  } else {
    const double w = az > 268435456.0 ? 2.0 * az : az + sqrt(fma(az, az, 1.0));
    r = log_pos(w);
  }
  return copysign(r, z);
}

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

// Quasi-dynamic residual F(V) = strength(V) - (tauQS - eta V)
// (ComputeVel_qd::getResid, fault.cpp:1038-1059).  jac_mode 0 keeps the
// reference's slope expression A*V/sqrt(B^2V^2+1)+eta; 1 uses A*B/sqrt(..)+eta.
//
// Evaluated with the per-solve constants A = a sN and B = exp(psi/a) / (2 v0)
// formed once, z = B V, the library-free asinh_qd above and the slope's
// 1/sqrt(z^2 + 1) as one reciprocal square root: the same function and slope
// to a few units in the last place, in about a third of the instructions of
// the direct form, which matters because the fault waves' Newton chains are
// the critical path of a QD stage.
struct QdResid {
  double psi, a, sN, eta, tauQS, v0;
  int jac_mode;
  double A, B;  // a sN and exp(psi / a) / (2 v0), fixed over the root solve
  // the strength at the last point evaluated, so the caller need not
  // evaluate it again
  mutable double strength = 0.0;
  __device__ QdResid(double psi_, double a_, double sN_, double eta_, double tauQS_, double v0_, int jac)
      : psi(psi_), a(a_), sN(sN_), eta(eta_), tauQS(tauQS_), v0(v0_), jac_mode(jac), A(a_ * sN_),
        B(exp(psi_ / a_) / (2. * v0_)) {}
  __device__ double strength_at(double V) const { return A * asinh_qd(B * V); }
  __device__ void operator()(double V, double* f, double* J) const {
    const double z = B * V;
    strength = A * asinh_qd(z);
    const double stress = tauQS - eta * V;
    *f = strength - stress;
    const double inv = rsqrt(fma(z, z, 1.0));
    *J = (jac_mode == 0 ? A * V : A * B) * inv + eta;
  }
  // residual only, at a bracket end (leaves `strength` alone)
  __device__ void at_bracket_end(double V, double* f) const { *f = strength_at(V) - (tauQS - eta * V); }
};

// the residual at a bracket end: the residual functor, or QdResid's cheaper
// form
template <class R>
__device__ __forceinline__ void bracket_end_resid(const R& r, double x, double* f) {
  double dummy;
  r(x, f, &dummy);
}
__device__ __forceinline__ void bracket_end_resid(const QdResid& r, double x, double* f) {
  r.at_bracket_end(x, f);
}

// Newton with bisection safeguard on a bracket (rootFinder.cpp:174-257);
// R(x, &f, &J) evaluates the residual and slope.  Returns 0 when |f| < atol.
template <class R>
__device__ int bracketed_newton(const R& resid, double left, double right, double x0, double atol,
                                int maxIts, double* out, int* its) {
  double f, fp, fLeft, fRight;
  // the residuals at both bracket ends and at the start are independent:
  // evaluated together, their latency chains overlap (the start's is used
  // only when neither end is already a root, as in the reference)
  bracket_end_resid(resid, left, &fLeft);
  bracket_end_resid(resid, right, &fRight);
  resid(x0, &f, &fp);
  *its = 0;
  if (sqrt(fLeft * fLeft) <= atol) { *out = left; return 0; }
  if (sqrt(fRight * fRight) <= atol) { *out = right; return 0; }
  if (fLeft > 0) {
    double t = left; left = right; right = t;
    t = fLeft; fLeft = fRight; fRight = t;
  }
  int n = 0;
  double x = x0, dxOld = fabs(right - left), dx = dxOld;
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

// The per-node inputs of Fault_qd::d_dt that do not depend on the stage's
// stress, loaded ahead of the work that forms the stress.
struct FaultNodeIn {
  double psi, a, sN, eta, lk, V, pre, b, Dc;
};
__device__ __forceinline__ FaultNodeIn fault_node_load(const sc_fault_qd_args& a, int k, double psi) {
  FaultNodeIn in;
  in.psi = psi;
  in.a = a.a_d[k];
  in.sN = a.sN_d[k];
  in.eta = a.eta_d[k];
  in.lk = a.locked_d ? a.locked_d[k] : 0.0;
  in.V = a.V_d[k];
  in.pre = a.prestress_d ? a.prestress_d[k] : 0.0;
  in.b = a.b_d[k];
  in.Dc = a.Dc_d[k];
  return in;
}

// Fault_qd::d_dt at one fault node (fault.cpp:861-918) from the quasi-static
// shear stress tq0 (before the prestress) and the node's loaded inputs: slip
// velocity from the root solve, state rate, shear stress and strength, formed
// in registers (fault_qd_eval) and then written (fault_qd_commit), so that a
// caller can form them before it knows whether it will keep them.
struct FaultNodeOut {
  double tq, V, dpsi, tau, strength;
  int bad;
};
__device__ __forceinline__ FaultNodeOut fault_qd_eval(const sc_fault_qd_args& a, double tq0,
                                                      const FaultNodeIn& in) {
  FaultNodeOut o;
  const double tq = tq0 + 1.0 * in.pre;
  o.tq = tq;
  const double psi = in.psi, aa = in.a, sN = in.sN, eta = in.eta;
  const double lk = in.lk;
  double V = in.V;
  int bad = 0;
  // the aging law's exponential does not depend on V: formed beside the
  // root solve
  const double agingA = a.stateLaw == SC_AGING_LAW ? exp((a.f0 - psi) / in.b) : 0.0;
  double strength = 0.0;
  bool haveStrength = false;
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
      // the last point the solve evaluated is the root unless it stopped at a
      // bracket end
      if (out != left && out != right) {
        strength = r.strength;
        haveStrength = true;
      }
    }
  }
  o.V = V;
  double dpsi;
  if (a.stateLaw == SC_AGING_LAW) {  // aging_law_psi with its exponential formed above
    dpsi = 0.;
    if (!isinf(agingA) && in.b > 1e-3) dpsi = (in.b * a.v0 / in.Dc) * (agingA - V / a.v0);
  } else {
    dpsi = slip_law_psi(psi, V, aa, in.b, a.f0, a.v0, in.Dc);
  }
  if (!isfinite(dpsi)) bad = SC_ERR_NONFINITE;
  o.dpsi = dpsi;
  o.tau = -1.0 * (eta * V) + tq;
  o.strength = haveStrength ? strength : strength_psi(sN, psi, V, aa, a.v0);
  o.bad = bad;
  return o;
}
// writes the node's outputs; returns V
__device__ __forceinline__ double fault_qd_commit(const sc_fault_qd_args& a, int k, const FaultNodeOut& o) {
  a.tauQS_d[k] = o.tq;
  a.V_d[k] = o.V;
  a.dpsi_d[k] = o.dpsi;
  a.tau_d[k] = o.tau;
  a.strength_d[k] = o.strength;
  if (o.bad && a.status_d) atomicOr(a.status_d, o.bad);
  return o.V;
}
__device__ __forceinline__ double fault_qd_node_in(const sc_fault_qd_args& a, int k, double tq0,
                                                   const FaultNodeIn& in) {
  return fault_qd_commit(a, k, fault_qd_eval(a, tq0, in));
}

// the same from memory: the node's inputs loaded here
__device__ __forceinline__ double fault_qd_node(const sc_fault_qd_args& a, int k, double tq0) {
  return fault_qd_node_in(a, k, tq0, fault_node_load(a, k, a.psi_d[k]));
}

// Fault_fd::d_dt for one node (fault.cpp:1436-1523, setPhi :1527-1575,
// ComputeVel_fd :1588-1685, ComputeAging_fd / ComputeSlipLaw_fd :1697-1879)
// from D2u at the node (d2uk): slip velocity, fault displacement with the
// frictional damping, implicit state update, slip and strength; a failed or
// non-finite solve sets a.status_d.  Returns the new fault displacement u.
__device__ __forceinline__ double fault_fd_node(const sc_fault_fd_args& a, int k, double d2uk) {
  const double dt = a.dt, fts = a.faultTypeScale, v0 = a.v0;
  const double psiC = a.psi_d[k], psiP = a.psiPrev_d[k];
  const double aa = a.a_d[k], bb = a.b_d[k], Dc = a.Dc_d[k], sN = a.sNEff_d[k];
  const double rho = a.rho_d[k], alphay = a.alphay_d[k], slip0 = a.slip0_d[k];
  int bad = 0;
  // u, uPrev from slip: (slip - slip0) / faultTypeScale
  const double uPrev = (-1.0 * slip0 + a.slipPrev_d[k]) * (1.0 / fts);
  double u = (-1.0 * slip0 + a.slip_d[k]) * (1.0 / fts);
  // setPhi
  const double an = d2uk + a.tau0_d[k] / alphay;
  const double Phi = 2.0 / dt * (u - uPrev) + dt * an / rho;
  const double fricPen = dt / alphay / rho;
  // slip velocity: fricPen*strength(V) = |Phi| - V on [0, |Phi|]
  const double lk = a.locked_d ? a.locked_d[k] : 0.0;
  double V;
  if (lk > 0.5) {
    V = 0.;
  } else if (lk < -0.5) {
    V = a.vL;
  } else {
    const double left = 0., right = fabs(Phi);
    if (fabs(left - right) < 1e-14) {
      V = left;
    } else {
      const double P = Phi < 0 ? -Phi : Phi;
      const double A = aa * sN, B = exp(psiC / aa) / (2. * v0);
      // the strength A asinh(B V) with the library-free asinh_qd and the
      // slope's 1/sqrt as one reciprocal square root (as the QD residual)
      auto resid = [&](double vel, double* f, double* J) {
        const double z = B * vel;
        *f = fricPen * (A * asinh_qd(z)) - (P - vel);
        *J = 1 + fricPen * A * B * rsqrt(fma(z, z, 1.0));
      };
      int its;
      if (bracketed_newton(resid, left, right, fabs(a.V_d[k]), a.rootTol, a.maxNumIts, &V, &its))
        bad = SC_ERR_NOCONV;
    }
  }
  // fault displacement with the frictional damping of the slip velocity
  if (!(V < 1e-14)) {
    const double fric = aa * sN * asinh_qd(exp(psiC / aa) / (2. * v0) * V);
    const double alpha = 1.0 / (rho * alphay) * fric / V;
    const double Aden = 1.0 + alpha * dt;
    V = Phi / (1. + dt * alpha);
    u = (2. * u + (an * dt * dt / rho) + (dt * alpha - 1.) * uPrev) / Aden;
  }
  // state: psiNext - psiPrev = 2 dt G((psiPrev + psiNext)/2, V) on [-2, 2]
  double psiN;
  {
    // one exponential per evaluation, shared by the residual and its slope;
    // the slip law's steady-state friction and |V| terms fixed over the solve
    const double absV = fabs(V);
    const double fss = absV > 0.0 ? a.f0 + (aa - bb) * log(absV / v0) : 0.0;
    const double xV = absV / 2. / v0, gV = absV / Dc;
    auto resid = [&](double state, double* f, double* J) {
      const double mid = (psiP + state) / 2.0;
      if (a.stateLaw == SC_AGING_LAW) {
        // aging_law_psi: zero where b <= 1e-3 or the exponential overflows
        const double E = exp((a.f0 - mid) / bb);
        const double G = (!isinf(E) && bb > 1e-3) ? (bb * v0 / Dc) * (E - V / v0) : 0.0;
        *f = -2 * dt * G + (state - psiP);
        *J = 1 + dt * v0 / Dc * E;
      } else {
        // slip_law_psi: zero at V = 0
        const double x = xV * exp(mid / aa);
        const double G = absV > 0.0 ? -gV * (aa * asinh_qd(x) - fss) : 0.0;
        *f = -2 * dt * G + (state - psiP);
        *J = 1 + dt * gV * x * rsqrt(fma(x, x, 1.0));
      }
    };
    int its;
    if (bracketed_newton(resid, -2., 2., psiC, a.rootTol, a.maxNumIts, &psiN, &its)) bad = SC_ERR_NOCONV;
  }
  const double slipN = fts * u + slip0;
  a.V_d[k] = V;
  a.psiNext_d[k] = psiN;
  a.slipNext_d[k] = slipN;
  if (a.strength_d) a.strength_d[k] = strength_psi(sN, psiN, V, aa, v0);
  if (a.uFault_d) a.uFault_d[k] = u;
  if (!isfinite(V) || !isfinite(psiN) || !isfinite(u)) bad = SC_ERR_NONFINITE;
  if (bad && a.status_d) atomicOr(a.status_d, bad);
  return u;
}

}  // namespace sc
