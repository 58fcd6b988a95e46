// Device-side rate-and-state functions (FP64).
#pragma once
#include <hip/hip_runtime.h>

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
  __device__ void operator()(double V, double* f, double* J) const {
    const double strength = strength_psi(sN, psi, V, a, v0);
    const double stress = tauQS - eta * V;
    *f = strength - stress;
    const double A = a * sN;
    const double B = exp(psi / a) / (2. * v0);
    const double den = sqrt(B * B * V * V + 1.);
    *J = (jac_mode == 0 ? A * V / den : A * B / den) + eta;
  }
};

}  // namespace sc
