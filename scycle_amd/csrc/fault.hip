// Fault-side kernels of the quasi-dynamic step: rate-and-state slip-velocity
// root solve, state evolution, shear stress and strength (Fault_qd::d_dt,
// fault.cpp:861-918), of the fully dynamic step (Fault_fd::d_dt,
// fault.cpp:1436-1523), plus the small vector kernels of the explicit
// Runge-Kutta driver (odeSolver.cpp) and the boundary-data update of
// StrikeSlip_LinearElastic_qd::d_dt (strikeSlip_linearElastic_qd.cpp:763-770).
// All of it is FP64: the argument of asinh reaches ~1e30.
#include <cmath>

#include "fault_dev.hpp"
#include "sc_internal.hpp"

namespace sc {

__global__ __launch_bounds__(256) void k_fault_qd(sc_fault_qd_args a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.N) return;
  fault_qd_node(a, k, a.tauQS_d[k]);
}

int fault_qd_rates(hipStream_t s, const sc_fault_qd_args& a) {
  if (a.N <= 0) return SC_OK;
  hipLaunchKernelGGL(k_fault_qd, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// ------------------------------------------------------------ FD fault --
// Fault_fd::d_dt for one node (fault.cpp:1436-1523, setPhi :1527-1575,
// ComputeVel_fd :1588-1685, ComputeAging_fd / ComputeSlipLaw_fd :1697-1879).
// uBody (may be null) receives the new fault displacement in the body array's
// type: the fault column of uNext (strikeSlip_linearElastic_fd.cpp:421-423).
template <typename TB, typename TD>
__global__ __launch_bounds__(256) void k_fault_fd(sc_fault_fd_args a, const TD* __restrict__ d2u,
                                                  TB* __restrict__ uBody) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.N) return;
  const double dt = a.dt, fts = a.faultTypeScale, v0 = a.v0;
  const double psiC = a.psi_d[k], psiP = a.psiPrev_d[k];
  const double aa = a.a_d[k], bb = a.b_d[k], Dc = a.Dc_d[k], sN = a.sNEff_d[k];
  const double rho = a.rho_d[k], alphay = a.alphay_d[k], slip0 = a.slip0_d[k];
  int bad = 0;
  // u, uPrev from slip: (slip - slip0) / faultTypeScale
  const double uPrev = (-1.0 * slip0 + a.slipPrev_d[k]) * (1.0 / fts);
  double u = (-1.0 * slip0 + a.slip_d[k]) * (1.0 / fts);
  // setPhi
  const double an = (double)d2u[k] + a.tau0_d[k] / alphay;
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
  if (uBody) uBody[k] = (TB)u;
  if (!isfinite(V) || !isfinite(psiN) || !isfinite(u)) bad = SC_ERR_NONFINITE;
  if (bad && a.status_d) atomicOr(a.status_d, bad);
}

template <typename TB, typename TD>
int fault_fd_step_t(hipStream_t s, const sc_fault_fd_args& a, const TD* d2u, TB* uBody) {
  if (a.N <= 0) return SC_OK;
  hipLaunchKernelGGL((k_fault_fd<TB, TD>), dim3((a.N + 255) / 256), dim3(256), 0, s, a, d2u, uBody);
  SC_HIP(hipGetLastError());
  return SC_OK;
}
int fault_fd_step64(hipStream_t s, const sc_fault_fd_args& a, double* uBody) {
  return fault_fd_step_t<double, double>(s, a, a.d2u_d, uBody);
}
int fault_fd_step32(hipStream_t s, const sc_fault_fd_args& a, const float* d2u32, float* uBody) {
  return fault_fd_step_t<float, float>(s, a, d2u32, uBody);
}

// ------------------------------------------------------ QD driver helpers
// bcL = slip / fts ; bcR = vL*t/fts + bcRShift
__global__ void k_qd_bc(int Nz, const double* __restrict__ slip, const double* __restrict__ shift,
                        double t, double vL, double fts, int remote, double* __restrict__ bcL,
                        double* __restrict__ bcR) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Nz) return;
  bcL[k] = slip[k] * (1.0 / fts);
  if (remote) bcR[k] = vL * t / fts + 1.0 * shift[k];
}

// Error measures over one vector pair (lo, hi): [sum (lo-hi)^2, sum hi^2, max |lo-hi|/hi]
__global__ __launch_bounds__(256) void k_err(int n, const double* __restrict__ lo,
                                             const double* __restrict__ hi, double* __restrict__ out) {
  __shared__ double s0[256], s1[256], s2[256];
  double a = 0, b = 0, m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double e = -1.0 * hi[i] + lo[i];
    a += e * e;
    b += hi[i] * hi[i];
    m = fmax(m, fabs(e) / hi[i]);
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
    out[2] = s2[0];
  }
}

int launch_qd_bc(hipStream_t s, int Nz, const double* slip, const double* shift, double t,
                 double vL, double fts, int remote, double* bcL, double* bcR) {
  hipLaunchKernelGGL(k_qd_bc, dim3((Nz + 255) / 256), dim3(256), 0, s, Nz, slip, shift, t, vL, fts,
                     remote, bcL, bcR);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// Long vectors (the viscous strains, N entries): per-block partial sums, then
// one block combines them in block order (deterministic).
constexpr int kErrBlocksMax = 512;

__global__ __launch_bounds__(256) void k_err_part(int n, const double* __restrict__ lo,
                                                  const double* __restrict__ hi,
                                                  double* __restrict__ part) {
  __shared__ double s0[256], s1[256], s2[256];
  double a = 0, b = 0, m = -INFINITY;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double e = -1.0 * hi[i] + lo[i];
    a += e * e;
    b += hi[i] * hi[i];
    m = fmax(m, fabs(e) / hi[i]);
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s0[0];
    part[3 * blockIdx.x + 1] = s1[0];
    part[3 * blockIdx.x + 2] = s2[0];
  }
}

__global__ void k_err_fin(int nb, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0, b = 0, m = -INFINITY;
  for (int i = 0; i < nb; ++i) {
    a += part[3 * i];
    b += part[3 * i + 1];
    m = fmax(m, part[3 * i + 2]);
  }
  out[0] = a;
  out[1] = b;
  out[2] = m;
}

int launch_err(hipStream_t s, int n, const double* lo, const double* hi, double* out3,
               double* partials) {
  if (n <= 16384 || !partials) {
    hipLaunchKernelGGL(k_err, dim3(1), dim3(256), 0, s, n, lo, hi, out3);
  } else {
    int nb = (n + 1023) / 1024;
    if (nb > kErrBlocksMax) nb = kErrBlocksMax;
    hipLaunchKernelGGL(k_err_part, dim3(nb), dim3(256), 0, s, n, lo, hi, partials);
    hipLaunchKernelGGL(k_err_fin, dim3(1), dim3(64), 0, s, nb, partials, out3);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc

extern "C" int sc_fault_fd_step(sc_ctx* h, const sc_fault_fd_args* args) {
  if (!h || !args) return sc::fail(SC_ERR_ARG, "sc_fault_fd_step: null argument");
  const sc_fault_fd_args& a = *args;
  if (!a.a_d || !a.b_d || !a.Dc_d || !a.sNEff_d || !a.rho_d || !a.alphay_d || !a.tau0_d ||
      !a.slip0_d || !a.d2u_d || !a.psi_d || !a.psiPrev_d || !a.slip_d || !a.slipPrev_d || !a.V_d ||
      !a.psiNext_d || !a.slipNext_d)
    return sc::fail(SC_ERR_ARG, "sc_fault_fd_step: null array");
  if (a.stateLaw != SC_AGING_LAW && a.stateLaw != SC_SLIP_LAW)
    return sc::fail(SC_ERR_ARG, "sc_fault_fd_step: stateLaw must be aging or slip");
  if (!(a.dt > 0)) return sc::fail(SC_ERR_ARG, "sc_fault_fd_step: dt must be > 0");
  return sc::fault_fd_step64(h->c.stream, a, nullptr);
}

extern "C" int sc_fault_qd_rates(sc_ctx* h, const sc_fault_qd_args* args) {
  if (!h || !args) return sc::fail(SC_ERR_ARG, "sc_fault_qd_rates: null argument");
  const sc_fault_qd_args& a = *args;
  if (!a.eta_d || !a.sN_d || !a.a_d || !a.b_d || !a.Dc_d || !a.psi_d || !a.tauQS_d || !a.V_d ||
      !a.dpsi_d || !a.tau_d || !a.strength_d)
    return sc::fail(SC_ERR_ARG, "sc_fault_qd_rates: null array");
  if (a.stateLaw != SC_AGING_LAW && a.stateLaw != SC_SLIP_LAW)
    return sc::fail(SC_ERR_ARG, "sc_fault_qd_rates: stateLaw must be aging or slip");
  return sc::fault_qd_rates(h->c.stream, a);
}
