// Fault-side kernels of the quasi-dynamic step: rate-and-state slip-velocity
// root solve, state evolution, shear stress and strength (Fault_qd::d_dt,
// fault.cpp:861-918), of the fully dynamic step (Fault_fd::d_dt,
// fault.cpp:1436-1523), plus the small vector kernels of the explicit
// Runge-Kutta driver (odeSolver.cpp).  All of it is FP64: the argument of
// asinh reaches ~1e30.
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
// Fault_fd::d_dt, one thread per node (the node's update is fault_fd_node;
// the FD driver runs it inside its wave-update launch instead).  uBody (may
// be null) receives the new fault displacement: the fault column of uNext
// (strikeSlip_linearElastic_fd.cpp:421-423).
__global__ __launch_bounds__(256) void k_fault_fd(sc_fault_fd_args a, const double* __restrict__ d2u,
                                                  double* __restrict__ uBody) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.N) return;
  const double u = fault_fd_node(a, k, d2u[k]);
  if (uBody) uBody[k] = u;
}

int fault_fd_step64(hipStream_t s, const sc_fault_fd_args& a, double* uBody) {
  if (a.N <= 0) return SC_OK;
  hipLaunchKernelGGL(k_fault_fd, dim3((a.N + 255) / 256), dim3(256), 0, s, a, a.d2u_d, uBody);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// ------------------------------------------------------ QD driver helpers
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
