// Power-law (dislocation creep) viscoelastic terms of the quasi-dynamic step
// (PowerLaw, powerLaw.cpp; StrikeSlip_PowerLaw_qd::solveMomentumBalance,
// strikeSlip_powerLaw_qd.cpp:787-823):
//   k_pl_source   rhs += B gVxy + C gVxz, the viscous-strain source of the
//                 momentum balance (computeViscStrainSourceTerms :1651-1672 with
//                 B, C from initializeMomBalMats :1239-1318)
//   k_pl_strain   total strains Dy u, Dz u -> stresses -> effective viscosity ->
//                 viscous strain rates, one pass per node (computeTotalStrains,
//                 computeStresses, computeViscosity, computeViscStrainRates and
//                 the SAT of computeViscousStrainRateSAT, :1674-1887), plus the
//                 shear stress on the fault and Fault_qd::d_dt there; the
//                 creep factor 1e3 A exp(-B/T) comes formed once per mediator
//                 (k_pl_factor)
//   k_pl_rates    the pointwise update alone (sigma -> eta, strain rates) in
//                 FP64 or with FP32 arrays
// The creep law 1/eta = 1e3 A s^(n-1) exp(-B/T) and the capped sum are
// evaluated in FP64 in every variant (exponents reach e^-50).
#include <cmath>

#include "cg_dev.hpp"
#include "fault_dev.hpp"
#include "sbp_dev.hpp"

namespace sc {

// eta from the deviatoric stress, and the strain rates (powerLaw.cpp:459,
// 1690-1700, 1718-1728)
__device__ __forceinline__ void pl_rates(double sxy, double sxz, double T, double A, double B,
                                         double n, double mu, double sat, double cap,
                                         double* ev_out, double* gxy_t, double* gxz_t) {
  double sdev = sxy * sxy;
  sdev = sdev + 1.0 * (sxz * sxz);
  sdev = sqrt(fabs(sdev));
  const double inv = 1e3 * A * pow(sdev, n - 1.0) * exp(-B / T);
  double ev = 1.0 / cap;
  ev = ev + 1.0 * inv;
  ev = 1.0 / ev;
  *ev_out = ev;
  *gxy_t = (mu * sat + 1.0 * sxy) / ev;
  *gxz_t = sxz / ev;
}

template <typename TS>
__global__ __launch_bounds__(256) void k_pl_rates(int N, const TS* __restrict__ sxy,
                                                  const TS* __restrict__ sxz, const TS* __restrict__ T,
                                                  const TS* __restrict__ A, const TS* __restrict__ B,
                                                  const TS* __restrict__ n, const TS* __restrict__ mu,
                                                  const TS* __restrict__ sat, double cap,
                                                  TS* __restrict__ ev, TS* __restrict__ gxy_t,
                                                  TS* __restrict__ gxz_t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double e, gy, gz;
  pl_rates((double)sxy[i], sxz ? (double)sxz[i] : 0.0, (double)T[i], (double)A[i], (double)B[i],
           (double)n[i], (double)mu[i], sat ? (double)sat[i] : 0.0, cap, &e, &gy, &gz);
  if (ev) ev[i] = (TS)e;
  gxy_t[i] = (TS)gy;
  if (gxz_t) gxz_t[i] = (TS)gz;
}

// D1 of a direction at line index i applied to f: the interior row in its
// factored form ia (f[i+1] - f[i-1]) + ib (f[i+2] - f[i-2]) (four loads, no
// table), the boundary rows from the table (clamped loads)
template <class F>
__device__ __forceinline__ double d1_line(const DevDir<double>& d, int n, int i, const F& f) {
  if (i >= 4 && i < n - 4) return d.ia * (f(i + 1) - f(i - 1)) + d.ib * (f(i + 2) - f(i - 2));
  const double* __restrict__ D1 = d.D1;
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < kBand; ++q) {
    const int m = i + q - 4;
    const double v = f(min(max(m, 0), n - 1));
    const double c = D1[i * kBand + q];
    s += (m >= 0 && m < n) ? c * v : 0.0;
  }
  return s;
}

// rhs = setRhs(boundary data) + B gxy + C gxz, B = yq zr H Dy mu (+ Neumann
// SATs on L/R), C likewise with Dz (T/B): the boundary-data update of the
// mediator's d_dt (bcL = slip/fts, bcR = vL t/fts + shift when remotely
// loaded; strikeSlip_powerLaw_qd.cpp d_dt), SbpOps::setRhs and the viscous
// source in one pass.  Every thread forms the boundary data it reads from the
// stage's slip (no thread reads what another writes); the j = 0 row writes
// bcL and bcR for the later kernels.  The same arithmetic as the boundary
// kernel, setRhs and the source term launched one after the other.
__global__ __launch_bounds__(256) void k_pl_source(DevOp<double> p, const double* __restrict__ gxy,
                                                   const double* __restrict__ gxz,
                                                   double* __restrict__ rhs, PlRhsArgs b) {
  const int Ny = p.Ny, Nz = p.Nz;
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= (long long)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I - (long long)j * Nz);
  auto bcL = [&](int i) { return b.slip[i] * (1.0 / b.fts); };
  auto bcR = [&](int i) { return b.remote ? b.vL * b.t / b.fts + 1.0 * b.shift[i] : b.bcR[i]; };
  const double r0 = rhs_node(p, j, k, [&](int side, int i) {
    return side == 0 ? bcL(i) : side == 1 ? bcR(i) : side == 2 ? b.bcT[i] : b.bcB[i];
  });
  const double* mu = p.mu;
  const double wgt = p.y.met[j] * p.z.met[k] * p.y.H[j] * p.z.H[k];
  const double sy = d1_line(p.y, Ny, j, [&](int m) { return mu[(size_t)m * Nz + k] * gxy[(size_t)m * Nz + k]; });
  const double sz = d1_line(p.z, Nz, k, [&](int m) { return mu[(size_t)j * Nz + m] * gxz[(size_t)j * Nz + m]; });
  double v = wgt * (p.y.imet[j] * sy) + wgt * (p.z.imet[k] * sz);
  const double muI = mu[I];
  if (j == 0 && p.bcL == SC_NEUMANN) v += p.z.met[k] * p.z.H[k] * muI * gxy[I];
  if (j == Ny - 1 && p.bcR == SC_NEUMANN) v -= p.z.met[k] * p.z.H[k] * muI * gxy[I];
  if (k == 0 && p.bcT == SC_NEUMANN) v += p.y.met[j] * p.y.H[j] * muI * gxz[I];
  if (k == Nz - 1 && p.bcB == SC_NEUMANN) v -= p.y.met[j] * p.y.H[j] * muI * gxz[I];
  rhs[I] = r0 + 1.0 * v;
  if (j == 0) {
    b.bcLOut[k] = bcL(k);
    if (b.remote) b.bcR[k] = bcR(k);
  }
}

// K = 1e3 A exp(-B/T), the dislocation-creep factor that does not change
// with the stress (formed once per mediator)
__global__ __launch_bounds__(256) void k_pl_factor(long long N, const double* __restrict__ A,
                                                   const double* __restrict__ B, const double* __restrict__ T,
                                                   double* __restrict__ K) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I < N) K[I] = (1e3 * A[I]) * exp(-B[I] / T[I]);
}

int launch_pl_factor(const Ctx& c, const double* A, const double* B, const double* T, double* K, hipStream_t s) {
  const long long N = (long long)c.N;
  hipLaunchKernelGGL(k_pl_factor, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, A, B, T, K);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// Strains, stresses, viscosity and strain rates in one pass per node, with
// the creep factor K precomputed: 1/eta = 1/cap + K s^(n-1).
struct PlStrainArgs {
  const double *u, *gVxy, *gVxz, *K, *n, *bcL, *bcR;
  double cap;
  double *dgxy, *dgxz, *tauQS, *effVisc, *sxyOut;
  // Fault_qd::d_dt fused in at the fault nodes (fault = 1): the thread that
  // forms sigma_xy at (0, k) runs the node's root solve with it
  int fault;
  sc_fault_qd_args fa;
  double* dslip;
  // the solve's deferred guess check (sc non-null): one more block
  PendingFinish fin;
};

__global__ __launch_bounds__(256) void k_pl_strain(DevOp<double> p, PlStrainArgs a, int nGrid) {
  if ((int)blockIdx.x >= nGrid) {
    finish_irres_block(a.fin);
    return;
  }
  const int Ny = p.Ny, Nz = p.Nz;
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= (long long)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I - (long long)j * Nz);
  const double* u = a.u;
  const double gTxy = p.y.imet[j] * d1_line(p.y, Ny, j, [&](int m) { return u[(size_t)m * Nz + k]; });
  const double gTxz = p.z.imet[k] * d1_line(p.z, Nz, k, [&](int m) { return u[(size_t)j * Nz + m]; });
  const double muI = p.mu[I];
  const double sxy = (gTxy + (-1.0) * a.gVxy[I]) * muI;
  const double sxz = (gTxz + (-1.0) * a.gVxz[I]) * muI;
  // SAT of the viscous strain rate on Dirichlet y-sides, scaled by qy
  double sat = 0.0;
  if (j == 0 && p.bcL == SC_DIRICHLET) {
    const double hi = p.y.Hinv[0];
    sat = sat + 1.0 * (hi * u[I]);
    sat = sat + (-1.0) * (hi * a.bcL[k]);
  }
  if (j == Ny - 1 && p.bcR == SC_DIRICHLET) {
    const double hi = p.y.Hinv[Ny - 1];
    sat = sat + (-1.0) * (hi * u[I]);
    sat = sat + 1.0 * (hi * a.bcR[k]);
  }
  sat = p.y.imet[j] * sat;
  double sdev = sxy * sxy;
  sdev = sdev + 1.0 * (sxz * sxz);
  sdev = sqrt(fabs(sdev));
  const double inv = a.K[I] * pow(sdev, a.n[I] - 1.0);
  double ev = 1.0 / a.cap;
  ev = ev + 1.0 * inv;
  ev = 1.0 / ev;
  const double gy = (muI * sat + 1.0 * sxy) / ev;
  const double gz = sxz / ev;
  a.dgxy[I] = gy;
  a.dgxz[I] = gz;
  if (a.effVisc) a.effVisc[I] = ev;
  if (a.sxyOut) a.sxyOut[I] = sxy;
  if (j == 0) {
    a.tauQS[k] = sxy;
    if (a.fault) a.dslip[k] = fault_qd_node(a.fa, k, sxy);
  }
}

int launch_pl_source(const Ctx& c, const double* gxy, const double* gxz, double* rhs, const PlRhsArgs& b,
                     hipStream_t s) {
  const long long N = (long long)c.N;
  hipLaunchKernelGGL(k_pl_source, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, gxy, gxz, rhs, b);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// The 2-D output fields of PowerLaw::writeStep2D (powerLaw.cpp:2312-2372) at
// the current state: total strains (computeTotalStrains :1806-1820),
// stresses and their deviatoric norm (computeStresses :1823-1887), the
// dislocation-creep inverse viscosity and effective viscosity
// (computeViscosity :1674-1700) and the deviatoric strain rate of the
// dislocation creep (computeDevViscStrainRates :1919-1938).
__global__ __launch_bounds__(256) void k_pl_fields(DevOp<double> p, PlFields f) {
  const int Ny = p.Ny, Nz = p.Nz;
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= (long long)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I - (long long)j * Nz);
  const double* u = f.u;
  const double gTxy = p.y.imet[j] * d1_line(p.y, Ny, j, [&](int m) { return u[(size_t)m * Nz + k]; });
  const double gTxz = p.z.imet[k] * d1_line(p.z, Nz, k, [&](int m) { return u[(size_t)j * Nz + m]; });
  const double muI = p.mu[I];
  const double sxy = (gTxy + (-1.0) * f.gVxy[I]) * muI;
  const double sxz = (gTxz + (-1.0) * f.gVxz[I]) * muI;
  double sdev = sxy * sxy;
  sdev = sdev + 1.0 * (sxz * sxz);
  sdev = sqrt(fabs(sdev));
  const double inv = 1e3 * f.A[I] * pow(sdev, f.n[I] - 1.0) * exp(-f.B[I] / f.T[I]);
  double ev = 1.0 / f.cap;
  ev = ev + 1.0 * inv;
  ev = 1.0 / ev;
  const double ry = sxy * inv, rz = sxz * inv;
  f.gTxy[I] = gTxy;
  f.gTxz[I] = gTxz;
  f.sxy[I] = sxy;
  f.sxz[I] = sxz;
  f.sdev[I] = sdev;
  f.effVisc[I] = ev;
  f.invDisl[I] = inv;
  f.dgVdev[I] = sqrt(ry * ry + rz * rz);
}

int launch_pl_fields(const Ctx& c, const PlFields& f, hipStream_t s) {
  const long long N = (long long)c.N;
  hipLaunchKernelGGL(k_pl_fields, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, f);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int launch_pl_strain(const Ctx& c, const double* u, const double* gVxy, const double* gVxz,
                     const double* K, const double* n, const double* bcL, const double* bcR, double cap,
                     double* dgxy, double* dgxz, double* tauQS, double* effVisc, double* sxyOut,
                     const sc_fault_qd_args* fault, double* dslip, const PendingFinish* fin, hipStream_t s) {
  const long long N = (long long)c.N;
  PlStrainArgs a{u, gVxy, gVxz, K, n, bcL, bcR, cap, dgxy, dgxz, tauQS, effVisc, sxyOut};
  if (fin && fin->sc) a.fin = *fin;
  a.fault = fault ? 1 : 0;
  if (fault) {
    if (fault->N != c.d.Nz) return fail(SC_ERR_ARG, "power-law strain: the fault has Nz nodes");
    a.fa = *fault;
    a.dslip = dslip;
  }
  const int nGrid = (int)((N + 255) / 256);
  hipLaunchKernelGGL(k_pl_strain, dim3((unsigned)nGrid + (a.fin.sc ? 1u : 0u)), dim3(256), 0, s, c.op64, a,
                     nGrid);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc

using sc::fail;

extern "C" int sc_powerlaw_rates(sc_ctx* h, int prec, const sc_powerlaw_args* a) {
  if (!h || !a) return fail(SC_ERR_ARG, "sc_powerlaw_rates: null argument");
  if (!a->sxy_d || !a->T_d || !a->A_d || !a->B_d || !a->n_d || !a->mu_d || !a->gxy_t_d)
    return fail(SC_ERR_ARG, "sc_powerlaw_rates: null array");
  if (!(a->effViscCap > 0)) return fail(SC_ERR_ARG, "sc_powerlaw_rates: effViscCap must be > 0");
  if (a->N <= 0) return SC_OK;
  hipStream_t s = h->c.stream;
  const dim3 grid((a->N + 255) / 256), blk(256);
  if (prec == SC_FP64) {
    hipLaunchKernelGGL(sc::k_pl_rates<double>, grid, blk, 0, s, a->N, (const double*)a->sxy_d,
                       (const double*)a->sxz_d, (const double*)a->T_d, (const double*)a->A_d,
                       (const double*)a->B_d, (const double*)a->n_d, (const double*)a->mu_d,
                       (const double*)a->sat_d, a->effViscCap, (double*)a->effVisc_d,
                       (double*)a->gxy_t_d, (double*)a->gxz_t_d);
  } else if (prec == SC_FP32) {
    hipLaunchKernelGGL(sc::k_pl_rates<float>, grid, blk, 0, s, a->N, (const float*)a->sxy_d,
                       (const float*)a->sxz_d, (const float*)a->T_d, (const float*)a->A_d,
                       (const float*)a->B_d, (const float*)a->n_d, (const float*)a->mu_d,
                       (const float*)a->sat_d, a->effViscCap, (float*)a->effVisc_d,
                       (float*)a->gxy_t_d, (float*)a->gxz_t_d);
  } else {
    return fail(SC_ERR_ARG, "sc_powerlaw_rates: prec must be SC_FP64 or SC_FP32");
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}
