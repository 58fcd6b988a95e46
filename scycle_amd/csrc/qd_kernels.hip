// Fused kernels of the quasi-dynamic driver: each one replaces a chain of
// small launches of one d_dt or Runge-Kutta stage with a single one.
//   k_qd_rhs_frame  boundary data from slip and time (StrikeSlip_LinearElastic_qd
//                   ::d_dt, strikeSlip_linearElastic_qd.cpp:763-770) and the
//                   right-hand side on the boundary frame (SbpOps::setRhs);
//                   with a solution basis, also the projections y_m . rhs
//                   that start the solver's projected guess
//   k_qd_fault      fault shear stress mu Dy u at j = 0 (LinearElastic::
//                   getStresses) and Fault_qd::d_dt (fault.cpp:861-918), with
//                   dslip = V
//   k_vars_lincomb  one Runge-Kutta combination over every integrated field
//   k_vars_lohi_err the two solutions of an embedded pair and their error
//                   sums, one block per short field
#include <algorithm>
#include <cmath>

#include "fault_dev.hpp"
#include "qd_internal.hpp"
#include "reduce_dev.hpp"
#include "sbp_dev.hpp"

namespace sc {

// bcL = slip / fts on the fault, bcR = vL t / fts + shift (remote loading);
// nodes of row 0 (and row Ny-1) also store the data they computed.
__global__ __launch_bounds__(256) void k_qd_rhs_frame(DevOp<double> p, FrameMap fm, double* __restrict__ rhs,
                                                      const double* __restrict__ slip,
                                                      const double* __restrict__ shift, double t, double vL,
                                                      double fts, int remote, double* __restrict__ bcL,
                                                      double* __restrict__ bcR,
                                                      const double* __restrict__ bcT,
                                                      const double* __restrict__ bcB) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= fm.count()) return;
  int j, k;
  fm.node(q, j, k);
  const double l = slip[k] * (1.0 / fts);
  const double r = remote ? vL * t / fts + 1.0 * shift[k] : bcR[k];
  rhs[(size_t)j * p.Nz + k] = rhs_node(p, j, k, [&](int side, int i) {
    return side == 0 ? l : side == 1 ? r : side == 2 ? bcT[i] : bcB[i];
  });
  if (j == 0) bcL[k] = l;
  if (j == p.Ny - 1 && remote) bcR[k] = r;
}

// With the top and bottom data zero (the QD driver never sets them) the rhs
// is nonzero only in the kBand rows next to each y-side: this kernel covers
// those rows (the rest of the frame stays zero), over a grid-stride loop, and
// forms the block partials of y_m . rhs for the M basis vectors.
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_qd_rhs_frame_md(DevOp<double> p, int nLow, int nHigh,
                                                                 double* __restrict__ rhs,
                                                                 const double* __restrict__ slip,
                                                                 const double* __restrict__ shift, double t,
                                                                 double vL, double fts, int remote,
                                                                 double* __restrict__ bcL,
                                                                 double* __restrict__ bcR, VecList Y,
                                                                 double* mpart) {
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  const unsigned Nz = (unsigned)p.Nz;
  const unsigned n = (unsigned)(nLow + nHigh) * Nz;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const unsigned row = q / Nz;
    const int k = (int)(q - row * Nz);
    const int j = (int)row < nLow ? (int)row : p.Ny - nHigh + ((int)row - nLow);
    const double l = slip[k] * (1.0 / fts);
    const double r = remote ? vL * t / fts + 1.0 * shift[k] : bcR[k];
    const size_t i = (size_t)j * Nz + k;
    const double v = rhs_node(p, j, k, [&](int side, int) { return side == 0 ? l : side == 1 ? r : 0.0; });
    rhs[i] = v;
    if (j == 0) bcL[k] = l;
    if (j == p.Ny - 1 && remote) bcR[k] = r;
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += Y.v[m][i] * v;
  }
  block_partials(acc, mpart);
}

template <int M>
void launch_rhs_md(const Ctx& c, int nLow, int nHigh, unsigned nb, double* rhs, const double* slip,
                   const double* shift, double t, double vL, double fts, int remote, double* bcL, double* bcR,
                   const ProjFramePlan& plan, hipStream_t s) {
  hipLaunchKernelGGL(k_qd_rhs_frame_md<M>, dim3(nb), dim3(kRedThreads), 0, s, c.op64, nLow, nHigh, rhs, slip,
                     shift, t, vL, fts, remote, bcL, bcR, plan.Y, plan.mpart);
}

int launch_qd_rhs_frame(const Ctx& c, double* rhs, const double* slip, const double* shift, double t,
                        double vL, double fts, int remote, double* bcL, double* bcR, const double* bcT,
                        const double* bcB, hipStream_t s, const ProjFramePlan* plan, int* mnb) {
  const FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
  const long long n = fm.count();
  if (mnb) *mnb = 0;
  if (!plan || plan->M == 0) {
    hipLaunchKernelGGL(k_qd_rhs_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c.op64, fm, rhs,
                       slip, shift, t, vL, fts, remote, bcL, bcR, bcT, bcB);
    SC_HIP(hipGetLastError());
    return SC_OK;
  }
  // the y-band rows only: the QD driver never sets the top/bottom data, so
  // the rest of the frame is zero (as the plan-less launch above leaves it)
  const int Ny = c.d.Ny;
  const int nLow = std::min(kBand, Ny);
  const int nHigh = Ny > kBand ? std::min(kBand, Ny - kBand) : 0;
  const long long nb_rows = (long long)(nLow + nHigh) * c.d.Nz;
  const unsigned nb = (unsigned)std::max<long long>(
      1, std::min<long long>((nb_rows + kRedThreads - 1) / kRedThreads, plan->maxBlocks));
  switch (plan->M) {
    case 1: launch_rhs_md<1>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 2: launch_rhs_md<2>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 3: launch_rhs_md<3>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 4: launch_rhs_md<4>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 5: launch_rhs_md<5>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 6: launch_rhs_md<6>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    case 7: launch_rhs_md<7>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
    default: launch_rhs_md<8>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, s); break;
  }
  SC_HIP(hipGetLastError());
  if (mnb) *mnb = (int)nb;
  return SC_OK;
}

// u set: tauQS = mu Dy u on the fault; u null: tauQS already in a.tauQS_d
__global__ __launch_bounds__(256) void k_qd_fault(DevOp<double> p, const double* __restrict__ u,
                                                  sc_fault_qd_args a, double* __restrict__ dslip) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.N) return;
  double tq;
  if (u) {
    tq = p.mu[k] * dy_row0(p, u, k);
  } else {
    tq = a.tauQS_d[k];
  }
  dslip[k] = fault_qd_node(a, k, tq);
}

int launch_qd_fault(const Ctx& c, const double* u, const sc_fault_qd_args& a, double* dslip, hipStream_t s) {
  if (a.N <= 0) return SC_OK;
  hipLaunchKernelGGL(k_qd_fault, dim3((a.N + 255) / 256), dim3(256), 0, s, c.op64, u, a, dslip);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// out_v = base_v + sum_i c[i] f_v[i] for every field v (the order of k_lincomb)
__global__ __launch_bounds__(256) void k_vars_lincomb(VarsLinComb L) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int v = 0;
  while (v < L.nv && i >= L.n[v]) i -= L.n[v++];
  if (v >= L.nv) return;
  double x = L.c[0] * L.f[v][0][i] + L.base[v][i];
  for (int t = 1; t < L.nt; ++t) x = x + L.c[t] * L.f[v][t][i];
  L.out[v][i] = x;
}

int launch_vars_lincomb(hipStream_t s, const VarsLinComb& L) {
  long long n = 0;
  for (int v = 0; v < L.nv; ++v) n += L.n[v];
  if (n == 0) return SC_OK;
  hipLaunchKernelGGL(k_vars_lincomb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// Block v: lo = base + sum_t clo[t] flo[t], hi likewise (the order of
// k_vars_lincomb), then [sum (lo-hi)^2, sum hi^2, max |lo-hi|/hi] over the
// pair.  kLoHiThreads threads: a field of up to that many entries takes one
// pass.
constexpr int kLoHiThreads = 1024;
__global__ __launch_bounds__(kLoHiThreads) void k_vars_lohi_err(VarsLoHi L) {
  __shared__ double s0[kLoHiThreads], s1[kLoHiThreads], s2[kLoHiThreads];
  const int v = blockIdx.x;
  const int n = L.n[v];
  double a = 0, b = 0, m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double base = L.base[v][i];
    double lo = L.clo[0] * L.flo[v][0][i] + base;
    for (int t = 1; t < L.ntlo; ++t) lo = lo + L.clo[t] * L.flo[v][t][i];
    double hi = L.chi[0] * L.fhi[v][0][i] + base;
    for (int t = 1; t < L.nthi; ++t) hi = hi + L.chi[t] * L.fhi[v][t][i];
    L.lo[v][i] = lo;
    L.hi[v][i] = hi;
    const double e = -1.0 * hi + lo;
    a += e * e;
    b += hi * hi;
    m = fmax(m, fabs(e) / hi);
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = m;
  __syncthreads();
  for (int w = kLoHiThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    L.out[v][0] = s0[0];
    L.out[v][1] = s1[0];
    L.out[v][2] = s2[0];
  }
}

int launch_vars_lohi_err(hipStream_t s, const VarsLoHi& L) {
  if (L.nv == 0) return SC_OK;
  hipLaunchKernelGGL(k_vars_lohi_err, dim3(L.nv), dim3(kLoHiThreads), 0, s, L);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// strikeSlip_linearElastic_qd_fd::checkSwitchRegime (:598-608): R = eta V,
// R = R / tauQS, VecMax(R); one block, grid-stride over the fault
constexpr int kRatioThreads = 1024;
__global__ __launch_bounds__(kRatioThreads) void k_max_ratio(const double* __restrict__ eta,
                                                             const double* __restrict__ V,
                                                             const double* __restrict__ tauQS, int n,
                                                             double* out) {
  __shared__ double sh[kRatioThreads];
  double m = -INFINITY;
  for (int k = threadIdx.x; k < n; k += kRatioThreads) m = fmax(m, (eta[k] * V[k]) / tauQS[k]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kRatioThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int launch_max_ratio(hipStream_t s, const double* eta, const double* V, const double* tauQS, int n,
                     double* out) {
  hipLaunchKernelGGL(k_max_ratio, dim3(1), dim3(kRatioThreads), 0, s, eta, V, tauQS, n, out);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc
