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

#include "cg_dev.hpp"
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

// out_v = base_v + sum_t c[t] f_v[t] at i (the order of k_vars_lincomb)
__device__ __forceinline__ double lincomb_at(const VarsLinComb& L, const double (&c)[6], int v, int i) {
  double x = c[0] * L.f[v][0][i] + L.base[v][i];
  for (int t = 1; t < L.nt; ++t) x = x + c[t] * L.f[v][t][i];
  return x;
}

// With the top and bottom data zero (the QD driver never sets them) the rhs
// is nonzero only in the kBand rows next to each y-side: this kernel covers
// those rows (the rest of the frame stays zero), over a grid-stride loop of
// its nbWork work blocks, and forms the block partials of y_m . rhs for the
// M basis vectors.  With L.nv > 0 it also forms the Runge-Kutta stage state
// of every field (k_vars_lincomb's work) and takes the slip of the rhs from
// field preSlip of that combination; with fin.sc set, one more block runs a
// pending residual test of the previous solve.
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_qd_rhs_frame_md(DevOp<double> p, int nLow, int nHigh,
                                                                 double* __restrict__ rhs,
                                                                 const double* __restrict__ slip,
                                                                 const double* __restrict__ shift, double t,
                                                                 double vL, double fts, int remote,
                                                                 double* __restrict__ bcL,
                                                                 double* __restrict__ bcR, VecList Y,
                                                                 double* mpart, int nbWork, VarsLinComb L,
                                                                 int preSlip, PendingFinish fin, StepGate gate,
                                                                 double tfrac) {
  if ((int)blockIdx.x >= nbWork) {
    // (run also when the step gate is closed: a repeated test of the same
    // partials sets the same flag)
    if (!(gate.ctl && gate.ctl->halt)) finish_irres_block(fin);
    return;
  }
  if (gate_closed(gate)) return;
  double cc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) cc[i] = L.c[i];
  if (gate.ctl) {  // this attempt's step and stage time
    const double h = gate.ctl->dt;
    t = gate.ctl->t + tfrac * h;
#pragma unroll
    for (int i = 0; i < 6; ++i) cc[i] = (L.c[i] * h) / L.cd[i];
  }
  const unsigned stride = (unsigned)nbWork * blockDim.x;
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (L.nv) {
    unsigned total = 0;
    for (int v = 0; v < L.nv; ++v) total += (unsigned)L.n[v];
    for (unsigned q = gid; q < total; q += stride) {
      int v = 0;
      unsigned i = q;
      while (i >= (unsigned)L.n[v]) i -= (unsigned)L.n[v++];
      L.out[v][i] = lincomb_at(L, cc, v, (int)i);
    }
  }
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  const unsigned Nz = (unsigned)p.Nz;
  const unsigned n = (unsigned)(nLow + nHigh) * Nz;
  for (unsigned q = gid; q < n; q += stride) {
    const unsigned row = q / Nz;
    const int k = (int)(q - row * Nz);
    const int j = (int)row < nLow ? (int)row : p.Ny - nHigh + ((int)row - nLow);
    const double sk = L.nv ? lincomb_at(L, cc, preSlip, k) : slip[k];
    const double l = sk * (1.0 / fts);
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
                   const ProjFramePlan& plan, const VarsLinComb& L, int preSlip, const PendingFinish& fin,
                   StepGate gate, double tfrac, hipStream_t s) {
  const unsigned grid = nb + (fin.sc ? 1u : 0u);
  hipLaunchKernelGGL(k_qd_rhs_frame_md<M>, dim3(grid), dim3(kRedThreads), 0, s, c.op64, nLow, nHigh, rhs, slip,
                     shift, t, vL, fts, remote, bcL, bcR, plan.Y, plan.mpart, (int)nb, L, preSlip, fin, gate,
                     tfrac);
}

int launch_qd_rhs_frame(const Ctx& c, double* rhs, const double* slip, const double* shift, double t,
                        double vL, double fts, int remote, double* bcL, double* bcR, const double* bcT,
                        const double* bcB, hipStream_t s, const ProjFramePlan* plan, int* mnb,
                        const VarsLinComb* pre, int preSlip, const PendingFinish* fin, StepGate gate,
                        double tfrac) {
  const FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
  const long long n = fm.count();
  if (mnb) *mnb = 0;
  if (!plan || plan->M == 0) {
    if (gate.ctl) return fail(SC_ERR_ARG, "a step-gated rhs needs a solution basis");
    if (fin && fin->sc) {
      const int e = launch_pending_finish(s, *fin);
      if (e) return e;
    }
    if (pre) {
      const int e = launch_vars_lincomb(s, *pre);
      if (e) return e;
    }
    hipLaunchKernelGGL(k_qd_rhs_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c.op64, fm, rhs,
                       slip, shift, t, vL, fts, remote, bcL, bcR, bcT, bcB);
    SC_HIP(hipGetLastError());
    return SC_OK;
  }
  // the y-band rows only: the QD driver never sets the top/bottom data, so
  // the rest of the frame is zero (as the plan-less launch above leaves it)
  const int Ny = c.d.Ny;
  int nLow, nHigh;
  qd_band_rows(Ny, &nLow, &nHigh);
  const long long nb_rows = (long long)(nLow + nHigh) * c.d.Nz;
  const unsigned nb = (unsigned)std::max<long long>(
      1, std::min<long long>((nb_rows + kRedThreads - 1) / kRedThreads, plan->maxBlocks));
  const VarsLinComb L = pre ? *pre : VarsLinComb{};
  const PendingFinish F = fin ? *fin : PendingFinish{};
#define SC_RHS_MD(m) \
  case m: launch_rhs_md<m>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, L, preSlip, F, gate, tfrac, s); break;
  switch (plan->M) {
    SC_RHS_MD(1) SC_RHS_MD(2) SC_RHS_MD(3) SC_RHS_MD(4) SC_RHS_MD(5) SC_RHS_MD(6) SC_RHS_MD(7)
    default: launch_rhs_md<8>(c, nLow, nHigh, nb, rhs, slip, shift, t, vL, fts, remote, bcL, bcR, *plan, L, preSlip, F, gate, tfrac, s); break;
  }
#undef SC_RHS_MD
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
__global__ __launch_bounds__(kLoHiThreads) void k_vars_lohi_err(VarsLoHi L, PendingFinish fin, StepGate gate) {
  if ((int)blockIdx.x >= L.nv) {  // a pending residual test of the last solve
    if (!(gate.ctl && gate.ctl->halt)) finish_irres_block(fin);
    return;
  }
  if (gate_closed(gate)) return;
  double clo[6], chi[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    clo[i] = L.clo[i];
    chi[i] = L.chi[i];
  }
  if (gate.ctl) {
    const double h = gate.ctl->dt;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      clo[i] = (L.clo[i] * h) / L.cdlo[i];
      chi[i] = (L.chi[i] * h) / L.cdhi[i];
    }
  }
  __shared__ double s0[kLoHiThreads], s1[kLoHiThreads], s2[kLoHiThreads];
  const int v = blockIdx.x;
  const int n = L.n[v];
  double a = 0, b = 0, m = -INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double base = L.base[v][i];
    double lo = clo[0] * L.flo[v][0][i] + base;
    for (int t = 1; t < L.ntlo; ++t) lo = lo + clo[t] * L.flo[v][t][i];
    double hi = chi[0] * L.fhi[v][0][i] + base;
    for (int t = 1; t < L.nthi; ++t) hi = hi + chi[t] * L.fhi[v][t][i];
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

int launch_vars_lohi_err(hipStream_t s, const VarsLoHi& L, const PendingFinish* fin, StepGate gate) {
  const PendingFinish F = fin ? *fin : PendingFinish{};
  const int grid = L.nv + (F.sc ? 1 : 0);
  if (grid == 0) return SC_OK;
  hipLaunchKernelGGL(k_vars_lohi_err, dim3(grid), dim3(kLoHiThreads), 0, s, L, F, gate);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// computeStepSize (odeSolver.cpp:395-445) on the control block's state, with
// the host's std::min / std::max order (a NaN ratio keeps the bound)
__device__ double step_size_dev(const StepCtl& c, double totErr) {
  double ratio;
  if (c.control == 0 || c.stepCount - c.pidBase < 4) {
    ratio = c.kappa * pow(c.tol / totErr, 1. / (1. + c.ord));
  } else {
    const double al = 0.49 / c.ord, be = 0.34 / c.ord, ga = 0.1 / c.ord;
    ratio = c.kappa * pow(c.tol / totErr, al) * pow(c.errA0 / c.tol, be) * pow(c.tol / c.errA1, ga);
  }
  double dt = ratio * c.dt;
  dt = (dt < c.dt * 5.0) ? dt : c.dt * 5.0;
  dt = (dt < c.maxDt) ? dt : c.maxDt;
  dt = (c.minDt < dt) ? dt : c.minDt;
  return dt;
}

// One attempt's end (the tail of the RK32 attempt loop and the accepted-step
// update of odeSolver.cpp:587-627): thread 0 updates the control block, the
// block copies an accepted solution into var.  A set fault status or
// missed-guess flag halts the loop with the state of the attempt's start.
__global__ __launch_bounds__(1024) void k_step_control(StepCtl* c, const double* __restrict__ err,
                                                       const int* __restrict__ flags, VarsCopy cp) {
  __shared__ int acc;
  if (threadIdx.x == 0) {
    acc = 0;
    c->accepted = 0;
    if (!c->halt && !c->done) {
      if (flags[0] || flags[1]) {
        c->halt = 1;
      } else {
        double tot = 0;
        for (int s = 0; s < c->nsel; ++s) {
          const double* v = err + 3 * s;
          if (c->normType == 0) tot += sqrt(v[0]) / (sqrt((double)c->selN[s]) * 1.0);
          else if (c->ord == 3) tot += sqrt(v[0]) / (sqrt(v[1]) * 1.0);
          else tot += v[2] / 1.0;
        }
        bool accept = c->ord == 3 ? tot <= c->tol : tot < c->tol;
        c->attempts++;
        if (!accept) {
          c->dt = step_size_dev(*c, tot);
          c->rejected++;
          if (c->minDt == c->dt || c->attempts >= 100) accept = true;
          else if (c->t + c->dt > c->finalT) c->dt = c->finalT - c->t;
        }
        if (accept) {
          c->t = c->t + c->dt;
          if (tot != 0.0) c->newDt = step_size_dev(*c, tot);
          c->errA1 = c->errA0;
          c->errA0 = tot;
          if (c->nOut < c->dtsCap) c->dts[c->nOut] = c->dt;
          c->nOut++;
          c->dt = c->newDt;
          c->accepted = 1;
          c->attempts = 0;
          if (!(c->stepCount < c->maxSteps && c->t < c->finalT)) {
            c->done = 1;
          } else {
            c->stepCount++;
            if (c->t + c->dt > c->finalT) c->dt = c->finalT - c->t;
          }
          acc = 1;
        }
      }
    }
  }
  __syncthreads();
  if (!acc) return;
  for (int v = 0; v < cp.nv; ++v)
    for (int i = threadIdx.x; i < cp.n[v]; i += blockDim.x) cp.dst[v][i] = cp.src[v][i];
}

int launch_step_control(hipStream_t s, StepCtl* ctl, const double* err, const int* flags, const VarsCopy& cp) {
  hipLaunchKernelGGL(k_step_control, dim3(1), dim3(1024), 0, s, ctl, err, flags, cp);
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
