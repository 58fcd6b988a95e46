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
//   k_step_end      the end of a device-controlled RK32 attempt: the last
//                   stage's residual test, the pair, the step control
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "cg_dev.hpp"
#include "dd_dev.hpp"
#include "fault_dev.hpp"
#include "qd_internal.hpp"
#include "reduce_dev.hpp"
#include "sbp_dev.hpp"
#include "step_dev.hpp"

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
  const unsigned stride = (unsigned)nbWork * blockDim.x;
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned Nz = (unsigned)p.Nz;
  const unsigned n = (unsigned)(nLow + nHigh) * Nz;
  unsigned total = 0;
  for (int v = 0; v < L.nv; ++v) total += (unsigned)L.n[v];
  // The first combination entry's and the first band node's operands load
  // while the gate and the step arrive with the control block (they are in
  // bounds either way); later grid-stride passes load as they go.
  double pf[6] = {}, pb = 0.0;  // combination entry gid
  int pv = 0;
  unsigned pi = gid;
  if (gid < total) {
    while (pi >= (unsigned)L.n[pv]) pi -= (unsigned)L.n[pv++];
    pb = L.base[pv][pi];
#pragma unroll
    for (int u = 0; u < 6; ++u)
      if (u < L.nt) pf[u] = L.f[pv][u][pi];
  }
  double sf[6] = {}, sb = 0.0, sh0 = 0.0, y0[M];  // band node gid
#pragma unroll
  for (int m = 0; m < M; ++m) y0[m] = 0.0;
  if (gid < n) {
    const unsigned row = gid / Nz;
    const int k = (int)(gid - row * Nz);
    const int j = (int)row < nLow ? (int)row : p.Ny - nHigh + ((int)row - nLow);
    const size_t i = (size_t)j * Nz + k;
    if (L.nv) {
      sb = L.base[preSlip][k];
#pragma unroll
      for (int u = 0; u < 6; ++u)
        if (u < L.nt) sf[u] = L.f[preSlip][u][k];
    } else {
      sb = slip[k];
    }
    sh0 = remote ? shift[k] : bcR[k];
#pragma unroll
    for (int m = 0; m < M; ++m) y0[m] = Y.v[m][i];
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
  // the combination (k_vars_lincomb's order)
  auto comb = [&](const double* f, double base) {
    double x = cc[0] * f[0] + base;
#pragma unroll
    for (int u = 1; u < 6; ++u)
      if (u < L.nt) x = x + cc[u] * f[u];
    return x;
  };
  if (L.nv) {
    for (unsigned q = gid; q < total; q += stride) {
      if (q == gid) {
        L.out[pv][pi] = comb(pf, pb);
        continue;
      }
      int v = 0;
      unsigned i = q;
      while (i >= (unsigned)L.n[v]) i -= (unsigned)L.n[v++];
      L.out[v][i] = lincomb_at(L, cc, v, (int)i);
    }
  }
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  for (unsigned q = gid; q < n; q += stride) {
    const unsigned row = q / Nz;
    const int k = (int)(q - row * Nz);
    const int j = (int)row < nLow ? (int)row : p.Ny - nHigh + ((int)row - nLow);
    const bool first = q == gid;
    const double sk = L.nv ? (first ? comb(sf, sb) : lincomb_at(L, cc, preSlip, k)) : (first ? sb : slip[k]);
    const double l = sk * (1.0 / fts);
    const double shk = first ? sh0 : (remote ? shift[k] : bcR[k]);
    const double r = remote ? vL * t / fts + 1.0 * shk : shk;
    const size_t i = (size_t)j * Nz + k;
    const double v = rhs_node(p, j, k, [&](int side, int) { return side == 0 ? l : side == 1 ? r : 0.0; });
    rhs[i] = v;
    if (j == 0) bcL[k] = l;
    if (j == p.Ny - 1 && remote) bcR[k] = r;
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += (first ? y0[m] : Y.v[m][i]) * v;
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

// Block v: the pair's field v; a block past the fields runs a pending
// residual test of the last solve
__global__ __launch_bounds__(kRedThreads) void k_vars_lohi_err(VarsLoHi L, PendingFinish fin, StepGate gate) {
  if ((int)blockIdx.x >= L.nv) {
    if (!(gate.ctl && gate.ctl->halt)) finish_irres_block(fin);
    return;
  }
  if (gate_closed(gate)) return;
  __shared__ double sh[3 * (kRedThreads / 64)];
  double clo[6], chi[6], a, b, m;
  lohi_coefs(L, gate.ctl, clo, chi);
  lohi_accum(L, blockIdx.x, clo, chi, threadIdx.x, a, b, m);
  lohi_reduce(a, b, m, threadIdx.x, true, sh, L.out[blockIdx.x]);
}

int launch_vars_lohi_err(hipStream_t s, const VarsLoHi& L, const PendingFinish* fin, StepGate gate) {
  const PendingFinish F = fin ? *fin : PendingFinish{};
  const int grid = L.nv + (F.sc ? 1 : 0);
  if (grid == 0) return SC_OK;
  hipLaunchKernelGGL(k_vars_lohi_err, dim3(grid), dim3(kRedThreads), 0, s, L, F, gate);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// One block of kRedThreads (1 + nv) threads: group 0 the pending residual
// test (not once halted), group 1 + v field v of the pair (not once halted
// or done), side by side; then thread 0 updates the control block and an
// accepted solution goes into var.  The sums are those of k_vars_lohi_err
// and k_finish bit for bit.  The control block and the flags are read once
// into LDS, the error sums pass through LDS, and var takes hi from the
// registers that formed it (fields of up to kLoHiChunk kRedThreads entries),
// so the chain of dependent memory accesses stays short.
constexpr int kStepEndVars = 1024 / kRedThreads - 1;
constexpr int kCtlWords = sizeof(StepCtl) / sizeof(unsigned long long);
static_assert(sizeof(StepCtl) % sizeof(unsigned long long) == 0, "StepCtl is copied as 8-byte words");
__global__ __launch_bounds__(1024) void k_step_end(StepEnd e) {
  __shared__ double sh[kStepEndVars][3 * (kRedThreads / 64)];
  __shared__ unsigned long long cw[kCtlWords];
  __shared__ double errS[6 * kMaxVars];  // field sums, then the same by error slot
  __shared__ int flg[2], acc;
  StepCtl* cs = reinterpret_cast<StepCtl*>(cw);
  if (threadIdx.x < kCtlWords) cw[threadIdx.x] = reinterpret_cast<const unsigned long long*>(e.ctl)[threadIdx.x];
  if (threadIdx.x == kCtlWords) {
    flg[0] = e.flags[0];
    flg[1] = e.flags[1];
  }
  __syncthreads();
  const bool halted = cs->halt != 0;
  const bool closed = halted || cs->done != 0;
  const int grp = threadIdx.x / kRedThreads, tid = threadIdx.x % kRedThreads;
  const int v = grp - 1;
  const bool pair = v >= 0 && v < e.L.nv && !closed;
  double keep[kLoHiChunk], a = 0, b = 0, m = 0;
  if (pair) {
    double clo[6], chi[6];
    lohi_coefs(e.L, cs, clo, chi);
    lohi_accum(e.L, v, clo, chi, tid, a, b, m, keep);
  }
  // only the first kRedThreads threads accumulate in the test
  if (e.fin.sc && !halted) finish_irres_block(e.fin);
  if (threadIdx.x == 0 && e.fin.sc && !halted && !e.fin.sc->done) flg[1] = 1;
  lohi_reduce(a, b, m, tid, pair, sh[pair ? v : 0], pair ? errS + 3 * v : nullptr);
  __syncthreads();
  if (pair && tid < 3) e.L.out[v][tid] = errS[3 * v + tid];
  // the sums at their error slots: field q's go to out[q] = err + 3 slot
  // (slots below kMaxVars are the selected fields)
  double* sel = errS + 3 * kMaxVars;
  if (threadIdx.x < 3 * kMaxVars) sel[threadIdx.x] = 0.0;
  __syncthreads();
  if (threadIdx.x < 3 * e.L.nv) {
    const int q = threadIdx.x / 3;
    const long long slot = (e.L.out[q] - e.err) / 3;
    if (slot >= 0 && slot < kMaxVars) sel[3 * slot + threadIdx.x % 3] = errS[threadIdx.x];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const bool ok = step_control_wave(cs, sel, flg[0] || flg[1]);
    if (threadIdx.x == 0) acc = ok ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x < kCtlWords) reinterpret_cast<unsigned long long*>(e.ctl)[threadIdx.x] = cw[threadIdx.x];
  if (!acc) return;
  if (pair && e.L.n[v] <= kLoHiChunk * kRedThreads && e.cp.dst[v] == e.L.base[v]) {
#pragma unroll
    for (int q = 0; q < kLoHiChunk; ++q) {
      const int i = tid + q * kRedThreads;
      if (i < e.L.n[v]) e.cp.dst[v][i] = keep[q];
    }
    return;
  }
  if (grp != 0 || pair) return;  // a field too long for the registers: group 0 copies it
  for (int q = 0; q < e.cp.nv; ++q) {
    if (q < e.L.nv && e.L.n[q] <= kLoHiChunk * kRedThreads && e.cp.dst[q] == e.L.base[q]) continue;
    for (int i = tid; i < e.cp.n[q]; i += kRedThreads) e.cp.dst[q][i] = e.cp.src[q][i];
  }
}

int launch_step_end(hipStream_t s, const StepEnd& e, StepGate gate) {
  if (!gate.ctl || gate.ctl != e.ctl || gate.mode != 1)
    return fail(SC_ERR_ARG, "launch_step_end: the gate must be the attempt's gate on the control block");
  if (e.L.nv > kStepEndVars || e.cp.nv != e.L.nv) return fail(SC_ERR_ARG, "launch_step_end: field lists");
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(kRedThreads * (1 + e.L.nv)), 0, s, e);
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


// ---- One kernel per Runge-Kutta stage (the device-controlled RK32 loop) ----
//
// The QD rhs is linear in the boundary data and nonzero only on the band
// rows, so the projections of a stage's rhs on the basis are
//   y_m . b = sum_k PL_m(k) l(k) + sum_k PR_m(k) r(k),
// PL_m(k) / PR_m(k) = sum over the band rows j of y_m(j, k) times the rhs at
// (j, k) for a unit left / right datum, l = slip / fts, r = vL t / fts +
// shift (remote loading) or the fixed bcR.  With PL, sum_k PR_m and
// sum_k PR_m r_fix formed once per basis (k_stage_plan), every block of the
// stage kernel forms the guess coefficients from Nz entries, and the stage
// needs no rhs launch: the streaming blocks form b on the band rows where
// the residual needs it from each band node's rhs for unit data (also formed
// once per basis), b = aL l + aR r.  b and the coefficients round
// differently from the rhs kernel's (to the last bit).

constexpr int kStageThreads = 512;
constexpr int kStageNzMax = 2048;
constexpr int kStagePos = 3;  // stage 2, stage 3, rates at the accepted state
constexpr int kStageNT = 2;   // terms of a stage combination (RK32: 1 and 2)

// thread per k: PL and PR over the band rows, and each band node's rhs for
// unit left and right data (aL, aR: band-row order, rows 0..nLow-1 then
// Ny-nHigh..Ny-1)
template <int M>
__global__ __launch_bounds__(256) void k_stage_plan(DevOp<double> p, int nLow, int nHigh, VecList Y,
                                                    double* __restrict__ PL, double* __restrict__ PR,
                                                    double* __restrict__ aL, double* __restrict__ aR) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int Nz = p.Nz;
  if (k >= Nz) return;
  double al[M], ar[M];
#pragma unroll
  for (int m = 0; m < M; ++m) al[m] = ar[m] = 0.0;
  for (int row = 0; row < nLow + nHigh; ++row) {
    const int j = row < nLow ? row : p.Ny - nHigh + (row - nLow);
    const double uL = rhs_node(p, j, k, [](int side, int) { return side == 0 ? 1.0 : 0.0; });
    const double uR = rhs_node(p, j, k, [](int side, int) { return side == 1 ? 1.0 : 0.0; });
    aL[(size_t)row * Nz + k] = uL;
    aR[(size_t)row * Nz + k] = uR;
    const size_t i = (size_t)j * Nz + k;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      al[m] += Y.v[m][i] * uL;
      ar[m] += Y.v[m][i] * uR;
    }
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    PL[(size_t)m * Nz + k] = al[m];
    PR[(size_t)m * Nz + k] = ar[m];
  }
}

// G_mn = <A y_m, A y_n> over the nodes off the band rows (where b is zero),
// in double-double: per-block partials (grid-stride), then one block sums
// them.  A y_m is the rhs of an earlier solve up to its residual there, so
// these entries are of the size of solve residuals squared, and
// c^T G c of the guess's coefficients keeps its digits.
constexpr int kGramBlocks = 256;
template <int M>
__global__ __launch_bounds__(256) void k_stage_gram(VecList Z, size_t zLo, size_t zHi, double* __restrict__ part) {
  constexpr int P = M * (M + 1) / 2;
  DD g[P];
#pragma unroll
  for (int q = 0; q < P; ++q) g[q] = DD{0.0, 0.0};
  for (size_t i = zLo + (size_t)blockIdx.x * 256 + threadIdx.x; i < zHi; i += (size_t)kGramBlocks * 256) {
    double z[M];
#pragma unroll
    for (int m = 0; m < M; ++m) z[m] = Z.v[m][i];
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
      for (int n = m; n < M; ++n) {
        const int q = m * M - m * (m - 1) / 2 + (n - m);  // pair (m, n), n >= m, row by row
        g[q] = dd_add(g[q], dd_prod(z[m], z[n]));
      }
  }
  __shared__ double sh[P][4][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const DD w = dd_wave_sum(g[q]);
    if (lane == 0) {
      sh[q][wid][0] = w.hi;
      sh[q][wid][1] = w.lo;
    }
  }
  __syncthreads();
  if (threadIdx.x < P) {
    DD t{0.0, 0.0};
    for (int w = 0; w < 4; ++w) t = dd_add(t, DD{sh[threadIdx.x][w][0], sh[threadIdx.x][w][1]});
    part[((size_t)blockIdx.x * P + threadIdx.x) * 2] = t.hi;
    part[((size_t)blockIdx.x * P + threadIdx.x) * 2 + 1] = t.lo;
  }
}

// one block: G (M x M, both triangles) from the partials
template <int M>
__global__ __launch_bounds__(64) void k_stage_gram_sum(const double* __restrict__ part, double* __restrict__ G) {
  constexpr int P = M * (M + 1) / 2;
  const int q = threadIdx.x;
  if (q >= P) return;
  DD t{0.0, 0.0};
  for (int b = 0; b < kGramBlocks; ++b) t = dd_add(t, DD{part[((size_t)b * P + q) * 2], part[((size_t)b * P + q) * 2 + 1]});
  int m = 0, r = q;
  while (r >= M - m) {
    r -= M - m;
    ++m;
  }
  const int n = m + r;
  G[((size_t)m * M + n) * 2] = t.hi;
  G[((size_t)m * M + n) * 2 + 1] = t.lo;
  G[((size_t)n * M + m) * 2] = t.hi;
  G[((size_t)n * M + m) * 2 + 1] = t.lo;
}

// one block: sR[m] = sum_k PR_m(k), sR[M + m] = sum_k PR_m(k) rfix(k)
template <int M>
__global__ __launch_bounds__(256) void k_stage_plan_sums(const double* __restrict__ PR, const double* __restrict__ rfix,
                                                         int Nz, double* __restrict__ sR) {
  __shared__ double sh[2 * M][4];
  double a[2 * M];
#pragma unroll
  for (int q = 0; q < 2 * M; ++q) a[q] = 0.0;
  for (int k = threadIdx.x; k < Nz; k += 256) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double v = PR[(size_t)m * Nz + k];
      a[m] += v;
      a[M + m] += v * rfix[k];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 2 * M; ++q) {
    const double t = wave_sum(a[q]);
    if (lane == 0) sh[q][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x < 2 * M) sR[threadIdx.x] = ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
}

struct QdStageArgs {
  int Nz, nfb, nstream;
  size_t N, zLo, zHi;
  VecList Z, T;
  const double* w;   // projection weights
  const double* PL;  // M x Nz
  const double* sR;  // 2 M
  const double* aL;  // band nodes' rhs for a unit left datum
  const double* aR;  // and a unit right datum
  // the off-band Gram matrix G_mn = <A y_m, A y_n> (M x M, double-double)
  const double* G;
  double* coefOut;   // the coefficients, for proj_materialize
  double* parts;     // this stage's residual partials (2 per streaming block)
  CgScal* sc;        // this stage's scalars
  double tol2;
  QdFuse fuse;       // fault rates: psi_d the stage's psi, dpsi_d / dslip its rates
  VarsLinComb L;     // the stage combination (nv = 0: the state itself)
  int psiIdx, slipIdx;
  const double* slipVar;  // the slip when L.nv = 0
  // the stage's slip and psi operands resolved by the host (no indexed
  // argument loads in the kernel): base, the combination's rate terms (the
  // base again past ntS terms), the outputs of the combination
  int ntS;  // terms of the combination (0: the state itself)
  const double* sBase;
  const double* sF[kStageNT];
  const double* pBase;
  const double* pF[kStageNT];
  double* sOut;
  double* pOut;
  double fts, vL;
  int remote;
  const double* rfix;  // shift (remote loading) or the fixed bcR
  double* bcL;
  double* bcR;
  StepGate gate;
  double tfrac, t0;   // t0: the time without a gate
  PendingFinish fin;  // the previous stage's residual test (the last block)
  // the attempt's end (END kernels; fields 0 psi, 1 slip)
  const StepCtl* ctlIn;
  StepCtl* ctlOut;
  int* flags;          // fault status, missed guess
  const int* flagEnd;  // the previous attempt's end's root-solve status
  PendingFinish fin3;  // the last stage's residual test
  const double* pv[2];      // the attempt's state (read)
  double* pvOut[2];         // the next attempt's state (written)
  const double* pf[2][3];  // the rates of the attempt's three stages
  double clo[2], cdlo[2], chi[3], cdhi[3];
  int selSlot[2];
};

// Blocks [0, nfb): the stage state and fault rates of 64 fault nodes each
// (one wave: the root solves are FP64 VALU work, spread one wave per CU);
// [nfb, nfb + nstream): the guess residual's r.r and b.b partials over the
// band rows, b = aL l + aR r there, the first adding the nodes off the band
// (b = 0 there) as c^T G c from the plan's off-band Gram matrix, so that no
// block passes over the grid; one more block: the previous stage's residual
// test.
// the gate words a stage launch without a gate reads (never written)
__device__ StepCtl g_no_gate;

// KQ: fault rows per thread in the prologue (Nz <= KQ * kStageThreads)
template <int M, bool END, int KQ>
__global__ __launch_bounds__(kStageThreads, 2) void k_qd_stage(QdStageArgs a) {
  const int blk = blockIdx.x;
  if (!END && blk >= a.nfb + a.nstream) {
    if (!(a.gate.ctl && a.gate.ctl->halt)) finish_irres_block(a.fin);
    return;
  }
  __shared__ double sS[kStageNzMax], rS[kStageNzMax];
  __shared__ double pS[END ? kStageNzMax : 1];
  __shared__ double red[M > 2 ? M : 2][kStageThreads / 64];
  __shared__ double coef[M];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int Nz = a.Nz;
  const int sblk = blk - a.nfb;  // < 0: a fault block
  constexpr int kCtlW = sizeof(StepCtl) / sizeof(unsigned long long);
  // The control block's words go out first: loads return in the order they
  // were issued, so a gate read issued after the prefetches below would wait
  // for all of them.  Plain stages read the gate's words; the attempt's end
  // reads the whole block (one word a thread) and the flags.
  // (loaded unconditionally, from a zeroed record when there is no gate,
  // so that no branch makes the compiler wait for them here)
  int gHalt = 0, gAcc = 0, gDone = 0;
  double gDt = 0.0, gT = 0.0;
  unsigned long long cwReg = 0;
  int flgReg0 = 0, flgReg1 = 0;
  if constexpr (END) {
    if (tid < kCtlW) cwReg = reinterpret_cast<const unsigned long long*>(a.ctlIn)[tid];
    if (tid == kCtlW) {
      flgReg0 = a.flags[0] | a.flagEnd[0];
      flgReg1 = a.flags[1];
    }
  } else {
    const StepCtl* g = a.gate.ctl ? a.gate.ctl : &g_no_gate;
    gHalt = g->halt;
    gAcc = g->accepted;
    gDone = g->done;
    gDt = g->dt;
    gT = g->t;
  }
  __builtin_amdgcn_sched_barrier(0);
  // Every load that does not depend on the step goes out before the gate is
  // waited for: the residual blocks' band nodes (A y_m and the unit-datum
  // rhs), and each thread's fault nodes' slip operands and PL entries.
  constexpr int kBN = 2;  // band nodes a thread of a residual block
  const size_t nBand = a.zLo + (a.N - a.zHi);
  double zb[kBN][M], ab[kBN][2];
  size_t bi_[kBN];
#pragma unroll
  for (int u = 0; u < kBN; ++u) {
    const size_t q = (size_t)(sblk >= 0 ? sblk : 0) * (kBN * kStageThreads) + (size_t)u * kStageThreads + tid;
    const bool use = sblk >= 0 && q < nBand;
    const size_t qc = use ? q : 0;  // (unconditional loads: none waits in a branch)
    const size_t ic = qc < a.zLo ? qc : a.zHi + (qc - a.zLo);
    bi_[u] = use ? ic : a.N;
#pragma unroll
    for (int m = 0; m < M; ++m) zb[u][m] = a.Z.v[m][ic];
    ab[u][0] = a.aL[qc];
    ab[u][1] = a.aR[qc];
  }
  constexpr int kQ = KQ;
  static_assert(KQ * kStageThreads <= kStageNzMax, "the LDS rows hold kStageNzMax nodes");
  double sb[kQ], sf[kQ][kStageNT], pl[kQ][M], rk[kQ];
  // the attempt's end: its state and its stages' rates (base, f0, f1, f2 of
  // psi and slip), and the last stage's first residual partials
  double ev[END ? kQ : 1][2][4];
  double w0 = 0.0, w1 = 0.0;
  if constexpr (END) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int k = min(tid + q * kStageThreads, Nz - 1);
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        ev[q][f][0] = a.pv[f][k];
        ev[q][f][1] = a.pf[f][0][k];
        ev[q][f][2] = a.pf[f][1][k];
        ev[q][f][3] = a.pf[f][2][k];
      }
    }
    const int b0 = min(tid, a.fin3.nb - 1);
    const double p0 = a.fin3.partials[(size_t)b0 * 2], p1 = a.fin3.partials[(size_t)b0 * 2 + 1];
    if (tid < a.fin3.nb) {
      w0 = p0;
      w1 = p1;
    }
  }
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    // (clamped rows: every load is unconditional, so that none waits in a
    // branch; rows past Nz are not used)
    const int k = min(tid + q * kStageThreads, Nz - 1);
    if constexpr (END) {
      sb[q] = 0.0;  // the state is the attempt's solution, formed below
#pragma unroll
      for (int u = 0; u < kStageNT; ++u) sf[q][u] = 0.0;
    } else {
      sb[q] = a.sBase[k];
#pragma unroll
      for (int u = 0; u < kStageNT; ++u) sf[q][u] = a.sF[u][k];
    }
    rk[q] = a.rfix[k];
#pragma unroll
    for (int m = 0; m < M; ++m) pl[q][m] = a.PL[(size_t)m * Nz + k];
  }
  // fault lanes: the node's step-independent operands (tractions of the
  // basis, friction inputs, the previous V) and the operands of the stage's
  // psi, so that none of them is a memory round trip after the gate
  const int kf = blk * 64 + tid;
  const bool faultLane = sblk < 0 && tid < 64 && kf < Nz;
  FaultNodeIn fnode{};
  double tm[M], psiB = 0.0, psiF[kStageNT];
#pragma unroll
  for (int m = 0; m < M; ++m) tm[m] = 0.0;
#pragma unroll
  for (int u = 0; u < kStageNT; ++u) psiF[u] = 0.0;
  if (faultLane) {
#pragma unroll
    for (int m = 0; m < M; ++m) tm[m] = a.T.v[m][kf];
    fnode = fault_node_load(a.fuse.a, kf, 0.0);
    if constexpr (!END) {
      psiB = a.pBase[kf];
#pragma unroll
      for (int u = 0; u < kStageNT; ++u) psiF[u] = a.pF[u][kf];
    }
  }
  // the coefficient step's per-basis plan sums and weights
  double wP = 0.0, sR0 = 0.0, sR1 = 0.0;
  if (tid < M) {
    wP = a.w[tid];
    sR0 = a.sR[tid];
    sR1 = a.sR[M + tid];
  }
  double cc[kStageNT];
#pragma unroll
  for (int i = 0; i < kStageNT; ++i) cc[i] = a.L.c[i];
  double t = a.t0;
  // the attempt's end: the control block's copy
  __shared__ unsigned long long cw[END ? kCtlW : 1];
  if constexpr (END) {
    // The attempt's end, formed by every block alike (the same loads and
    // arithmetic, so every block reaches the same decision): the last
    // stage's residual test (k_finish's order), the RK32 pair and its error
    // sums, the step control on a copy of the control block, by the second
    // wave.  Beside the decision the other waves form the rates (fault
    // lanes) and the guess residual (residual blocks) at the pair's solution
    // as if the attempt were accepted, in registers; the writes follow the
    // decision: an accepted attempt's rates and next state (var <- hi), or a
    // rejected or closed attempt's start passed on.  Block 0 writes the
    // updated control block to ctlOut (the next attempt reads it; this
    // attempt's kernels read ctlIn), the test's scalars and flag.
    __shared__ double rred[2][kStageThreads / 64];
    __shared__ double sel[3 * kMaxVars];
    __shared__ int accS, fastS;
    __shared__ double tAccS;
    double wPm[M], sRm0[M], sRm1[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      wPm[m] = a.w[m];
      sRm0[m] = a.sR[m];
      sRm1[m] = a.sR[M + m];
    }
    __shared__ double eS[2][3][kStageThreads / 64], fS[2][kRedThreads / 64];
    __shared__ int flg[2];
    StepCtl* cs = reinterpret_cast<StepCtl*>(cw);
    if (tid < kCtlW) cw[tid] = cwReg;
    if (tid == kCtlW) {
      flg[0] = flgReg0;
      flg[1] = flgReg1;
    }
    if (tid < kRedThreads)
      for (int bq = tid + kRedThreads; bq < a.fin3.nb; bq += kRedThreads) {
        w0 += a.fin3.partials[(size_t)bq * 2];
        w1 += a.fin3.partials[(size_t)bq * 2 + 1];
      }
    __syncthreads();
    // a closed or rejected attempt passes the state and its rates on
    auto pass_state = [&]() {
      const int k = blk * 64 + tid;
      if (blk < a.nfb && tid < 64 && k < Nz) {
        a.pvOut[0][k] = a.pv[0][k];
        a.pvOut[1][k] = a.pv[1][k];
        a.fuse.a.dpsi_d[k] = a.pf[0][0][k];
        a.fuse.dslip[k] = a.pf[1][0][k];
      }
    };
    if (cs->halt || cs->done) {  // closed: pass the control block on
      if (blk == 0 && tid == 0) {
        StepCtl c = *cs;
        c.accepted = 0;
        *a.ctlOut = c;
      }
      pass_state();
      return;
    }
    w0 = wave_sum(w0);
    w1 = wave_sum(w1);
    if (lane == 0 && wid < kRedThreads / 64) {
      fS[0][wid] = w0;
      fS[1][wid] = w1;
    }
    const double h = cs->dt;
    double clo[2], chi[3];
#pragma unroll
    for (int i = 0; i < 2; ++i) clo[i] = (a.clo[i] * h) / a.cdlo[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) chi[i] = (a.chi[i] * h) / a.cdhi[i];
    const double tSpec = cs->t + cs->dt;  // an accepted step's end (the fast path's)
    double ea[2] = {0.0, 0.0}, eb[2] = {0.0, 0.0}, em[2] = {-INFINITY, -INFINITY};
    double acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0.0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int k = tid + q * kStageThreads;
      if (k < Nz) {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          const double base = ev[q][f][0], f0 = ev[q][f][1], f1 = ev[q][f][2], f2 = ev[q][f][3];
          double lo = clo[0] * f0 + base;  // y + h/2 (f0 + f2)
          lo = lo + clo[1] * f2;
          double hi = chi[0] * f0 + base;  // y + h (f0/6 + 2 f1/3 + f2/6)
          hi = hi + chi[1] * f1;
          hi = hi + chi[2] * f2;
          const double e = -1.0 * hi + lo;
          ea[f] += e * e;
          eb[f] += hi * hi;
          em[f] = fmax(em[f], fabs(e) / hi);
          (f == 0 ? pS : sS)[k] = hi;
        }
        // the pair's solution as the next stage state (the stage path's
        // arithmetic): right datum and this thread's share of the projections
        const double l = sS[k] * (1.0 / a.fts);
        rS[k] = a.remote ? a.vL * tSpec / a.fts + 1.0 * rk[q] : rk[q];
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] += pl[q][m] * l;
      }
    }
    // only the sums the norm reads (computeError, odeSolver.cpp:449-501):
    // L2_absolute sum (lo-hi)^2; L2_relative with RK32 also sum hi^2;
    // otherwise max |lo-hi|/hi (the others stay zero)
    const bool needB = cs->normType != 0 && cs->ord == 3, needM = cs->normType != 0 && cs->ord != 3;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double x = wave_sum(ea[f]);
      const double y = needB ? wave_sum(eb[f]) : 0.0;
      const double z = needM ? wave_max(em[f]) : 0.0;
      if (lane == 0) {
        eS[f][0][wid] = x;
        eS[f][1][wid] = y;
        eS[f][2][wid] = z;
      }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double sm = wave_sum(acc[m]);
      if (lane == 0) red[m][wid] = sm;
    }
    __syncthreads();
    // the coefficients at the pair's solution, every thread (the stage path's order)
    double c[M];
    auto coefs_at = [&](double tt) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        double yb = 0.0;
        for (int q = 0; q < kStageThreads / 64; ++q) yb += red[m][q];
        yb += a.remote ? (a.vL * tt / a.fts) * sRm0[m] + sRm1[m] : sRm1[m];
        c[m] = -wPm[m] * yb;
      }
    };
    coefs_at(tSpec);
    // the rates (fault lanes) or the guess residual (residual blocks) at c
    FaultNodeOut fo{};
    double r2 = 0.0, b2 = 0.0;
    auto spec = [&]() {
      if (sblk < 0) {
        if (faultLane) {
          double tq = 0.0;
#pragma unroll
          for (int m = 0; m < M; ++m) tq += c[m] * tm[m];
          fnode.psi = pS[kf];
          fo = fault_qd_eval(a.fuse.a, tq, fnode);
        }
      } else {
        r2 = 0.0;
        b2 = 0.0;
#pragma unroll
        for (int u = 0; u < kBN; ++u) {
          if (bi_[u] < a.N) {
            double az = 0.0;
#pragma unroll
            for (int m = 0; m < M; ++m) az += c[m] * zb[u][m];
            const int k = (int)(bi_[u] % (size_t)Nz);
            const double bi = ab[u][0] * (sS[k] * (1.0 / a.fts)) + ab[u][1] * rS[k];
            const double ri = bi - az;
            r2 += ri * ri;
            b2 += bi * bi;
          }
        }
      }
    };
    if (wid == 1) {
      // the decision: the last stage's residual test and the selected error
      // sums (each lane forms them), then the controller.  An attempt its
      // error accepts is accepted at once, block 0 forming the next step
      // size and control block here; a flagged, rejected or forced attempt
      // runs the whole controller in every block.
      double t0 = 0.0, t1 = 0.0;
      for (int ww = 0; ww < kRedThreads / 64; ++ww) t0 += fS[0][ww];
      for (int ww = 0; ww < kRedThreads / 64; ++ww) t1 += fS[1][ww];
      const bool miss3 = !(t0 <= a.tol2 * t1 || t1 == 0.0);
      // (every lane writes the same values: the wave reads them back below)
#pragma unroll
      for (int q = 0; q < 3 * kMaxVars; ++q) sel[q] = 0.0;
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const int sl = a.selSlot[f];
        if (sl < 0) continue;
        double x = eS[f][0][0], y = eS[f][1][0], z = eS[f][2][0];
        for (int ww = 1; ww < kStageThreads / 64; ++ww) {
          x += eS[f][0][ww];
          y += eS[f][1][ww];
          z = fmax(z, eS[f][2][ww]);
        }
        sel[3 * sl] = x;
        sel[3 * sl + 1] = y;
        sel[3 * sl + 2] = z;
      }
      if (blk == 0 && lane == 0) {  // the test's scalars and flag, for the host
        a.fin3.sc->rr = t0;
        a.fin3.sc->bb = t1;
        a.fin3.sc->done = miss3 ? 0 : 1;
        if (miss3 && a.fin3.flag) *a.fin3.flag = 1;
      }
      const bool flagged = flg[0] || flg[1] || miss3;
      if (!flagged && step_accepts(cs, sel)) {
        if (lane == 0) {
          accS = 1;
          fastS = 1;
          tAccS = cs->t + cs->dt;
        }
        if (blk == 0) {
          (void)step_control_wave(cs, sel, false, true);
          if (lane == 0) *a.ctlOut = *cs;
        }
      } else {
        const bool ok = step_control_wave(cs, sel, flagged, blk == 0);
        if (lane == 0) {
          accS = ok ? 1 : 0;
          fastS = 0;
        }
      }
    }
    spec();
    __syncthreads();
    if (!fastS && blk == 0 && tid < kCtlW) reinterpret_cast<unsigned long long*>(a.ctlOut)[tid] = cw[tid];
    if (!accS) {
      pass_state();
      return;
    }
    const double tEnd = fastS ? tAccS : cs->t;
    if (a.remote && tEnd != tSpec) {
      // an acceptance the controller forced ended the step at another time:
      // the right datum, the coefficients and the rates there
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        const int k = tid + q * kStageThreads;
        if (k < Nz) rS[k] = a.vL * tEnd / a.fts + 1.0 * rk[q];
      }
      __syncthreads();
      coefs_at(tEnd);
      spec();
    }
    if (sblk < 0) {
      if (faultLane) {
        const int k = kf;
        a.pvOut[0][k] = pS[k];
        a.pvOut[1][k] = sS[k];
        a.fuse.dslip[k] = fault_qd_commit(a.fuse.a, k, fo);
        a.bcL[k] = sS[k] * (1.0 / a.fts);
        if (a.remote) a.bcR[k] = rS[k];
      }
      return;
    }
    if (sblk == 0 && tid == 0) {
      a.sc->tol2 = a.tol2;
      a.sc->maxit = 0;
      a.sc->done = 0;
      a.sc->iters = 0;
      a.sc->cnt = 0;
    }
    if (sblk == 0 && tid < M) a.coefOut[tid] = c[tid];
    r2 = wave_sum(r2);
    b2 = wave_sum(b2);
    if (lane == 0) {
      rred[0][wid] = r2;
      rred[1][wid] = b2;
    }
    __syncthreads();
    if (tid < 2) {
      double v = 0.0;
      for (int q = 0; q < kStageThreads / 64; ++q) v += rred[tid][q];
      if (tid == 0 && sblk == 0) {
        // the nodes off the band rows, where b = 0: ||sum c_m A y_m||^2 =
        // c^T G c with the off-band Gram matrix (double-double)
        DD q2{0.0, 0.0};
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
          for (int n = 0; n < M; ++n) {
            const size_t o = ((size_t)m * M + n) * 2;
            q2 = dd_add(q2, dd_mul_d(dd_mul_d(DD{a.G[o], a.G[o + 1]}, c[m]), c[n]));
          }
        const double off = q2.hi + q2.lo;
        v += off > 0.0 ? off : 0.0;
      }
      a.parts[(size_t)sblk * 2 + tid] = v;
    }
    return;
  } else {
    // gate_closed
    if (a.gate.ctl && (gHalt || (a.gate.mode == 2 ? !gAcc : (a.gate.mode == 1 && gDone)))) return;
    if (a.gate.ctl) {  // this attempt's step and stage time
      const double h = gDt;
      t = gT + a.tfrac * h;
#pragma unroll
      for (int i = 0; i < kStageNT; ++i) cc[i] = (a.L.c[i] * h) / a.L.cd[i];
    }
  }
  // (the plain stage from here: the attempt's end returned above)
  // the stage's slip and right datum at every fault node (into LDS), and
  // this thread's share of sum_k PL_m(k) l(k)
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int k = tid + q * kStageThreads;
    if (k < Nz) {
      double sk = sb[q];
      if (a.ntS > 0) {  // the combination (k_vars_lincomb's order)
        sk = cc[0] * sf[q][0] + sb[q];
#pragma unroll
        for (int u = 1; u < kStageNT; ++u)
          if (u < a.ntS) sk = sk + cc[u] * sf[q][u];
      }
      const double l = sk * (1.0 / a.fts);
      const double r = a.remote ? a.vL * t / a.fts + 1.0 * rk[q] : rk[q];
      sS[k] = sk;
      rS[k] = r;
#pragma unroll
      for (int m = 0; m < M; ++m) acc[m] += pl[q][m] * l;
    }
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const double s = wave_sum(acc[m]);
    if (lane == 0) red[m][wid] = s;
  }
  __syncthreads();
  if (tid < M) {
    double yb = 0.0;
    for (int q = 0; q < kStageThreads / 64; ++q) yb += red[tid][q];
    yb += a.remote ? (a.vL * t / a.fts) * sR0 + sR1 : sR1;
    coef[tid] = -wP * yb;
  }
  __syncthreads();
  double c[M];
#pragma unroll
  for (int m = 0; m < M; ++m) c[m] = coef[m];
  if (sblk < 0) {
    // fault nodes: the stage state, tau = sum c_m T_m, the rates
    const int k = kf;
    if (faultLane) {
      double psi = psiB;
      if (a.ntS > 0) {  // psi of the stage (k_vars_lincomb's order)
        double x = cc[0] * psiF[0] + psiB;
#pragma unroll
        for (int u = 1; u < kStageNT; ++u)
          if (u < a.ntS) x = x + cc[u] * psiF[u];
        psi = x;
        a.pOut[k] = x;
        a.sOut[k] = sS[k];
      }
      double tq = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) tq += c[m] * tm[m];
      fnode.psi = psi;
      a.fuse.dslip[k] = fault_qd_node_in(a.fuse.a, k, tq, fnode);
      a.bcL[k] = sS[k] * (1.0 / a.fts);
      if (a.remote) a.bcR[k] = rS[k];
    }
    return;
  }
  if (sblk == 0 && tid == 0) {
    a.sc->tol2 = a.tol2;
    a.sc->maxit = 0;
    a.sc->done = 0;
    a.sc->iters = 0;
    a.sc->cnt = 0;
  }
  if (sblk == 0 && tid < M) a.coefOut[tid] = c[tid];
  // ||b - sum c_m A y_m||^2 on the band nodes, node by node, b = aL l + aR r
  double r2 = 0.0, b2 = 0.0;
#pragma unroll
  for (int u = 0; u < kBN; ++u) {
    if (bi_[u] < a.N) {
      double az = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) az += c[m] * zb[u][m];
      const int k = (int)(bi_[u] % (size_t)Nz);
      const double bi = ab[u][0] * (sS[k] * (1.0 / a.fts)) + ab[u][1] * rS[k];
      const double ri = bi - az;
      r2 += ri * ri;
      b2 += bi * bi;
    }
  }
  r2 = wave_sum(r2);
  b2 = wave_sum(b2);
  __syncthreads();  // red is reused
  if (lane == 0) {
    red[0][wid] = r2;
    red[1][wid] = b2;
  }
  __syncthreads();
  if (tid < 2) {
    double v = 0.0;
    for (int q = 0; q < kStageThreads / 64; ++q) v += red[tid][q];
    if (tid == 0 && sblk == 0) {
      // the nodes off the band rows, where b = 0: ||sum c_m A y_m||^2 =
      // c^T G c with the off-band Gram matrix (double-double)
      DD q2{0.0, 0.0};
#pragma unroll
      for (int m = 0; m < M; ++m)
#pragma unroll
        for (int n = 0; n < M; ++n) {
          const size_t o = ((size_t)m * M + n) * 2;
          q2 = dd_add(q2, dd_mul_d(dd_mul_d(DD{a.G[o], a.G[o + 1]}, c[m]), c[n]));
        }
      const double off = q2.hi + q2.lo;
      v += off > 0.0 ? off : 0.0;
    }
    a.parts[(size_t)sblk * 2 + tid] = v;
  }
}

// the gram buffer: G (2 SC_REUSE_MAX^2 doubles), the Gram kernel's partials
static size_t gram_doubles() {
  constexpr size_t R = SC_REUSE_MAX;
  return 2 * R * R + (size_t)kGramBlocks * (R * (R + 1) / 2) * 2;
}

int stage_bufs_ensure(QdStageBufs& b, int Nz, int nBandRows) {
  // residual blocks: 2 band nodes a thread
  const int nstream = (int)(((size_t)nBandRows * Nz + 2 * kStageThreads - 1) / (2 * kStageThreads));
  if (b.plan && b.nz == Nz && b.nstream == nstream) return SC_OK;
  stage_bufs_free(b);
  // PL, PR (SC_REUSE_MAX x Nz each), sums (2 SC_REUSE_MAX), aL, aR (2 kBand x Nz each)
  SC_HIP(hipMalloc(&b.plan, (2 * (size_t)SC_REUSE_MAX * Nz + 2 * SC_REUSE_MAX + 4 * (size_t)kBand * Nz) *
                                sizeof(double)));
  SC_HIP(hipMalloc(&b.parts, (size_t)kStagePos * 2 * nstream * sizeof(double)));
  SC_HIP(hipMalloc(&b.scal, kStagePos * sizeof(CgScal)));
  SC_HIP(hipMemset(b.scal, 0, kStagePos * sizeof(CgScal)));
  SC_HIP(hipMalloc(&b.alt, 4 * (size_t)Nz * sizeof(double)));
  SC_HIP(hipMalloc(&b.gram, gram_doubles() * sizeof(double)));
  b.nz = Nz;
  b.nstream = nstream;
  return SC_OK;
}

void stage_bufs_free(QdStageBufs& b) {
  if (b.plan) (void)hipFree(b.plan);
  if (b.parts) (void)hipFree(b.parts);
  if (b.scal) (void)hipFree(b.scal);
  if (b.alt) (void)hipFree(b.alt);
  if (b.gram) (void)hipFree(b.gram);
  b = QdStageBufs{};
}

bool stage_fits(const Ctx& c) { return c.d.Nz <= kStageNzMax && c.projCount > 0; }

__global__ __launch_bounds__(256) void k_resid_norms(const double* __restrict__ b, const double* __restrict__ Au,
                                                     size_t N, double* __restrict__ parts) {
  double v[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
    const double r = b[i] - Au[i];
    v[0] += r * r;
    v[1] += b[i] * b[i];
  }
  block_partials(v, parts);
}

int launch_resid_norms(const double* b, const double* Au, size_t N, double* parts, int nb, hipStream_t s) {
  if (nb < 1 || nb > 1024) return fail(SC_ERR_ARG, "resid norms: 1..1024 blocks");
  hipLaunchKernelGGL(k_resid_norms, dim3(nb), dim3(256), 0, s, b, Au, N, parts);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

PendingFinish stage_finish(const QdStageBufs& b, int pos, int* miss) {
  return PendingFinish{(CgScal*)b.scal + pos, b.parts + (size_t)pos * 2 * b.nstream, b.nstream, miss};
}

struct GramPtrs {
  double *G, *part;
};
static GramPtrs gram_ptrs(const QdStageBufs& b) {
  constexpr size_t R = SC_REUSE_MAX;
  GramPtrs p;
  p.G = b.gram;
  p.part = p.G + 2 * R * R;
  return p;
}

int launch_stage_plan(const Ctx& c, QdStageBufs& b, const double* rfix, hipStream_t s) {
  const int M = c.projCount, Nz = c.d.Nz;
  if (!stage_fits(c)) return fail(SC_ERR_ARG, "stage kernel: no basis or Nz too large");
  int nLow, nHigh;
  qd_band_rows(c.d.Ny, &nLow, &nHigh);
  VecList Y{}, Z{};
  for (int m = 0; m < M; ++m) {
    Y.v[m] = c.projY[m];
    Z.v[m] = c.projZ[m];
  }
  double* PL = b.plan;
  double* PR = b.plan + (size_t)SC_REUSE_MAX * Nz;
  double* sR = b.plan + 2 * (size_t)SC_REUSE_MAX * Nz;
  double* aL = sR + 2 * SC_REUSE_MAX;
  double* aR = aL + 2 * (size_t)kBand * Nz;
  GramPtrs gp = gram_ptrs(b);
  const size_t zLo = (size_t)nLow * Nz, zHi = (size_t)(c.d.Ny - nHigh) * Nz;
  const dim3 g((Nz + 255) / 256), bl(256);
  switch (M) {
#define SC_SP(m)                                                                                    \
  case m:                                                                                           \
    hipLaunchKernelGGL(k_stage_plan<m>, g, bl, 0, s, c.op64, nLow, nHigh, Y, PL, PR, aL, aR);       \
    hipLaunchKernelGGL(k_stage_plan_sums<m>, dim3(1), dim3(256), 0, s, (const double*)PR, rfix, Nz, sR); \
    hipLaunchKernelGGL(k_stage_gram<m>, dim3(kGramBlocks), dim3(256), 0, s, Z, zLo, zHi, gp.part);  \
    hipLaunchKernelGGL(k_stage_gram_sum<m>, dim3(1), dim3(64), 0, s, (const double*)gp.part, gp.G); \
    break;
    SC_SP(1) SC_SP(2) SC_SP(3) SC_SP(4) SC_SP(5) SC_SP(6) SC_SP(7) SC_SP(8)
#undef SC_SP
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int launch_qd_stage(const Ctx& c, const QdStageBufs& b, int pos, const QdStageIn& in, hipStream_t s,
                    std::vector<unsigned char>* rec, hipEvent_t evStart, hipEvent_t evStop) {
  const int M = c.projCount, Nz = c.d.Nz;
  if (!stage_fits(c) || b.nz != Nz || pos < 0 || pos >= kStagePos)
    return fail(SC_ERR_ARG, "stage kernel: buffers or position");
  if (in.L.nv && (in.L.nv != 2 || in.L.nt > kStageNT || in.psiIdx < 0 || in.slipIdx < 0 || in.psiIdx == in.slipIdx))
    return fail(SC_ERR_ARG, "stage kernel: the combination must hold psi and slip with at most 2 terms");
  // the arguments live in zeroed storage so that their bytes (padding
  // included) can serve as a graph key
  alignas(QdStageArgs) unsigned char abuf[sizeof(QdStageArgs)];
  std::memset(abuf, 0, sizeof abuf);
  QdStageArgs& a = *new (abuf) QdStageArgs{};
  int nLow, nHigh;
  qd_band_rows(c.d.Ny, &nLow, &nHigh);
  a.Nz = Nz;
  a.nfb = (Nz + 63) / 64;
  a.nstream = b.nstream;
  a.N = (size_t)c.N;
  a.zLo = (size_t)nLow * Nz;
  a.zHi = (size_t)(c.d.Ny - nHigh) * Nz;
  a.aL = b.plan + 2 * (size_t)SC_REUSE_MAX * Nz + 2 * SC_REUSE_MAX;
  a.aR = a.aL + 2 * (size_t)kBand * Nz;
  a.G = gram_ptrs(b).G;
  for (int m = 0; m < M; ++m) {
    a.Z.v[m] = c.projZ[m];
    a.T.v[m] = c.projT[m];
  }
  a.w = c.projW;
  a.PL = b.plan;
  a.sR = b.plan + 2 * (size_t)SC_REUSE_MAX * Nz;
  a.coefOut = c.projW + 2 * SC_REUSE_MAX;
  a.parts = b.parts + (size_t)pos * 2 * b.nstream;
  a.sc = (CgScal*)b.scal + pos;
  a.tol2 = in.tol2;
  a.fuse = in.fuse;
  a.L = in.L;
  {
    // the stage's operands, resolved here (no indexed argument loads in the kernel)
    const bool comb = in.L.nv > 0;
    a.ntS = comb ? in.L.nt : 0;
    a.sBase = comb ? in.L.base[in.slipIdx] : in.slipVar;
    a.pBase = comb ? in.L.base[in.psiIdx] : in.fuse.a.psi_d;
    for (int u = 0; u < kStageNT; ++u) {
      a.sF[u] = comb && u < in.L.nt ? in.L.f[in.slipIdx][u] : a.sBase;
      a.pF[u] = comb && u < in.L.nt ? in.L.f[in.psiIdx][u] : a.pBase;
    }
    a.sOut = comb ? in.L.out[in.slipIdx] : nullptr;
    a.pOut = comb ? in.L.out[in.psiIdx] : nullptr;
    if (!in.end && (!a.sBase || !a.pBase))
      return fail(SC_ERR_ARG, "stage kernel: the stage's slip or psi is missing");
  }
  a.psiIdx = in.psiIdx;
  a.slipIdx = in.slipIdx;
  a.slipVar = in.slipVar;
  a.fts = in.fts;
  a.vL = in.vL;
  a.remote = in.remote;
  a.rfix = in.rfix;
  a.bcL = in.bcL;
  a.bcR = in.bcR;
  a.gate = in.gate;
  a.tfrac = in.tfrac;
  a.t0 = in.t0;
  a.fin = in.fin;
  if (in.end) {
    if (!in.ctlOut || !in.gate.ctl || !in.flags || !in.flagEnd || !in.fin3.sc || in.fin.sc || in.L.nv)
      return fail(SC_ERR_ARG, "stage kernel: an attempt's end needs its control blocks, flags and test");
    // its own root solves must not report into a word that it reads
    if (in.fuse.a.status_d == in.flags || in.fuse.a.status_d == in.flagEnd)
      return fail(SC_ERR_ARG, "stage kernel: an attempt's end reports its root solves into a word it reads");
    a.ctlIn = in.gate.ctl;
    a.ctlOut = in.ctlOut;
    a.flags = in.flags;
    a.flagEnd = in.flagEnd;
    a.fin3 = in.fin3;
    if (in.fin3.nb < 1 || !in.fin3.partials)
      return fail(SC_ERR_ARG, "stage kernel: an attempt's end needs the last stage's residual partials");
    for (int f = 0; f < 2; ++f) {
      a.pv[f] = in.pv[f];
      a.pvOut[f] = in.pvOut[f];
      if (!a.pv[f] || !a.pvOut[f] || a.pv[f] == a.pvOut[f])
        return fail(SC_ERR_ARG, "stage kernel: an attempt's end reads and writes different states");
      for (int r = 0; r < 3; ++r) {
        a.pf[f][r] = in.pf[f][r];
        if (!a.pf[f][r]) return fail(SC_ERR_ARG, "stage kernel: an attempt's end needs its three stages' rates");
      }
      a.selSlot[f] = in.selSlot[f];
    }
    for (int i = 0; i < 2; ++i) {
      a.clo[i] = in.clo[i];
      a.cdlo[i] = in.cdlo[i];
    }
    for (int i = 0; i < 3; ++i) {
      a.chi[i] = in.chi[i];
      a.cdhi[i] = in.cdhi[i];
    }
  }
  const dim3 g(a.nfb + a.nstream + (in.fin.sc && !in.end ? 1 : 0)), bl(kStageThreads);
  if (rec) {
    const int head[3] = {M, in.end ? 1 : 0, (int)g.x};
    rec->insert(rec->end(), (const unsigned char*)head, (const unsigned char*)head + sizeof head);
    rec->insert(rec->end(), abuf, abuf + sizeof abuf);
    return SC_OK;
  }
  const bool ev = evStart || evStop;
  const bool kq2 = Nz <= 2 * kStageThreads;
  switch (M) {
#define SC_STQ(m, q)                                                                                  \
  if (ev) {                                                                                           \
    if (in.end) hipExtLaunchKernelGGL((k_qd_stage<m, true, q>), g, bl, 0, s, evStart, evStop, 0, a); \
    else hipExtLaunchKernelGGL((k_qd_stage<m, false, q>), g, bl, 0, s, evStart, evStop, 0, a);       \
  } else if (in.end) {                                                                                \
    hipLaunchKernelGGL((k_qd_stage<m, true, q>), g, bl, 0, s, a);                                     \
  } else {                                                                                            \
    hipLaunchKernelGGL((k_qd_stage<m, false, q>), g, bl, 0, s, a);                                    \
  }
#define SC_ST(m)       \
  case m:              \
    if (kq2) {         \
      SC_STQ(m, 2)     \
    } else {           \
      SC_STQ(m, 4)     \
    }                  \
    break;
    SC_ST(1) SC_ST(2) SC_ST(3) SC_ST(4) SC_ST(5) SC_ST(6) SC_ST(7) SC_ST(8)
#undef SC_ST
#undef SC_STQ
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc
