// Device parts of the RK32 attempt's end (the tail of the attempt loop of
// odeSolver.cpp:505-627): the lo/hi pair with its error sums, the step-size
// controller and the accept/reject update of the control block, run by
// k_vars_lohi_err and k_step_end (qd_kernels.hip).
#pragma once
#include "cg_dev.hpp"

namespace sc {

// NTLO / NTHI terms (the RK32 pair: 2 and 3), kChunk entries per thread
// with every load of a chunk issued before its stores (one memory round trip
// per chunk); the sums run in entry order as the plain loop's
// keep (optional, kLoHiChunk entries): hi of this thread's first chunk
// (entries tid + q kRedThreads), the whole field when n <= kLoHiChunk kRedThreads
constexpr int kLoHiChunk = 4;
template <int NTLO, int NTHI>
__device__ inline void lohi_accum_t(const VarsLoHi& L, int v, const double* clo, const double* chi, int tid,
                                    double& a, double& b, double& m, double* keep) {
  constexpr int kChunk = kLoHiChunk;
  const int n = L.n[v];
  for (int i0 = tid; i0 < n; i0 += kChunk * kRedThreads) {
    double base[kChunk], flo[kChunk][NTLO], fhi[kChunk][NTHI];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int i = i0 + q * kRedThreads;
      if (i < n) {
        base[q] = L.base[v][i];
#pragma unroll
        for (int t = 0; t < NTLO; ++t) flo[q][t] = L.flo[v][t][i];
#pragma unroll
        for (int t = 0; t < NTHI; ++t) fhi[q][t] = L.fhi[v][t][i];
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int i = i0 + q * kRedThreads;
      if (i < n) {
        double lo = clo[0] * flo[q][0] + base[q];
#pragma unroll
        for (int t = 1; t < NTLO; ++t) lo = lo + clo[t] * flo[q][t];
        double hi = chi[0] * fhi[q][0] + base[q];
#pragma unroll
        for (int t = 1; t < NTHI; ++t) hi = hi + chi[t] * fhi[q][t];
        const double e = -1.0 * hi + lo;
        a += e * e;
        b += hi * hi;
        m = fmax(m, fabs(e) / hi);
        flo[q][0] = lo;
        fhi[q][0] = hi;
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int i = i0 + q * kRedThreads;
      if (i < n) {
        L.lo[v][i] = flo[q][0];
        L.hi[v][i] = fhi[q][0];
        if (keep && i0 == tid) keep[q] = fhi[q][0];
      }
    }
  }
}

// Field v of the pair at local thread tid (0..kRedThreads-1): lo = base +
// sum_t clo[t] flo[t], hi likewise (the order of k_vars_lincomb), written
// out, and this thread's share of [sum (lo-hi)^2, sum hi^2, max |lo-hi|/hi]
// into s0[tid], s1[tid], s2[tid]
__device__ inline void lohi_accum(const VarsLoHi& L, int v, const double* clo, const double* chi, int tid,
                                  double* s0, double* s1, double* s2, double* keep = nullptr) {
  double a = 0, b = 0, m = -INFINITY;
  if (L.ntlo == 2 && L.nthi == 3) {
    lohi_accum_t<2, 3>(L, v, clo, chi, tid, a, b, m, keep);
  } else {
    const int n = L.n[v];
    for (int i = tid; i < n; i += kRedThreads) {
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
  }
  s0[tid] = a;
  s1[tid] = b;
  s2[tid] = m;
}

// The tree over kRedThreads shares of lohi_accum, the sums into out[0..2].
// Every thread of the block calls it (a barrier per level); threads with
// act = false only take the barriers.
__device__ inline void lohi_tree(double* s0, double* s1, double* s2, int tid, bool act, double* out) {
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (act && tid < w) {
      s0[tid] += s0[tid + w];
      s1[tid] += s1[tid + w];
      s2[tid] = fmax(s2[tid], s2[tid + w]);
    }
    __syncthreads();
  }
  if (act && tid == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
    out[2] = s2[0];
  }
}

// The pair's coefficients: constants, or under a step gate templates of the
// control block's step h, (c[t] h) / cd[t] (the order of the host's h/6.0)
__device__ inline void lohi_coefs(const VarsLoHi& L, const StepCtl* ctl, double* clo, double* chi) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    clo[i] = L.clo[i];
    chi[i] = L.chi[i];
  }
  if (!ctl) return;
  const double h = ctl->dt;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    clo[i] = (L.clo[i] * h) / L.cdlo[i];
    chi[i] = (L.chi[i] * h) / L.cdhi[i];
  }
}

// computeStepSize (odeSolver.cpp:395-445) on the control block's state, with
// the host's std::min / std::max order (a NaN ratio keeps the bound)
__device__ inline double step_size_dev(const StepCtl& c, double totErr) {
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
// update of odeSolver.cpp:587-627) on the control block c (one thread): err
// holds the selected fields' sums, flagged a set fault status or missed
// guess, which halts the loop with the state of the attempt's start.
// Returns whether the attempt was accepted (its solution goes into var).
__device__ inline bool step_control_update(StepCtl* c, const double* err, bool flagged) {
  c->accepted = 0;
  if (c->halt || c->done) return false;
  if (flagged) {
    c->halt = 1;
    return false;
  }
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
  if (!accept) return false;
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
  return true;
}

}  // namespace sc
