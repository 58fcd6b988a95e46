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
// out, and this thread's share (a, b, m) of [sum (lo-hi)^2, sum hi^2,
// max |lo-hi|/hi]
__device__ inline void lohi_accum(const VarsLoHi& L, int v, const double* clo, const double* chi, int tid,
                                  double& a, double& b, double& m, double* keep = nullptr) {
  a = 0;
  b = 0;
  m = -INFINITY;
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
}

// The sums over a group of kRedThreads threads (kRedThreads / 64 waves,
// group-aligned) of their lohi_accum shares: a butterfly in each wave, then
// the wave values in wave order, into out[0..2].  sh: kRedThreads / 64 x 3
// doubles of LDS for the group.  Every thread of the block calls it (one
// barrier); threads with act = false only take the barrier.
__device__ inline void lohi_reduce(double a, double b, double m, int tid, bool act, double* sh, double* out) {
  constexpr int kW = kRedThreads / 64;
  a = wave_sum(a);
  b = wave_sum(b);
  m = wave_max(m);
  const int w = tid >> 6;
  if (act && (tid & 63) == 0) {
    sh[w] = a;
    sh[kW + w] = b;
    sh[2 * kW + w] = m;
  }
  __syncthreads();
  if (act && tid == 0) {
    double sa = sh[0], sb = sh[kW], sm = sh[2 * kW];
    for (int q = 1; q < kW; ++q) {
      sa += sh[q];
      sb += sh[kW + q];
      sm = fmax(sm, sh[2 * kW + q]);
    }
    out[0] = sa;
    out[1] = sb;
    out[2] = sm;
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

// computeStepSize (odeSolver.cpp:395-445) by one wave: the controller's
// powers run on lanes 0..2 side by side (P control: lane 0 only) and the
// ratio multiplies them in the host's order; every lane returns the step.
// The host's std::min / std::max order (a NaN ratio keeps the bound).
__device__ inline double step_size_wave(bool pCtl, int ord, double kappa, double tol, double errA0, double errA1,
                                        double dt0, double minDt, double maxDt, double totErr) {
  const int lane = threadIdx.x & 63;
  double base = tol / totErr, ex = pCtl ? 1. / (1. + ord) : 0.49 / ord;
  if (!pCtl && lane == 1) {
    base = errA0 / tol;
    ex = 0.34 / ord;
  } else if (!pCtl && lane == 2) {
    base = tol / errA1;
    ex = 0.1 / ord;
  }
  const double pw = pow(base, ex);
  const double p0 = __shfl(pw, 0, 64), p1 = __shfl(pw, 1, 64), p2 = __shfl(pw, 2, 64);
  const double ratio = pCtl ? kappa * p0 : kappa * p0 * p1 * p2;
  double dt = ratio * dt0;
  dt = (dt < dt0 * 5.0) ? dt : dt0 * 5.0;
  dt = (dt < maxDt) ? dt : maxDt;
  dt = (minDt < dt) ? dt : minDt;
  return dt;
}

// One attempt's end (the tail of the RK32 attempt loop and the accepted-step
// update of odeSolver.cpp:587-627) on the control block c (in LDS), run by
// the 64 lanes of one wave on register copies of its fields, lane 0 writing
// the changed ones back: err holds the selected fields' sums, flagged a set
// fault status or missed guess, which halts the loop with the state of the
// attempt's start.  Returns (on every lane) whether the attempt was accepted
// (its solution goes into var).  The proposed step is the same for a
// rejected attempt's retry and an accepted step's next size (computeStepSize
// sees the same state in both), so it is computed once; an attempt accepted
// at minDt or the attempt limit proposes again from the step it took.
// computeError's total over the selected fields from their sums (err: 3 per
// field)
__device__ inline double step_error_total(const StepCtl* c, const double* err) {
  const int ord = c->ord, nsel = c->nsel, normType = c->normType;
  double tot = 0;
  for (int q = 0; q < nsel; ++q) {
    const double* v = err + 3 * q;
    if (normType == 0) tot += sqrt(v[0]) / (sqrt((double)c->selN[q]) * 1.0);
    else if (ord == 3) tot += sqrt(v[0]) / (sqrt(v[1]) * 1.0);
    else tot += v[2] / 1.0;
  }
  return tot;
}

// the attempt is accepted by its error alone (not flagged, not forced)
__device__ inline bool step_accepts(const StepCtl* c, const double* err) {
  const double tot = step_error_total(c, err);
  return c->ord == 3 ? tot <= c->tol : tot < c->tol;
}

__device__ inline bool step_control_wave(StepCtl* c, const double* err, bool flagged, bool writeDts = true) {
  const bool lead = (threadIdx.x & 63) == 0;
  if (c->halt || c->done) {
    if (lead) c->accepted = 0;
    return false;
  }
  if (flagged) {
    if (lead) {
      c->accepted = 0;
      c->halt = 1;
    }
    return false;
  }
  const int ord = c->ord;
  const double tot = step_error_total(c, err);
  const double tol = c->tol, minDt = c->minDt, finalT = c->finalT;
  double t = c->t, dt = c->dt, newDt = c->newDt;
  int attempts = c->attempts + 1, rejected = c->rejected, stepCount = c->stepCount;
  const bool pCtl = c->control == 0 || stepCount - c->pidBase < 4;
  const double kappa = c->kappa, errA0 = c->errA0, errA1 = c->errA1, maxDt = c->maxDt;
  const double sz = step_size_wave(pCtl, ord, kappa, tol, errA0, errA1, dt, minDt, maxDt, tot);
  bool accept = ord == 3 ? tot <= tol : tot < tol, forced = false;
  if (!accept) {
    dt = sz;
    rejected++;
    if (minDt == dt || attempts >= 100) accept = forced = true;
    else if (t + dt > finalT) dt = finalT - t;
  }
  if (!accept) {
    if (lead) {
      c->accepted = 0;
      c->attempts = attempts;
      c->dt = dt;
      c->rejected = rejected;
    }
    return false;
  }
  t = t + dt;
  // a forced acceptance proposes from the step it just took
  if (tot != 0.0) newDt = forced ? step_size_wave(pCtl, ord, kappa, tol, errA0, errA1, dt, minDt, maxDt, tot) : sz;
  const int nOut = c->nOut;
  if (lead) {
    if (writeDts && nOut < c->dtsCap) c->dts[nOut] = dt;
    c->errA1 = errA0;
    c->errA0 = tot;
  }
  dt = newDt;
  int done = 0;
  if (!(stepCount < c->maxSteps && t < finalT)) {
    done = 1;
  } else {
    stepCount++;
    if (t + dt > finalT) dt = finalT - t;
  }
  if (lead) {
    c->t = t;
    c->newDt = newDt;
    c->nOut = nOut + 1;
    c->dt = dt;
    c->accepted = 1;
    c->attempts = 0;
    c->rejected = rejected;
    c->done = done;
    c->stepCount = stepCount;
  }
  return true;
}

}  // namespace sc
