// Device-side scalars of the CG solves and the residual test of a projected
// guess run by a whole block of another kernel (cg.hip launches it on its
// own as k_finish<FIN_IRRES, 2>; a deferred check may instead run as the
// extra block of the next launch on the stream, see PendingFinish).
#pragma once
#include "reduce_dev.hpp"
#include "sc_internal.hpp"

namespace sc {

struct CgScal {
  double rz, pq, rr, bb, alpha, beta, tol2;
  int done, iters, maxit, pad;
  unsigned int cnt;
};

// rr and bb from the partials of f.nb blocks, the test rr <= tol2 bb and the
// missed-guess flag: the arithmetic and summation order of
// k_finish<FIN_IRRES, 2> (the first kRedThreads threads accumulate, the rest
// of a larger block add nothing), so the flag is the same bit for bit.
// Every thread of the block must call it.
__device__ inline void finish_irres_block(const PendingFinish& f) {
  __shared__ double sh[2][kRedThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double w0 = 0.0, w1 = 0.0;
  if (tid < kRedThreads)
    for (int b = tid; b < f.nb; b += kRedThreads) {
      w0 += f.partials[(size_t)b * 2];
      w1 += f.partials[(size_t)b * 2 + 1];
    }
  const double s0 = wave_sum(w0), s1 = wave_sum(w1);
  if (lane == 0 && wid < kRedThreads / 64) {
    sh[0][wid] = s0;
    sh[1][wid] = s1;
  }
  __syncthreads();
  if (tid == 0) {
    double t0 = 0, t1 = 0;
    for (int ww = 0; ww < kRedThreads / 64; ++ww) t0 += sh[0][ww];
    for (int ww = 0; ww < kRedThreads / 64; ++ww) t1 += sh[1][ww];
    CgScal* sc = f.sc;
    sc->rr = t0;
    sc->bb = t1;
    sc->done = (t0 <= sc->tol2 * t1 || t1 == 0.0) ? 1 : 0;
    if (f.flag && !sc->done) *f.flag = 1;
  }
}

// true when a step kernel under this gate has nothing to do (uniform over
// the grid: the control block does not change during the launch)
__device__ __forceinline__ bool gate_closed(const StepGate& g) {
  if (!g.ctl) return false;
  if (g.ctl->halt) return true;
  return g.mode == 2 ? !g.ctl->accepted : (g.mode == 1 && g.ctl->done);
}

}  // namespace sc
