// Linear solve A u = rhs replacing KSPSolve in LinearElastic::computeU
// (linearElastic.cpp:465-489).
//
// SC_SOLVE_CG64: Jacobi-preconditioned CG in FP64 on the matrix-free operator.
// SC_SOLVE_IR32: FP64 iterative refinement; each sweep solves the correction
//                equation with PCG in FP32 (operator, vectors and
//                preconditioner in FP32, dot products accumulated in FP64):
//                Jacobi, or the fast-diagonalisation inverse with FP32 or
//                bf16 GEMMs (flexible PCG); the residual b - A u is always
//                formed in FP64.
//
// All scalars (alpha, beta, norms, the converged flag) stay on the device;
// every kernel of an iteration returns at once when the flag is set, so the
// host launches iterations in batches and only reads the flag between
// batches.  Reductions are two-stage and deterministic: per-block partial
// sums, then a one-block kernel sums the partials in block order and updates
// the scalars.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "cg_dev.hpp"
#include "fdiag.hpp"
#include "fault_dev.hpp"
#include "reduce_dev.hpp"
#include "sbp_dev.hpp"
#include "sc_internal.hpp"

namespace sc {

enum FinishOp { FIN_INIT, FIN_PQ, FIN_UPDATE, FIN_IRRES, FIN_UPDATE_R, FIN_RZ };

// Sum the partials of nb blocks in a fixed order (deterministic for a given
// grid) and update the CG scalars.  One block of kRedThreads threads.
template <int OP, int NV>
__global__ __launch_bounds__(kRedThreads) void k_finish(CgScal* sc, const double* __restrict__ partials,
                                                        int nb, int first, int* flag) {
  if ((OP == FIN_PQ || OP == FIN_UPDATE || OP == FIN_UPDATE_R || OP == FIN_RZ) && sc->done) return;
  __shared__ double sh[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double w[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) w[i] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRedThreads) {
#pragma unroll
    for (int i = 0; i < NV; ++i) w[i] += partials[(size_t)b * NV + i];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(w[i]);
    if (lane == 0) sh[i][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0;
    for (int ww = 0; ww < kRedThreads / 64; ++ww) s += sh[i][ww];
    t[i] = s;
  }
  if constexpr (OP == FIN_INIT) {  // rz, rr, bb
    sc->rz = t[0];
    sc->rr = t[1];
    sc->bb = t[2];
    sc->iters = 0;
    sc->done = (t[1] <= sc->tol2 * t[2] || t[2] == 0.0) ? 1 : 0;
  } else if constexpr (OP == FIN_PQ) {  // pq
    sc->pq = t[0];
    sc->alpha = sc->rz / t[0];
    if (!(t[0] != 0.0) || !isfinite(sc->alpha)) sc->done = 2;  // breakdown
  } else if constexpr (OP == FIN_UPDATE) {  // rz', rr
    sc->beta = t[0] / sc->rz;
    sc->rz = t[0];
    sc->rr = t[1];
    sc->iters += 1;
    if (t[1] <= sc->tol2 * sc->bb) sc->done = 1;
    else if (sc->iters >= sc->maxit) sc->done = 3;
    else if (!isfinite(t[1])) sc->done = 4;
  } else if constexpr (OP == FIN_IRRES) {  // rr, bb
    sc->rr = t[0];
    sc->bb = t[1];
    sc->done = (t[0] <= sc->tol2 * t[1] || t[1] == 0.0) ? 1 : 0;
    if (flag && !sc->done) *flag = 1;  // a deferred check records the miss
  } else if constexpr (OP == FIN_UPDATE_R) {  // rr
    sc->rr = t[0];
    sc->iters += 1;
    if (t[0] <= sc->tol2 * sc->bb) sc->done = 1;
    else if (sc->iters >= sc->maxit) sc->done = 3;
    else if (!isfinite(t[0])) sc->done = 4;
  } else {  // FIN_RZ: r.z, r.z_old
    sc->beta = first ? 0.0 : (t[0] - t[1]) / sc->rz;
    sc->rz = t[0];
    if (!isfinite(t[0])) sc->done = 4;
  }
}

// r = b - q ; z = dinv*r ; p = z ; rz = r.z ; rr = r.r ; bb = b.b
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_init(size_t N, const T* __restrict__ b,
                                                         const T* __restrict__ q,
                                                         const T* __restrict__ dinv,
                                                         T* __restrict__ r, T* __restrict__ p,
                                                         double* partials) {
  double v[3] = {0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const T bi = b[i];
    const T ri = bi - q[i];
    const T zi = dinv[i] * ri;
    r[i] = ri;
    p[i] = zi;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
    v[2] += (double)bi * (double)bi;
  }
  block_partials(v, partials);
}

// pq = p.q ; alpha = rz/pq
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_dot_pq(size_t N, const T* __restrict__ p,
                                                           const T* __restrict__ q, CgScal* sc,
                                                           double* partials) {
  if (sc->done) return;
  double v[1] = {0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    v[0] += (double)p[i] * (double)q[i];
  block_partials(v, partials);
}

// x += alpha p ; r -= alpha q ; rz' = r.(dinv r) ; rr = r.r ; beta = rz'/rz
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_update(size_t N, T* __restrict__ x,
                                                           T* __restrict__ r,
                                                           const T* __restrict__ p,
                                                           const T* __restrict__ q,
                                                           const T* __restrict__ dinv, CgScal* sc,
                                                           double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  double v[2] = {0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  }
  block_partials(v, partials);
}

// p = dinv r + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_cg_pvec(size_t N, const T* __restrict__ r,
                                                 const T* __restrict__ dinv, T* __restrict__ p,
                                                 const CgScal* sc) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// FP64 residual for refinement: r64 = b - q ; rr = r.r ; bb = b.b ;
// r32 = (float)(r64 * s) is done separately once the norm is known.
__global__ __launch_bounds__(kRedThreads) void k_ir_residual(size_t N, const double* __restrict__ b,
                                                             const double* __restrict__ q,
                                                             double* __restrict__ r,
                                                             double* partials) {
  double v[2] = {0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const double bi = b[i];
    const double ri = bi - q[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += bi * bi;
  }
  block_partials(v, partials);
}

__global__ __launch_bounds__(256) void k_ir_to32(size_t N, const double* __restrict__ r,
                                                 const CgScal* sc, float* __restrict__ r32,
                                                 float* __restrict__ d32) {
  const double s = 1.0 / sqrt(sc->rr);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    r32[i] = (float)(r[i] * s);
    d32[i] = 0.0f;
  }
}

// u += d32 * ||r||  (the correction was solved for the unit-norm residual)
__global__ __launch_bounds__(256) void k_ir_correct(size_t N, const float* __restrict__ d32,
                                                    const CgScal* sc, double* __restrict__ u) {
  const double s = sqrt(sc->rr);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    u[i] += (double)d32[i] * s;
}

// x += alpha p ; r -= alpha q ; rr = r.r   (preconditioner applied separately)
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_update_r(size_t N, T* __restrict__ x,
                                                             T* __restrict__ r,
                                                             const T* __restrict__ p,
                                                             const T* __restrict__ q,
                                                             CgScal* sc, double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  double v[1] = {0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
  }
  block_partials(v, partials);
}

// FP32 start of an inner solve from x = 0: r = b ; rr = r.r ; bb = b.b
__global__ __launch_bounds__(kRedThreads) void k_cg32_start(size_t N, const float* __restrict__ b,
                                                            float* __restrict__ x, float* __restrict__ r,
                                                            double* partials) {
  double v[2] = {0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const float bi = b[i];
    x[i] = 0.0f;
    r[i] = bi;
    v[0] += (double)bi * (double)bi;
    v[1] += (double)bi * (double)bi;
  }
  block_partials(v, partials);
}

// Flexible (Polak-Ribiere) beta, robust to a preconditioner that is only
// approximately symmetric (FP32 / bf16 GEMMs):
//   beta = r.(z - z_old) / rz_old ; rz = r.z ; z_old <- z
// (first call: beta = 0, p = z)
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_rz(size_t N, const T* __restrict__ r,
                                                       const T* __restrict__ z,
                                                       T* __restrict__ zold, int first,
                                                       CgScal* sc, double* partials) {
  if (sc->done) return;
  double v[2] = {0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const double ri = r[i], zi = z[i];
    v[0] += ri * zi;
    if (!first) v[1] += ri * (double)zold[i];
    zold[i] = z[i];
  }
  block_partials(v, partials);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_pz(size_t N, const T* __restrict__ z,
                                               T* __restrict__ p, const CgScal* sc, int first) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

// ----------------------------------------------- initial guess from reuse --
// Fischer's projection for successive right-hand sides: the basis y_i is
// A-orthogonal and w_i = 1 / (y_i^T (-A) y_i) (A is negative definite), so
// the (-A)-norm best approximation of A^-1 b in span{y_i} is
//   x0 = sum_i w_i (y_i^T (-A) x*) y_i = sum_i (-w_i y_i^T b) y_i.
// After a solve with solution x (A x = b - r) the new basis vector is
//   y = x - sum_i beta_i y_i,  beta_i = -w_i y_i^T (b - r),
// and y^T (-A) y = -y^T (b - r) since y is A-orthogonal to every y_i.

// partial sums of y_m . (a - b)   (b may be null)
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_multidot(size_t N, const double* __restrict__ a,
                                                          const double* __restrict__ b, VecList Y,
                                                          double* partials) {
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const double v = b ? a[i] - b[i] : a[i];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += Y.v[m][i] * v;
  }
  block_partials(acc, partials);
}

// coef[m] = -w[m] * sum of the partials of column m (fixed order)
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_proj_coef(const double* __restrict__ partials, int nb,
                                                           double* wc) {
  __shared__ double sh[M][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRedThreads) {
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += partials[(size_t)b * M + m];
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const double t = wave_sum(acc[m]);
    if (lane == 0) sh[m][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double t = 0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += sh[threadIdx.x][w];
    wc[SC_REUSE_MAX + threadIdx.x] = -wc[threadIdx.x] * t;
  }
}

// partial sums of y_m . b over the boundary frame only, for a b that is zero
// elsewhere
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_multidot_frame(FrameMap fm, int Nz, const double* __restrict__ b,
                                                                VecList Y, double* partials) {
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  const long long n = fm.count();
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int j, k;
    fm.node(t, j, k);
    const size_t i = (size_t)j * Nz + k;
    const double v = b[i];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += Y.v[m][i] * v;
  }
  block_partials(acc, partials);
}

// The projected guess and its residual in one pass.  Every block first sums
// the mnb multidot partials y_m . b in the same fixed order, so all blocks
// hold the same coef[m] = -w[m] y_m . b; then x = sum_m coef[m] y_m and
// r = b - sum_m coef[m] z_m with z_m = A y_m kept beside the basis, partial
// sums of r.r and b.b (r itself is not stored: a solve that goes on from x
// recomputes it).  The first streaming block also starts the residual check
// (tolerance tol2).  FAULT: the first nfb blocks instead compute the fault
// rates of the QD step from tau = sum_m coef[m] T_m (QdFuse).
constexpr int kFaultWave = 64;  // fault nodes per fault block (its first wave)
struct GuessFault {
  QdFuse f;
  VecList T;
  int nfb;
};

// WX = false (a deferred guess whose caller reads the solution only later):
// x is not written (nor Y read); block 0 leaves the coefficients in coefOut
// and proj_materialize forms x = sum c_m y_m from them when it is needed.
// b is known to vanish on [zLo, zHi) (zLo = zHi: nowhere) and is not read
// there.
template <int M, bool FAULT, bool WX>
__global__ __launch_bounds__(kRedThreads) void k_proj_guess_res(size_t N, VecList Y, VecList Z,
                                                                const double* __restrict__ mpart, int mnb,
                                                                const double* __restrict__ w, CgScal* sc,
                                                                double tol2, const double* __restrict__ b,
                                                                double* __restrict__ x, double* partials,
                                                                GuessFault gf, StepGate gate, size_t zLo,
                                                                size_t zHi, double* coefOut) {
  __shared__ double shc[M][kRedThreads / 64];
  __shared__ double coef[M];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // Streaming blocks: kU entries per pass with every load of a pass issued
  // before its arithmetic; the first pass's loads go out before the
  // projection sums (they do not depend on the coefficients), and the sums
  // run in entry order as a plain strided loop's.
  constexpr int kU = M <= 2 ? 4 : (M <= 4 ? 2 : 1);
  const int sblk = (int)blockIdx.x - (FAULT ? gf.nfb : 0);  // streaming block index, < 0: a fault block
  const unsigned nstream = gridDim.x - (FAULT ? gf.nfb : 0);
  const size_t stride = (size_t)nstream * blockDim.x;
  double zv[kU][M], yv[kU][WX ? M : 1], bv[kU];
  auto load_pass = [&](size_t i0) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t i = i0 + u * stride;
      if (i < N) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          zv[u][m] = Z.v[m][i];
          if constexpr (WX) yv[u][m] = Y.v[m][i];
        }
        bv[u] = (i >= zLo && i < zHi) ? 0.0 : b[i];
      }
    }
  };
  const size_t iFirst = sblk >= 0 ? (size_t)sblk * blockDim.x + threadIdx.x : N;
  load_pass(iFirst);
  {
    // the projections' partials load while the gate is read (they are in
    // bounds either way)
    double a[M];
#pragma unroll
    for (int m = 0; m < M; ++m) a[m] = 0.0;
    for (int q = threadIdx.x; q < mnb; q += kRedThreads) {
#pragma unroll
      for (int m = 0; m < M; ++m) a[m] += mpart[(size_t)q * M + m];
    }
    if (gate_closed(gate)) return;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double t = wave_sum(a[m]);
      if (lane == 0) shc[m][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < M) {
      double t = 0;
      for (int ww = 0; ww < kRedThreads / 64; ++ww) t += shc[threadIdx.x][ww];
      coef[threadIdx.x] = -w[threadIdx.x] * t;
    }
    __syncthreads();
  }
  double c[M];
#pragma unroll
  for (int m = 0; m < M; ++m) c[m] = coef[m];
  int blk = blockIdx.x;
  if constexpr (FAULT) {
    if (blk < gf.nfb) {
      // one wave of fault nodes per block: the root solves are FP64 VALU
      // work, spread over as many CUs as there are waves of nodes
      const int k = blk * kFaultWave + threadIdx.x;
      if (threadIdx.x < kFaultWave && k < gf.f.a.N) {
        double tq = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) tq += c[m] * gf.T.v[m][k];
        gf.f.dslip[k] = fault_qd_node(gf.f.a, k, tq);
      }
      return;
    }
    blk -= gf.nfb;
  }
  if (blk == 0 && threadIdx.x == 0) {
    sc->tol2 = tol2;
    sc->maxit = 0;
    sc->done = 0;
    sc->iters = 0;
    sc->cnt = 0;
  }
  if (!WX && blk == 0 && threadIdx.x < M) coefOut[threadIdx.x] = c[threadIdx.x];
  double acc[2] = {0.0, 0.0};
  for (size_t i0 = iFirst; i0 < N; i0 += kU * stride) {
    if (i0 != iFirst) load_pass(i0);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t i = i0 + u * stride;
      if (i < N) {
        double v = 0.0, az = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if constexpr (WX) v += c[m] * yv[u][m];
          az += c[m] * zv[u][m];
        }
        const double bi = bv[u];
        const double ri = bi - az;
        if constexpr (WX) x[i] = v;
        acc[0] += ri * ri;
        acc[1] += bi * bi;
      }
    }
  }
  block_partials_at(acc, partials, blk);
}

// x = sum_m coef[m] y_m: the x a k_proj_guess_res<M, *, false> launch left
// unwritten, with its arithmetic (the same value bit for bit)
template <int M>
__global__ __launch_bounds__(256) void k_proj_x(size_t N, VecList Y, const double* __restrict__ coef,
                                                double* __restrict__ x) {
  double c[M];
#pragma unroll
  for (int m = 0; m < M; ++m) c[m] = coef[m];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) v += c[m] * Y.v[m][i];
    x[i] = v;
  }
}

// fault traction of basis vector k: T = mu Dy y on row 0 (the QD step's S y)
__global__ __launch_bounds__(256) void k_basis_traction(DevOp<double> p, const double* __restrict__ y,
                                                        double* __restrict__ T) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.Nz) return;
  T[k] = p.mu[k] * dy_row0(p, y, k);
}

// New basis vector, A-orthogonal to the others: y = x - sum_m coef[m] y_m and
// z = A x - sum_m coef[m] z_m (= A y), partial sums of y.z and x.(A x).  y, z
// may be the slot being replaced: every element is read before it is written
// by the same thread.
template <int M>
__global__ __launch_bounds__(kRedThreads) void k_proj_update_az(size_t N, VecList Y, VecList Z,
                                                                const double* __restrict__ wc,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ ax,
                                                                double* y, double* z, double* partials) {
  double c[M];
#pragma unroll
  for (int m = 0; m < M; ++m) c[m] = wc[SC_REUSE_MAX + m];
  double acc[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const double xi = x[i], axi = ax[i];
    double v = xi, w = axi;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      v -= c[m] * Y.v[m][i];
      w -= c[m] * Z.v[m][i];
    }
    y[i] = v;
    z[i] = w;
    acc[0] += v * w;
    acc[1] += xi * axi;
  }
  block_partials(acc, partials);
}

// w[slot] = 1 / (y^T (-A) y), or 0 when y vanishes against x^T (-A) x (x was
// already in the span to rounding): the slot is then ignored.  Partials hold
// y.(b - r) and, for NV = 2, x.(b - r) (NV = 1: y = x).
template <int NV>
__global__ __launch_bounds__(kRedThreads) void k_proj_weight(const double* __restrict__ partials, int nb,
                                                             double* wc, int slot) {
  __shared__ double sh[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRedThreads) {
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] += partials[(size_t)b * NV + i];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double t = wave_sum(acc[i]);
    if (lane == 0) sh[i][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      t[i] = 0;
      for (int w = 0; w < kRedThreads / 64; ++w) t[i] += sh[i][w];
    }
    const double nu2 = -t[0], mu2 = -t[NV - 1];
    wc[slot] = (nu2 > 1e-20 * mu2 && isfinite(nu2) && mu2 > 0.0) ? 1.0 / nu2 : 0.0;
  }
}

template <typename T>
__global__ void k_set_scal(CgScal* sc, double tol2, int maxit) {
  sc->tol2 = tol2;
  sc->maxit = maxit;
  sc->done = 0;
  sc->iters = 0;
  sc->cnt = 0;
}

__global__ __launch_bounds__(kRedThreads) void k_pending_finish(PendingFinish f) { finish_irres_block(f); }

int launch_pending_finish(hipStream_t s, const PendingFinish& f) {
  if (!f.sc) return SC_OK;
  hipLaunchKernelGGL(k_pending_finish, dim3(1), dim3(kRedThreads), 0, s, f);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

namespace {

int ensure_workspace(Ctx& c, bool need32) {
  const size_t N = (size_t)c.N;
  if (!c.r64) {
    SC_HIP(hipMalloc(&c.r64, N * sizeof(double)));
    SC_HIP(hipMalloc(&c.p64, N * sizeof(double)));
    SC_HIP(hipMalloc(&c.q64, N * sizeof(double)));
    SC_HIP(hipMalloc(&c.partials, (size_t)kRedBlocks * 4 * sizeof(double)));
    SC_HIP(hipMalloc(&c.scal, 2 * sizeof(CgScal)));
    SC_HIP(hipHostMalloc(&c.scal_h, 2 * sizeof(CgScal), hipHostMallocDefault));
    SC_HIP(hipMemsetAsync(c.scal, 0, 2 * sizeof(CgScal), c.stream));
  }
  if (need32 && !c.r32) {
    SC_HIP(hipMalloc(&c.r32, N * sizeof(float)));
    SC_HIP(hipMalloc(&c.p32, N * sizeof(float)));
    SC_HIP(hipMalloc(&c.q32, N * sizeof(float)));
    SC_HIP(hipMalloc(&c.d32, N * sizeof(float)));
  }
  return SC_OK;
}

unsigned red_blocks(size_t N) {
  size_t b = (N + 4 * kRedThreads - 1) / (4 * kRedThreads);  // about 4 entries per thread
  if (b > (size_t)kRedBlocks) b = kRedBlocks;
  return (unsigned)(b ? b : 1);
}
unsigned ew_blocks(size_t N) {
  size_t b = (N + 255) / 256;
  if (b > 4096) b = 4096;
  return (unsigned)(b ? b : 1);
}

template <int OP, int NV>
void finish(hipStream_t s, CgScal* sc, const Ctx& c, size_t N, int first = 0, int* flag = nullptr) {
  hipLaunchKernelGGL((k_finish<OP, NV>), dim3(1), dim3(kRedThreads), 0, s, sc, c.partials,
                     (int)red_blocks(N), first, flag);
}

int read_scal(Ctx& c, CgScal* out, int which) {
  SC_HIP(hipMemcpyAsync(c.scal_h, (CgScal*)c.scal + which, sizeof(CgScal), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(out, c.scal_h, sizeof(CgScal));
  return SC_OK;
}

// Jacobi-PCG on A x = b from the given x.  `which` selects the scalar slot.
template <typename T>
int pcg(Ctx& c, const T* b, T* x, T* r, T* p, T* q, const T* dinv, double rtol, int maxit,
        int check_every, int which, int* iters_out, double* relres_out, int* status_out) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  CgScal* sc = (CgScal*)c.scal + which;
  auto apply = [&](const T* in, T* out) -> int {
    if constexpr (sizeof(T) == 8) return launch_apply64(c, (const double*)in, (double*)out, s);
    else return launch_apply32(c, (const float*)in, (float*)out, s);
  };
  hipLaunchKernelGGL(k_set_scal<T>, dim3(1), dim3(1), 0, s, sc, rtol * rtol, maxit);
  int e = apply(x, q);
  if (e) return e;
  hipLaunchKernelGGL(k_cg_init<T>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, b, q, dinv, r,
                     p, c.partials);
  finish<FIN_INIT, 3>(s, sc, c, N);
  SC_HIP(hipGetLastError());
  CgScal h;
  int batch = 4;
  int launched = 0;
  if (check_every <= 0) check_every = 64;
  while (true) {
    for (int i = 0; i < batch; ++i) {
      e = apply(p, q);
      if (e) return e;
      hipLaunchKernelGGL(k_cg_dot_pq<T>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, p, q,
                         sc, c.partials);
      finish<FIN_PQ, 1>(s, sc, c, N);
      hipLaunchKernelGGL(k_cg_update<T>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, x, r, p,
                         q, dinv, sc, c.partials);
      finish<FIN_UPDATE, 2>(s, sc, c, N);
      hipLaunchKernelGGL(k_cg_pvec<T>, dim3(ew_blocks(N)), dim3(256), 0, s, N, r, dinv, p, sc);
    }
    SC_HIP(hipGetLastError());
    launched += batch;
    e = read_scal(c, &h, which);
    if (e) return e;
    if (h.done || launched >= maxit) break;
    batch = batch * 2 < check_every ? batch * 2 : check_every;
  }
  *iters_out = h.iters;
  *relres_out = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  *status_out = h.done;
  return SC_OK;
}

// FP64 CG preconditioned with the fast-diagonalisation inverse of the
// separable operator; `fprec` selects the GEMM precision.  The host reads the
// device scalars once per iteration (iteration counts are small).
int pcg_fdiag(Ctx& c, const double* b, double* x, int fprec, double rtol, int maxit,
              int* iters_out, double* relres_out, int* status_out) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  CgScal* sc = (CgScal*)c.scal;
  double* r = c.r64;
  double* p = c.p64;
  double* q = c.q64;
  double* z = c.z64;
  hipLaunchKernelGGL(k_set_scal<double>, dim3(1), dim3(1), 0, s, sc, rtol * rtol, maxit);
  int e = launch_apply64(c, x, q, s);
  if (e) return e;
  hipLaunchKernelGGL(k_ir_residual, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, b, q, r,
                     c.partials);
  finish<FIN_IRRES, 2>(s, sc, c, N);
  SC_HIP(hipGetLastError());
  CgScal h;
  e = read_scal(c, &h, 0);
  if (e) return e;
  int first = 1;
  while (!h.done) {
    e = fdiag_apply(c, *c.fd, fprec, r, z);
    if (e) return e;
    hipLaunchKernelGGL(k_cg_rz<double>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, r, z, c.zold64,
                       first, sc, c.partials);
    finish<FIN_RZ, 2>(s, sc, c, N, first);
    hipLaunchKernelGGL(k_cg_pz<double>, dim3(ew_blocks(N)), dim3(256), 0, s, N, z, p, sc, first);
    first = 0;
    e = launch_apply64(c, p, q, s);
    if (e) return e;
    hipLaunchKernelGGL(k_cg_dot_pq<double>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, p, q,
                       sc, c.partials);
    finish<FIN_PQ, 1>(s, sc, c, N);
    hipLaunchKernelGGL(k_cg_update_r<double>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, x, r, p, q,
                       sc, c.partials);
    finish<FIN_UPDATE_R, 1>(s, sc, c, N);
    SC_HIP(hipGetLastError());
    e = read_scal(c, &h, 0);
    if (e) return e;
  }
  *iters_out = h.iters;
  *relres_out = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  *status_out = h.done;
  return SC_OK;
}

// The inner solve of iterative refinement with the fast-diagonalisation
// preconditioner: flexible PCG in FP32 on A x = b from x = 0, the
// preconditioner's GEMMs in FP32 or bf16, dot products accumulated in FP64,
// scalars in slot 1.  Vectors: r, p, q, z, zold (FP32, N each).  The host
// reads the scalars after every `batch` iterations; the vector kernels of
// an iteration past convergence return at once.
int pcg32_fdiag(Ctx& c, const float* b, float* x, float* r, float* p, float* q, float* z, float* zold,
                int fprec, double rtol, int maxit, int batch, int* iters_out, double* relres_out,
                int* status_out) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  CgScal* sc = (CgScal*)c.scal + 1;
  hipLaunchKernelGGL(k_set_scal<float>, dim3(1), dim3(1), 0, s, sc, rtol * rtol, maxit);
  hipLaunchKernelGGL(k_cg32_start, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, b, x, r, c.partials);
  finish<FIN_IRRES, 2>(s, sc, c, N);
  SC_HIP(hipGetLastError());
  CgScal h;
  int e = read_scal(c, &h, 1);
  if (e) return e;
  int first = 1, launched = 0;
  if (batch < 1) batch = 1;
  while (!h.done && launched < maxit) {
    for (int i = 0; i < batch; ++i) {
      e = fdiag_apply_f32(c, *c.fd, fprec, r, z);
      if (e) return e;
      hipLaunchKernelGGL(k_cg_rz<float>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, r, z, zold, first,
                         sc, c.partials);
      finish<FIN_RZ, 2>(s, sc, c, N, first);
      hipLaunchKernelGGL(k_cg_pz<float>, dim3(ew_blocks(N)), dim3(256), 0, s, N, z, p, sc, first);
      first = 0;
      e = launch_apply32(c, p, q, s);
      if (e) return e;
      hipLaunchKernelGGL(k_cg_dot_pq<float>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, p, q, sc,
                         c.partials);
      finish<FIN_PQ, 1>(s, sc, c, N);
      hipLaunchKernelGGL(k_cg_update_r<float>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, x, r, p, q,
                         sc, c.partials);
      finish<FIN_UPDATE_R, 1>(s, sc, c, N);
    }
    SC_HIP(hipGetLastError());
    launched += batch;
    e = read_scal(c, &h, 1);
    if (e) return e;
  }
  *iters_out = h.iters;
  *relres_out = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  *status_out = h.done;
  return SC_OK;
}

// multidot partials of the guess sit above the partials of the guess kernel
constexpr int kMdOffset = 2 * kRedBlocks;
constexpr int kMdDoubles = 2 * kRedBlocks;
constexpr int kMdBlocksMax = 256;

template <int M>
void proj_guess_m(Ctx& c, const VecList& Y, const VecList& Z, const double* b, double* x, double tol2,
                  bool rhs_frame, int pre_mnb, const QdFuse* fuse, StepGate gate, bool lazyX) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  double* mpart = c.partials + kMdOffset;
  static_assert(kMdBlocksMax * SC_REUSE_MAX <= kMdDoubles, "multidot partials overflow");
  unsigned mnb;
  if (pre_mnb > 0) {
    mnb = (unsigned)pre_mnb;  // the caller's kernel wrote them
  } else if (rhs_frame) {
    const FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
    mnb = (unsigned)std::min<long long>((fm.count() + kRedThreads - 1) / kRedThreads, kMdBlocksMax);
    hipLaunchKernelGGL(k_multidot_frame<M>, dim3(mnb), dim3(kRedThreads), 0, s, fm, c.d.Nz, b, Y, mpart);
  } else {
    mnb = std::min<unsigned>(red_blocks(N), kMdBlocksMax);
    hipLaunchKernelGGL(k_multidot<M>, dim3(mnb), dim3(kRedThreads), 0, s, N, b, (const double*)nullptr, Y,
                       mpart);
  }
  // the projections came from the QD driver's rhs kernel: b vanishes off
  // its band rows
  size_t zLo = 0, zHi = 0;
  if (pre_mnb > 0) {
    int nLow, nHigh;
    qd_band_rows(c.d.Ny, &nLow, &nHigh);
    zLo = (size_t)nLow * c.d.Nz;
    zHi = (size_t)(c.d.Ny - nHigh) * c.d.Nz;
  }
  double* coefOut = c.projW + 2 * SC_REUSE_MAX;
  if (fuse) {
    GuessFault gf{*fuse, {}, (fuse->a.N + kFaultWave - 1) / kFaultWave};
    for (int m = 0; m < M; ++m) gf.T.v[m] = c.projT[m];
    if (lazyX)
      hipLaunchKernelGGL((k_proj_guess_res<M, true, false>), dim3(red_blocks(N) + gf.nfb), dim3(kRedThreads), 0, s,
                         N, Y, Z, (const double*)mpart, (int)mnb, (const double*)c.projW, (CgScal*)c.scal, tol2, b,
                         x, c.partials, gf, gate, zLo, zHi, coefOut);
    else
      hipLaunchKernelGGL((k_proj_guess_res<M, true, true>), dim3(red_blocks(N) + gf.nfb), dim3(kRedThreads), 0, s,
                         N, Y, Z, (const double*)mpart, (int)mnb, (const double*)c.projW, (CgScal*)c.scal, tol2, b,
                         x, c.partials, gf, gate, zLo, zHi, coefOut);
  } else {
    hipLaunchKernelGGL((k_proj_guess_res<M, false, true>), dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, Y, Z,
                       (const double*)mpart, (int)mnb, (const double*)c.projW, (CgScal*)c.scal, tol2, b, x,
                       c.partials, GuessFault{}, gate, zLo, zHi, coefOut);
  }
}

template <int M>
void proj_add_m(Ctx& c, const VecList& Y, const VecList& Z, const double* x, const double* ax, double* y,
                double* z) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  const unsigned nb = red_blocks(N);
  hipLaunchKernelGGL(k_multidot<M>, dim3(nb), dim3(kRedThreads), 0, s, N, ax, (const double*)nullptr, Y,
                     c.partials);
  hipLaunchKernelGGL(k_proj_coef<M>, dim3(1), dim3(kRedThreads), 0, s, c.partials, (int)nb, c.projW);
  hipLaunchKernelGGL(k_proj_update_az<M>, dim3(nb), dim3(kRedThreads), 0, s, N, Y, Z, c.projW, x, ax, y, z,
                     c.partials);
}

// guess: x (p1) = projection of the solution of A x = b (p0) and a residual
// check of tolerance tol2; add: new basis pair from x (p0) and A x (p1) into
// y (p2), z (p3)
struct ProjArgs {
  const double* p0;
  double *p1, *p2, *p3;
  double tol2;
  bool rhs_frame;
  int pre_mnb;
  const QdFuse* fuse;
  StepGate gate;
  bool lazyX;
};

template <int M>
void proj_dispatch(bool guess, Ctx& c, const VecList& Y, const VecList& Z, const ProjArgs& a) {
  if (guess) proj_guess_m<M>(c, Y, Z, a.p0, a.p1, a.tol2, a.rhs_frame, a.pre_mnb, a.fuse, a.gate, a.lazyX);
  else proj_add_m<M>(c, Y, Z, a.p0, a.p1, a.p2, a.p3);
}

void proj_run(bool guess, int M, Ctx& c, const VecList& Y, const VecList& Z, const ProjArgs& a) {
  switch (M) {
    case 1: proj_dispatch<1>(guess, c, Y, Z, a); break;
    case 2: proj_dispatch<2>(guess, c, Y, Z, a); break;
    case 3: proj_dispatch<3>(guess, c, Y, Z, a); break;
    case 4: proj_dispatch<4>(guess, c, Y, Z, a); break;
    case 5: proj_dispatch<5>(guess, c, Y, Z, a); break;
    case 6: proj_dispatch<6>(guess, c, Y, Z, a); break;
    case 7: proj_dispatch<7>(guess, c, Y, Z, a); break;
    default: proj_dispatch<8>(guess, c, Y, Z, a); break;
  }
}

int proj_prepare(Ctx& c, int cap) {
  if (c.projCap == cap) return SC_OK;
  for (int i = 0; i < SC_REUSE_MAX; ++i) {
    for (double** v : {&c.projY[i], &c.projZ[i], &c.projU[i], &c.projZU[i], &c.projT[i]}) {
      if (*v) SC_HIP(hipFree(*v));
      *v = nullptr;
    }
  }
  for (int i = 0; i < cap; ++i) {
    for (double** v : {&c.projY[i], &c.projZ[i], &c.projU[i], &c.projZU[i]})
      SC_HIP(hipMalloc(v, (size_t)c.N * sizeof(double)));
    SC_HIP(hipMalloc(&c.projT[i], (size_t)c.d.Nz * sizeof(double)));
  }
  // w, the basis update's coefficients, the last lazy guess's coefficients
  if (!c.projW) SC_HIP(hipMalloc(&c.projW, 3 * SC_REUSE_MAX * sizeof(double)));
  c.projCap = cap;
  c.projCount = c.projNext = 0;
  return SC_OK;
}

VecList proj_list(const Ctx& c, bool az) {
  VecList Y{};
  for (int i = 0; i < c.projCount; ++i) Y.v[i] = az ? c.projZ[i] : c.projY[i];
  return Y;
}

int check_opts(const sc_solve_opts& o) {
  if (o.mode != SC_SOLVE_CG64 && o.mode != SC_SOLVE_IR32) return fail(SC_ERR_ARG, "unknown solve mode");
  if (!(o.rtol > 0)) return fail(SC_ERR_ARG, "rtol must be > 0");
  if (o.precond < SC_PREC_JACOBI || o.precond > SC_PREC_FDIAG_BF16)
    return fail(SC_ERR_ARG, "unknown preconditioner");
  if (o.reuse < 0 || o.reuse > SC_REUSE_MAX) return fail(SC_ERR_ARG, "reuse must be in [0, SC_REUSE_MAX]");
  return SC_OK;
}

}  // namespace

int solve_core(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, sc_solve_stats& st);

// Basis vector k from the raw pair (x, A x), A-orthogonal to basis vectors
// 0..k-1, and its weight.
int proj_basis_vector(Ctx& c, int k, const double* x, const double* ax) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  const int nb = (int)red_blocks(N);
  if (k == 0) {
    SC_HIP(hipMemcpyAsync(c.projY[0], x, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    SC_HIP(hipMemcpyAsync(c.projZ[0], ax, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_multidot<1>, dim3(nb), dim3(kRedThreads), 0, s, N, ax, (const double*)nullptr,
                       VecList{{x}}, c.partials);
    hipLaunchKernelGGL(k_proj_weight<1>, dim3(1), dim3(kRedThreads), 0, s, c.partials, nb, c.projW, 0);
  } else {
    VecList Y{}, Z{};
    for (int i = 0; i < k; ++i) {
      Y.v[i] = c.projY[i];
      Z.v[i] = c.projZ[i];
    }
    proj_run(false, k, c, Y, Z,
             ProjArgs{x, const_cast<double*>(ax), c.projY[k], c.projZ[k], 0.0, false, 0, nullptr, StepGate{}, false});
    hipLaunchKernelGGL(k_proj_weight<2>, dim3(1), dim3(kRedThreads), 0, s, c.partials, nb, c.projW, k);
  }
  hipLaunchKernelGGL(k_basis_traction, dim3((c.d.Nz + 255) / 256), dim3(256), 0, s, c.op64,
                     (const double*)c.projY[k], c.projT[k]);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// Put u (and A u) into the ring of recent solutions.  While the ring fills,
// u becomes the next basis vector; once it is full, u replaces the oldest
// solution and the basis is rebuilt from the ring oldest first, so that its
// span stays that of the last projCap solutions (replacing one basis vector
// in place would drop the evicted solution's components from the others).
int proj_add_solution(Ctx& c, const double* u) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  const int slot = c.projNext;
  SC_HIP(hipMemcpyAsync(c.projU[slot], u, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  int e = launch_apply64(c, u, c.projZU[slot], s);
  if (e) return e;
  c.projNext = (slot + 1) % c.projCap;
  if (c.projCount < c.projCap) {
    e = proj_basis_vector(c, c.projCount, c.projU[slot], c.projZU[slot]);
    ++c.projCount;
    return e;
  }
  for (int k = 0; k < c.projCap && !e; ++k) {
    const int r = (c.projNext + k) % c.projCap;  // oldest first
    e = proj_basis_vector(c, k, c.projU[r], c.projZU[r]);
  }
  return e;
}

// With o.reuse > 0 the solve starts from the A-orthogonal projection of the
// new solution onto up to o.reuse earlier ones (Fischer's projection), with
// A y kept beside each basis vector y so that the guess's residual costs no
// operator apply.  A guess that already meets the tolerance is the answer.
// With miss_flag set and a basis present the check is deferred: u gets the
// guess, *miss_flag is set to 1 on the device if it misses the tolerance, and
// *deferred tells the caller, who must test the flag before trusting u.
// Timing of the projected-guess kernel alone on the current basis (rhs and
// multidot partials as the last solve left them): mean ms per launch over
// reps launches, HIP events on the context stream; *words = 8-byte words the
// kernel moves per launch (b, M basis vectors, M products, x).
int proj_materialize(Ctx& c, double* u) {
  if (c.projCount == 0) return fail(SC_ERR_ARG, "proj_materialize: no basis");
  const size_t N = (size_t)c.N;
  const VecList Y = proj_list(c, false);
  const double* coef = c.projW + 2 * SC_REUSE_MAX;
  const dim3 g(ew_blocks(N)), b(256);
  switch (c.projCount) {
#define SC_PX(m) \
  case m: hipLaunchKernelGGL(k_proj_x<m>, g, b, 0, c.stream, N, Y, coef, u); break;
    SC_PX(1) SC_PX(2) SC_PX(3) SC_PX(4) SC_PX(5) SC_PX(6) SC_PX(7) SC_PX(8)
#undef SC_PX
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// The QD step's guess (x left to proj_materialize, b read on the rhs band
// only): M basis products and the band of b per launch
int bench_proj_guess(Ctx& c, int reps, double* ms, double* words) {
  if (c.projCount == 0) return fail(SC_ERR_ARG, "no projection basis: run a solve with reuse > 0 first");
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  int nLow, nHigh;
  qd_band_rows(c.d.Ny, &nLow, &nHigh);
  const size_t zLo = (size_t)nLow * c.d.Nz, zHi = (size_t)(c.d.Ny - nHigh) * c.d.Nz;
  const FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
  const int mnb = (int)std::min<long long>((fm.count() + kRedThreads - 1) / kRedThreads, kMdBlocksMax);
  const int M = c.projCount;
  const VecList Y = proj_list(c, false), Z = proj_list(c, true);
  auto launch = [&]() {
    const dim3 g(red_blocks(N)), b(kRedThreads);
    double* mp = c.partials + kMdOffset;
    CgScal* sc = (CgScal*)c.scal;
    switch (M) {
#define SC_GUESS(m) \
  case m: hipLaunchKernelGGL((k_proj_guess_res<m, false, false>), g, b, 0, s, N, Y, Z, (const double*)mp, mnb, \
                             (const double*)c.projW, sc, 0.0, (const double*)c.r64, c.p64, c.partials, \
                             GuessFault{}, StepGate{}, zLo, zHi, c.projW + 2 * SC_REUSE_MAX); break;
      SC_GUESS(1) SC_GUESS(2) SC_GUESS(3) SC_GUESS(4) SC_GUESS(5) SC_GUESS(6) SC_GUESS(7) SC_GUESS(8)
#undef SC_GUESS
    }
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  SC_HIP(hipEventCreate(&e0));
  SC_HIP(hipEventCreate(&e1));
  SC_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) launch();
  SC_HIP(hipEventRecord(e1, s));
  SC_HIP(hipEventSynchronize(e1));
  float t = 0;
  SC_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SC_HIP(hipGetLastError());
  *ms = t / reps;
  *words = (double)N * M + (double)(nLow + nHigh) * c.d.Nz;
  return SC_OK;
}

int proj_frame_plan(Ctx& c, const sc_solve_opts& o, ProjFramePlan* plan) {
  plan->M = 0;
  int e = check_opts(o);
  if (e || o.reuse == 0) return e;
  e = ensure_workspace(c, o.mode == SC_SOLVE_IR32);
  if (e) return e;
  e = proj_prepare(c, o.reuse);
  if (e || c.projCount == 0) return e;
  plan->M = c.projCount;
  plan->Y = proj_list(c, false);
  plan->mpart = c.partials + kMdOffset;
  plan->maxBlocks = kMdBlocksMax;
  return SC_OK;
}

int solve(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, sc_solve_stats& st,
          int* miss_flag, bool* deferred, bool rhs_frame, int pre_mnb, const QdFuse* fuse, bool* fused,
          PendingFinish* postpone, StepGate gate, bool* x_lazy) {
  if (postpone) *postpone = PendingFinish{};
  if (x_lazy) *x_lazy = false;
  if (deferred) *deferred = false;
  if (fused) *fused = false;
  int e = check_opts(o);
  if (e) return e;
  if (o.reuse == 0) return solve_core(c, rhs, u, o, st);
  e = ensure_workspace(c, o.mode == SC_SOLVE_IR32);
  if (e) return e;
  e = proj_prepare(c, o.reuse);
  if (e) return e;
  hipStream_t s = c.stream;
  const size_t N = (size_t)c.N;
  if (c.projCount > 0) {
    CgScal* sc = (CgScal*)c.scal;
    const QdFuse* f = miss_flag ? fuse : nullptr;  // only a deferred guess is taken as the answer
    // a deferred, fused guess may leave u to proj_materialize
    const bool lazy = f && x_lazy;
    proj_run(true, c.projCount, c, proj_list(c, false), proj_list(c, true),
             ProjArgs{rhs, u, nullptr, nullptr, o.rtol * o.rtol, rhs_frame, pre_mnb, f, miss_flag ? gate : StepGate{},
                      lazy});
    if (lazy) *x_lazy = true;
    if (f && fused) *fused = true;
    if (miss_flag && postpone)
      *postpone = PendingFinish{sc, c.partials, (int)red_blocks(N), miss_flag};
    else
      finish<FIN_IRRES, 2>(s, sc, c, N, 0, miss_flag);
    SC_HIP(hipGetLastError());
    if (miss_flag) {
      std::memset(&st, 0, sizeof(st));
      st.converged = 1;
      st.relres = -1.0;  // not read back
      if (deferred) *deferred = true;
      return SC_OK;
    }
    CgScal h;
    e = read_scal(c, &h, 0);
    if (e) return e;
    if (h.done) {
      std::memset(&st, 0, sizeof(st));
      st.converged = 1;
      st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
      // the answer lies in the span of the basis, which therefore stays as it is
      return SC_OK;
    }
  }
  e = solve_core(c, rhs, u, o, st);
  if (e) return e;
  return proj_add_solution(c, u);
}

// FP64 iterative refinement whose sweeps solve the correction equation with
// the FP32 fast-diagonalisation PCG (pcg32_fdiag): the residual b - A u is
// formed in FP64 with the FP64 apply, scaled to unit norm and rounded to
// FP32; the FP32 correction is added in FP64.  FP32 vectors of the inner
// solve: r32 (its rhs), d32 (its solution), q32, p32 and three more in the
// FP64 buffers p64, q64 (free while the inner solve runs).
int solve_ir_fdiag(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, int fprec,
                   sc_solve_stats& st) {
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  CgScal* sc0 = (CgScal*)c.scal;
  const double inner = o.inner_rtol > 0 ? o.inner_rtol : 1e-5;
  const int maxOuter = o.max_outer > 0 ? o.max_outer : 30;
  const int maxit = o.max_iters > 0 ? o.max_iters : 10000;
  float* zin = (float*)c.p64;
  float* zold = (float*)c.p64 + N;
  float* qin = (float*)c.q64;
  int total = 0;
  for (int outer = 0;; ++outer) {
    hipLaunchKernelGGL(k_set_scal<double>, dim3(1), dim3(1), 0, s, sc0, o.rtol * o.rtol, maxit);
    int e = launch_apply64(c, u, c.q64, s);
    if (e) return e;
    hipLaunchKernelGGL(k_ir_residual, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, rhs, c.q64,
                       c.r64, c.partials);
    finish<FIN_IRRES, 2>(s, sc0, c, N);
    SC_HIP(hipGetLastError());
    CgScal h;
    e = read_scal(c, &h, 0);
    if (e) return e;
    st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
    st.outer = outer;
    if (h.done) {
      st.converged = 1;
      break;
    }
    if (outer >= maxOuter || total >= maxit) {
      st.iters = total;
      return fail(SC_ERR_NOCONV, "IR32 (fast-diagonalisation) did not converge (relres " +
                                     std::to_string(st.relres) + ")");
    }
    hipLaunchKernelGGL(k_ir_to32, dim3(ew_blocks(N)), dim3(256), 0, s, N, c.r64, sc0, c.r32, c.d32);
    int it = 0, status = 0;
    double rr = 0;
    e = pcg32_fdiag(c, c.r32, c.d32, c.q32, c.p32, qin, zin, zold, fprec, inner, maxit - total, 2, &it, &rr,
                    &status);
    if (e) return e;
    total += it;
    if (status == 2 || status == 4) return fail(SC_ERR_NONFINITE, "IR32 inner CG breakdown");
    hipLaunchKernelGGL(k_ir_correct, dim3(ew_blocks(N)), dim3(256), 0, s, N, c.d32, sc0, u);
    SC_HIP(hipGetLastError());
  }
  st.iters = total;
  return SC_OK;
}

// the fast-diagonalisation factors for the operator as it stands (built on
// first use and again after the BC types or the coefficients change)
int ensure_fdiag(Ctx& c) {
  const int key = c.d.bcR | c.d.bcT << 1 | c.d.bcL << 2 | c.d.bcB << 3;
  if (!c.fd) c.fd = new FdDiag;
  if (c.fd->bcKey != key || c.fd_stale) {
    const int e = fdiag_build(c, *c.fd);
    if (e) return e;
    c.fd_stale = false;
  }
  return SC_OK;
}

__global__ void k_jacobi_apply(size_t n, const double* __restrict__ r, const double* __restrict__ dinv,
                               double* __restrict__ z) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    z[i] = r[i] * dinv[i];
}

// z = M^-1 r with one of the solve's preconditioners
int precond_apply(Ctx& c, int precond, const double* r, double* z) {
  if (precond == SC_PREC_JACOBI) {
    if (!c.dinv64) return fail(SC_ERR_ARG, "sc_precond_apply: the context has no diagonal");
    const size_t N = (size_t)c.N;
    hipLaunchKernelGGL(k_jacobi_apply, dim3(ew_blocks(N)), dim3(256), 0, c.stream, N, r, c.dinv64, z);
    SC_HIP(hipGetLastError());
    return SC_OK;
  }
  if (precond != SC_PREC_FDIAG64 && precond != SC_PREC_FDIAG32 && precond != SC_PREC_FDIAG_BF16)
    return fail(SC_ERR_ARG, "sc_precond_apply: unknown preconditioner");
  const int e = ensure_fdiag(c);
  if (e) return e;
  const int fprec = precond == SC_PREC_FDIAG64 ? FD_F64 : (precond == SC_PREC_FDIAG32 ? FD_F32 : FD_BF16);
  return fdiag_apply(c, *c.fd, fprec, r, z);
}

int solve_core(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, sc_solve_stats& st) {
  std::memset(&st, 0, sizeof(st));
  const bool ir = o.mode == SC_SOLVE_IR32;
  int e = ensure_workspace(c, ir);
  if (e) return e;
  if (o.precond != SC_PREC_JACOBI) {
    if (ir && o.precond == SC_PREC_FDIAG64)
      return fail(SC_ERR_ARG, "SC_SOLVE_IR32 takes the FP32 or bf16 fast-diagonalisation preconditioner");
    e = ensure_fdiag(c);
    if (e) return e;
    if (!c.z64) SC_HIP(hipMalloc(&c.z64, (size_t)c.N * sizeof(double)));
    if (!c.zold64) SC_HIP(hipMalloc(&c.zold64, (size_t)c.N * sizeof(double)));
    const int fprec = o.precond == SC_PREC_FDIAG64 ? FD_F64 : (o.precond == SC_PREC_FDIAG32 ? FD_F32 : FD_BF16);
    if (ir) return solve_ir_fdiag(c, rhs, u, o, fprec, st);
    int it = 0, status = 0;
    double rr = 0;
    e = pcg_fdiag(c, rhs, u, fprec, o.rtol, o.max_iters > 0 ? o.max_iters : 10000, &it, &rr, &status);
    if (e) return e;
    st.iters = it;
    st.relres = rr;
    st.converged = status == 1;
    if (!st.converged)
      return fail(SC_ERR_NOCONV, "FD-preconditioned CG did not converge (status " +
                                     std::to_string(status) + ", relres " + std::to_string(rr) + ")");
    return SC_OK;
  }
  const int maxit = o.max_iters > 0 ? o.max_iters : 100000;
  const size_t N = (size_t)c.N;
  hipStream_t s = c.stream;
  if (!ir) {
    int it = 0, status = 0;
    double rr = 0;
    e = pcg<double>(c, rhs, u, c.r64, c.p64, c.q64, c.dinv64, o.rtol, maxit, o.check_every, 0, &it,
                    &rr, &status);
    if (e) return e;
    st.iters = it;
    st.outer = 0;
    st.relres = rr;
    st.converged = status == 1;
    if (!st.converged)
      return fail(SC_ERR_NOCONV, "CG64 did not converge (status " + std::to_string(status) +
                                     ", relres " + std::to_string(rr) + ")");
    return SC_OK;
  }
  // FP64 iterative refinement with FP32 inner PCG
  CgScal* sc0 = (CgScal*)c.scal;
  const double inner = o.inner_rtol > 0 ? o.inner_rtol : 1e-5;
  const int maxOuter = o.max_outer > 0 ? o.max_outer : 30;
  int total = 0;
  for (int outer = 0;; ++outer) {
    hipLaunchKernelGGL(k_set_scal<double>, dim3(1), dim3(1), 0, s, sc0, o.rtol * o.rtol, maxit);
    e = launch_apply64(c, u, c.q64, s);
    if (e) return e;
    hipLaunchKernelGGL(k_ir_residual, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, rhs, c.q64,
                       c.r64, c.partials);
    finish<FIN_IRRES, 2>(s, sc0, c, N);
    SC_HIP(hipGetLastError());
    CgScal h;
    e = read_scal(c, &h, 0);
    if (e) return e;
    st.relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
    st.outer = outer;
    if (h.done) {
      st.converged = 1;
      break;
    }
    if (outer >= maxOuter || total >= maxit) {
      st.iters = total;
      return fail(SC_ERR_NOCONV, "IR32 did not converge (relres " + std::to_string(st.relres) + ")");
    }
    hipLaunchKernelGGL(k_ir_to32, dim3(ew_blocks(N)), dim3(256), 0, s, N, c.r64, sc0, c.r32, c.d32);
    // inner FP32 solve of A d = r/||r|| from d = 0 (slot 1 scalars)
    int it = 0, status = 0;
    double rr = 0;
    e = pcg<float>(c, c.r32, c.d32, c.q32 /*r*/, c.p32, (float*)c.p64 /*q (reuses FP64 buffer)*/,
                   c.dinv32, inner, maxit - total, o.check_every, 1, &it, &rr, &status);
    if (e) return e;
    total += it;
    if (status == 2 || status == 4) return fail(SC_ERR_NONFINITE, "IR32 inner CG breakdown");
    hipLaunchKernelGGL(k_ir_correct, dim3(ew_blocks(N)), dim3(256), 0, s, N, c.d32, sc0, u);
    SC_HIP(hipGetLastError());
  }
  st.iters = total;
  return SC_OK;
}

}  // namespace sc

extern "C" int sc_solve(sc_ctx* h, const double* rhs_d, double* u_d, const sc_solve_opts* opts,
                        sc_solve_stats* stats) {
  if (!h || !rhs_d || !u_d || !opts) return sc::fail(SC_ERR_ARG, "sc_solve: null argument");
  sc_solve_stats st;
  int e = sc::solve(h->c, rhs_d, u_d, *opts, st);
  if (stats) *stats = st;
  return e;
}

extern "C" int sc_precond_apply(sc_ctx* h, int precond, const double* r_d, double* z_d) {
  if (!h || !r_d || !z_d) return sc::fail(SC_ERR_ARG, "sc_precond_apply: null argument");
  return sc::precond_apply(h->c, precond, r_d, z_d);
}

// ------------------------------------------------ allocation-free solve --
// The scratch of sc_solve_noalloc lives in one file-scope device array: the
// CG vectors r, p, q for grids of up to kNoAllocMaxNodes nodes, then the
// reduction partials, then the two scalar blocks -- 25.2 MB that every GPU
// the library runs on holds until the process ends.  In host code `scratch`
// only names the symbol: its address on a device comes from
// hipGetSymbolAddress, resolved once per device with that device current,
// and that address is what the kernels receive.
namespace sc {
namespace {
constexpr size_t kNoAllocMaxNodes = (size_t)1 << 20;
constexpr size_t kNoAllocPartials = (size_t)kRedBlocks * 4;
constexpr size_t kNoAllocScal = 32;  // doubles reserved for two CgScal blocks
static_assert(2 * sizeof(CgScal) <= kNoAllocScal * sizeof(double), "scalar blocks exceed their slot");
constexpr size_t kNoAllocDoubles = 3 * kNoAllocMaxNodes + kNoAllocPartials + kNoAllocScal;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocDevices] = {};

// the scratch's address on the device of `stream`, resolved on first use
int noalloc_scratch(hipStream_t stream, double** out) {
  int dev = 0, cur = 0;
  SC_HIP(hipStreamGetDevice(stream, &dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(SC_ERR_ARG, "sc_solve_noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    SC_HIP(hipGetDevice(&cur));
    if (cur != dev) SC_HIP(hipSetDevice(dev));
    double* p = nullptr;
    const hipError_t he = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (cur != dev) SC_HIP(hipSetDevice(cur));
    if (he != hipSuccess || !p)
      return fail(SC_ERR_HIP, std::string("sc_solve_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(he));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SC_OK;
}

// zeroes the scalar blocks (a fill kernel, not a memset, so that the solve
// holds no memset node if a caller records it)
__global__ void k_zero_doubles(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace
}  // namespace sc

extern "C" int sc_solve_noalloc(sc_ctx* h, const double* rhs_d, double* u_d, const sc_solve_opts* opts,
                                sc_solve_stats* stats) {
  using namespace sc;
  if (!h || !rhs_d || !u_d || !opts) return fail(SC_ERR_ARG, "sc_solve_noalloc: null argument");
  const sc_solve_opts& o = *opts;
  if (o.mode != SC_SOLVE_CG64 || o.precond != SC_PREC_JACOBI || o.reuse != 0)
    return fail(SC_ERR_ARG, "sc_solve_noalloc: only FP64 Jacobi-PCG without reuse (SC_SOLVE_CG64, "
                            "SC_PREC_JACOBI, reuse 0) runs without allocating");
  if (!(o.rtol > 0)) return fail(SC_ERR_ARG, "sc_solve_noalloc: rtol must be > 0");
  Ctx& c = h->c;
  const size_t N = (size_t)c.N;
  if (N > kNoAllocMaxNodes)
    return fail(SC_ERR_ARG, "sc_solve_noalloc: the grid has " + std::to_string(N) +
                                " nodes, more than the allocation-free scratch holds (" +
                                std::to_string(kNoAllocMaxNodes) + ")");
  double* base = nullptr;
  int e = noalloc_scratch(c.stream, &base);
  if (e) return e;
  // the context's kernels run on the scratch for this call; its own
  // workspace (if any) is put back afterwards
  struct Swap {
    Ctx& c;
    double *r, *p, *q, *part;
    void *scal, *scal_h;
    ~Swap() {
      c.r64 = r;
      c.p64 = p;
      c.q64 = q;
      c.partials = part;
      c.scal = scal;
      c.scal_h = scal_h;
    }
  } keep{c, c.r64, c.p64, c.q64, c.partials, c.scal, c.scal_h};
  CgScal hostScal[2];
  c.r64 = base;
  c.p64 = base + kNoAllocMaxNodes;
  c.q64 = base + 2 * kNoAllocMaxNodes;
  c.partials = base + 3 * kNoAllocMaxNodes;
  c.scal = base + 3 * kNoAllocMaxNodes + kNoAllocPartials;
  c.scal_h = hostScal;
  hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, c.stream, (double*)c.scal, (int)kNoAllocScal);
  SC_HIP(hipGetLastError());
  int it = 0, status = 0;
  double rr = 0;
  e = pcg<double>(c, rhs_d, u_d, c.r64, c.p64, c.q64, c.dinv64, o.rtol, o.max_iters > 0 ? o.max_iters : 100000,
                  o.check_every, 0, &it, &rr, &status);
  if (e) return e;
  sc_solve_stats st{};
  st.iters = it;
  st.relres = rr;
  st.converged = status == 1;
  if (stats) *stats = st;
  if (!st.converged)
    return fail(SC_ERR_NOCONV, "sc_solve_noalloc: CG64 did not converge (status " + std::to_string(status) +
                                   ", relres " + std::to_string(rr) + ")");
  return SC_OK;
}
