// Linear solve A u = rhs replacing KSPSolve in LinearElastic::computeU
// (linearElastic.cpp:465-489).
//
// SC_SOLVE_CG64: Jacobi-preconditioned CG in FP64 on the matrix-free operator.
// SC_SOLVE_IR32: FP64 iterative refinement; each sweep solves the correction
//                equation with Jacobi-PCG in FP32 (operator, vectors and
//                preconditioner in FP32, dot products accumulated in FP64),
//                the residual b - A u is always formed in FP64.
//
// All scalars (alpha, beta, norms, the converged flag) stay on the device;
// every kernel of an iteration returns at once when the flag is set, so the
// host launches iterations in batches and only reads the flag between
// batches.  Reductions are two-stage and deterministic: per-block partial
// sums, then the last block to finish sums the partials in block order.
#include <cmath>
#include <cstring>

#include "sc_internal.hpp"

namespace sc {

struct CgScal {
  double rz, pq, rr, bb, alpha, beta, tol2;
  int done, iters, maxit, pad;
  unsigned int cnt;
};

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;  // grid-stride reduction kernels

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-reduce NV values, write partials, and let the last block combine them.
// Returns true in exactly one block (the last), with tot[] filled, on thread 0.
template <int NV>
__device__ bool block_reduce_last(double (&v)[NV], double* partials, unsigned int* cnt,
                                  double (&tot)[NV]) {
  __shared__ double sh[NV][kRedThreads / 64];
  __shared__ bool amLast;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) sh[i][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0;
      for (int w = 0; w < kRedThreads / 64; ++w) s += sh[i][w];
      partials[(size_t)blockIdx.x * NV + i] = s;
    }
    __threadfence();
    unsigned int prev = atomicAdd(cnt, 1u);
    amLast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!amLast) return false;
  // last block: sum partials in block order (deterministic)
  __threadfence();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0;
      for (unsigned b = 0; b < gridDim.x; ++b) s += ((volatile double*)partials)[(size_t)b * NV + i];
      tot[i] = s;
    }
    *cnt = 0;
  }
  return true;
}

// r = b - q ; z = dinv*r ; p = z ; rz = r.z ; rr = r.r ; bb = b.b
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_cg_init(size_t N, const T* __restrict__ b,
                                                         const T* __restrict__ q,
                                                         const T* __restrict__ dinv,
                                                         T* __restrict__ r, T* __restrict__ p,
                                                         CgScal* sc, double* partials) {
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
  double tot[3];
  if (block_reduce_last<3>(v, partials, &sc->cnt, tot) && threadIdx.x == 0) {
    sc->rz = tot[0];
    sc->rr = tot[1];
    sc->bb = tot[2];
    sc->iters = 0;
    sc->done = (tot[1] <= sc->tol2 * tot[2]) ? 1 : 0;
    if (tot[2] == 0.0) sc->done = 1;
  }
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
  double tot[1];
  if (block_reduce_last<1>(v, partials, &sc->cnt, tot) && threadIdx.x == 0) {
    sc->pq = tot[0];
    sc->alpha = sc->rz / tot[0];
    if (!(tot[0] != 0.0) || !isfinite(sc->alpha)) sc->done = 2;  // breakdown
  }
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
  double tot[2];
  if (block_reduce_last<2>(v, partials, &sc->cnt, tot) && threadIdx.x == 0) {
    sc->beta = tot[0] / sc->rz;
    sc->rz = tot[0];
    sc->rr = tot[1];
    sc->iters += 1;
    if (tot[1] <= sc->tol2 * sc->bb) sc->done = 1;
    else if (sc->iters >= sc->maxit) sc->done = 3;
    else if (!isfinite(tot[1])) sc->done = 4;
  }
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
                                                             double* __restrict__ r, CgScal* sc,
                                                             double* partials) {
  double v[2] = {0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const double bi = b[i];
    const double ri = bi - q[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += bi * bi;
  }
  double tot[2];
  if (block_reduce_last<2>(v, partials, &sc->cnt, tot) && threadIdx.x == 0) {
    sc->rr = tot[0];
    sc->bb = tot[1];
    sc->done = (tot[0] <= sc->tol2 * tot[1] || tot[1] == 0.0) ? 1 : 0;
  }
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

template <typename T>
__global__ void k_set_scal(CgScal* sc, double tol2, int maxit) {
  sc->tol2 = tol2;
  sc->maxit = maxit;
  sc->done = 0;
  sc->iters = 0;
  sc->cnt = 0;
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
    SC_HIP(hipMemset(c.scal, 0, 2 * sizeof(CgScal)));
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
  size_t b = (N + kRedThreads - 1) / kRedThreads;
  if (b > (size_t)kRedBlocks) b = kRedBlocks;
  return (unsigned)(b ? b : 1);
}
unsigned ew_blocks(size_t N) {
  size_t b = (N + 255) / 256;
  if (b > 4096) b = 4096;
  return (unsigned)(b ? b : 1);
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
                     p, sc, c.partials);
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
      hipLaunchKernelGGL(k_cg_update<T>, dim3(red_blocks(N)), dim3(kRedThreads), 0, s, N, x, r, p,
                         q, dinv, sc, c.partials);
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

}  // namespace

int solve(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, sc_solve_stats& st) {
  std::memset(&st, 0, sizeof(st));
  const bool ir = o.mode == SC_SOLVE_IR32;
  if (o.mode != SC_SOLVE_CG64 && !ir) return fail(SC_ERR_ARG, "unknown solve mode");
  if (!(o.rtol > 0)) return fail(SC_ERR_ARG, "rtol must be > 0");
  int e = ensure_workspace(c, ir);
  if (e) return e;
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
                       c.r64, sc0, c.partials);
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
