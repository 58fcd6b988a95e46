// Block reductions shared by the solver and the fused driver kernels.
#pragma once
#include "sc_internal.hpp"

namespace sc {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 2048;  // grid-stride reduction kernels (max blocks)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-reduce NV values and write this block's partial sums.  The partials
// are combined by a kernel launched after this one (k_finish, or the guess
// kernel of the projection): the kernel boundary orders the two, so no
// device-scope fence is needed.
template <int NV>
__device__ __forceinline__ void block_partials_at(double (&v)[NV], double* __restrict__ partials, int blk) {
  __shared__ double sh[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) sh[i][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kRedThreads / 64; ++w) s += sh[threadIdx.x][w];
    partials[(size_t)blk * NV + threadIdx.x] = s;
  }
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partials) {
  block_partials_at(v, partials, blockIdx.x);
}

}  // namespace sc
