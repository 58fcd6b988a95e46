// Double-double arithmetic (a value as an unevaluated sum hi + lo of two
// doubles, about 104 bits): Knuth's two-sum and the FMA two-product, with the
// renormalisation of Dekker's quick-two-sum.  Used where a short sum of
// large terms must keep a small difference: the guess residual of the QD
// stage kernel, ||b - sum c_m A y_m||^2 = ||b||^2 - 2 c.p + c^T G c, which
// the residual test compares with 1e-24 ||b||^2.
#pragma once
#include <hip/hip_runtime.h>

namespace sc {

struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return DD{s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ DD dd_quick(double s, double e) {
  const double h = s + e;
  return DD{h, e - (h - s)};
}

__device__ __forceinline__ DD dd_add(DD x, DD y) {
  const DD s = dd_two_sum(x.hi, y.hi);
  const DD t = dd_two_sum(x.lo, y.lo);
  const DD u = dd_quick(s.hi, s.lo + t.hi);
  return dd_quick(u.hi, u.lo + t.lo);
}

// a * b exactly
__device__ __forceinline__ DD dd_prod(double a, double b) {
  const double p = a * b;
  return DD{p, __builtin_fma(a, b, -p)};
}

__device__ __forceinline__ DD dd_mul_d(DD x, double y) {
  const DD p = dd_prod(x.hi, y);
  return dd_quick(p.hi, __builtin_fma(x.lo, y, p.lo));
}

__device__ __forceinline__ DD dd_neg(DD x) { return DD{-x.hi, -x.lo}; }

__device__ __forceinline__ DD dd_wave_sum(DD v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const DD o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
    v = dd_add(v, o);
  }
  return v;
}

}  // namespace sc
