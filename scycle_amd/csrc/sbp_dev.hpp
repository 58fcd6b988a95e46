// Device helpers shared by the SBP-SAT kernels and the fused kernels of the
// quasi-dynamic driver.
#pragma once
#include "sc_internal.hpp"

namespace sc {

// Linear index t over the nodes within kFrame of a boundary -> (j, k): the
// top kFrame rows, the bottom kFrame rows, then the left/right kFrame columns
// of the rows in between.
struct FrameMap {
  int Ny, Nz, nTop, nBot, nL, nR, nMid;
  __host__ __device__ static FrameMap make(int Ny, int Nz) {
    FrameMap f;
    f.Ny = Ny;
    f.Nz = Nz;
    f.nTop = Ny < kFrame ? Ny : kFrame;
    f.nBot = Ny > kFrame ? (Ny - kFrame < kFrame ? Ny - kFrame : kFrame) : 0;
    f.nMid = Ny - f.nTop - f.nBot;
    f.nL = Nz < kFrame ? Nz : kFrame;
    f.nR = Nz > kFrame ? (Nz - kFrame < kFrame ? Nz - kFrame : kFrame) : 0;
    return f;
  }
  __host__ __device__ long long count() const {
    return (long long)(nTop + nBot) * Nz + (long long)nMid * (nL + nR);
  }
  // 32-bit arithmetic: a frame holds fewer than 2^31 nodes for any grid the
  // kernels take (kFrame * 4 * side)
  __device__ void node(long long t64, int& j, int& k) const {
    unsigned t = (unsigned)t64;
    const unsigned nz = (unsigned)Nz;
    const unsigned a = (unsigned)nTop * nz;
    if (t < a) { j = (int)(t / nz); k = (int)(t - (unsigned)j * nz); return; }
    t -= a;
    const unsigned b = (unsigned)nBot * nz;
    if (t < b) { const unsigned r = t / nz; j = Ny - nBot + (int)r; k = (int)(t - r * nz); return; }
    t -= b;
    const unsigned w = (unsigned)(nL + nR);
    const unsigned r = t / w;
    j = nTop + (int)r;
    const int c = (int)(t - r * w);
    k = c < nL ? c : Nz - nR + (c - nL);
  }
};

// D1 row i of a direction applied along that direction.
template <typename T, class F>
__device__ __forceinline__ T d1_row(const DevDir<T>& d, int i, const F& f) {
  T s = 0;
  for (int q = 0; q < kBand; ++q) {
    const T w = d.D1[i * kBand + q];
    if (w != T(0)) s += w * f(i + q - 4);
  }
  return s;
}

// Dy u at fault node (0, k), scaled by the metric: D1 row 0 applied along y
// with unconditional loads (the entries of row 0 sit at q >= 4; zero weights
// add nothing), so that the nine loads issue together.
__device__ __forceinline__ double dy_row0(const DevOp<double>& p, const double* __restrict__ u, int k) {
  const int Nz = p.Nz;
  double s = 0.0;
#pragma unroll
  for (int q = 4; q < kBand; ++q) {
    const int m = min(q - 4, p.Ny - 1);
    s += p.y.D1[q] * u[(size_t)m * Nz + k];
  }
  return p.y.imet[0] * s;
}

// Right-hand side of A u = b at (j, k) from the boundary data
// (SbpOps::setRhs, sbpOps_m_varGrid.cpp:1046-1083).  Every nonzero lies
// within kFrame of a boundary (the BS rows are checked to vanish beyond it at
// set-up).  bc(side, i) gives the datum of side 0 L, 1 R (i = k), 2 T, 3 B
// (i = j); it is called only where the datum enters.
template <class BC>
__device__ __forceinline__ double rhs_node(const DevOp<double>& p, int j, int k, const BC& bc) {
  const int Ny = p.Ny, Nz = p.Nz;
  double v = 0;
  if (p.doY) {
    const DevDir<double>& d = p.y;
    const double pre = p.z.H[k] * p.z.met[k];
    if (j < kBand) {
      if (p.bcL == SC_DIRICHLET) {
        const double bs = d.BS0[j] + (j == 0 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[k] * d.imet[0])) * bc(0, k);
      } else if (j == 0) {
        v += pre * bc(0, k);
      }
    }
    if (j >= Ny - kBand) {
      if (p.bcR == SC_DIRICHLET) {
        const double bs = d.BSN[j - (Ny - kBand)] + (j == Ny - 1 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[(size_t)(Ny - 1) * Nz + k] * d.imet[Ny - 1])) * bc(1, k);
      } else if (j == Ny - 1) {
        v -= pre * bc(1, k);
      }
    }
  }
  if (p.doZ) {
    const DevDir<double>& d = p.z;
    const double pre = p.y.H[j] * p.y.met[j];
    const size_t row = (size_t)j * Nz;
    if (k < kBand) {
      if (p.bcT == SC_DIRICHLET) {
        const double bs = d.BS0[k] + (k == 0 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[row] * d.imet[0])) * bc(2, j);
      } else if (k == 0) {
        v += pre * bc(2, j);
      }
    }
    if (k >= Nz - kBand) {
      if (p.bcB == SC_DIRICHLET) {
        const double bs = d.BSN[k - (Nz - kBand)] + (k == Nz - 1 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[row + Nz - 1] * d.imet[Nz - 1])) * bc(3, j);
      } else if (k == Nz - 1) {
        v -= pre * bc(3, j);
      }
    }
  }
  return v;
}

}  // namespace sc
