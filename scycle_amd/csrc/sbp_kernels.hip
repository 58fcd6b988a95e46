// Matrix-free SBP-SAT operator kernels for gfx950.
//
// A = H*(D_yy^mu + D_zz^mu) + SAT_L + SAT_R + SAT_T + SAT_B, the operator
// SbpOps_m_varGrid / SbpOps_m_constGrid assemble as a PETSc AIJ matrix
// (sbpOps_m_varGrid.cpp:822-869, 1133-1504, 601-772), applied without ever
// forming it.  Per direction the operator is written in flux form:
//   L(u)[i] = H[i] * sum_m D1[i][m] g1[m] - sum_m D3[m][i] g3[m] - sum_m D4[m][i] g4[m]
//   g1 = c * D1 u,  g3 = C3/(18h) * c3 * D3 u,  g4 = C4/(144h) * c * D4 u   (order 4)
//   g3 = C2*h^3/4 * c * D2 u                                                (order 2)
// with c = mu*qy (y) or mu*rz (z) and c3 the two-point average the reference
// builds with global index I+1 (sbpOps_m_varGrid.cpp:1285-1311, 1428-1448).
//
// Two kernels share the work: an interior kernel (rows/columns at least
// kFrame from every boundary, where all stencils are the constant interior
// ones) and a generic table-driven kernel for the frame, which also carries
// the SAT boundary terms.
#include "sc_internal.hpp"

namespace sc {

// ------------------------------------------------------------ accessors ---
template <typename T>
struct FieldAcc {
  const T* __restrict__ u;
  int Nz;
  __device__ __forceinline__ T operator()(int j, int k) const { return u[(size_t)j * Nz + k]; }
};
template <typename T>
struct DeltaAcc {  // u = e_I, for extracting diag(A)
  int j0, k0;
  __device__ __forceinline__ T operator()(int j, int k) const { return (j == j0 && k == k0) ? T(1) : T(0); }
};

// Generic 1-D operator along y at (j, k) (no H_z*zr prefactor).
template <typename T, class U>
__device__ T line_y(const DevOp<T>& p, const U& u, int j, int k) {
  const DevDir<T>& d = p.y;
  const int Nz = p.Nz;
  const size_t N = (size_t)p.Ny * Nz;
  T acc1 = 0, accR = 0, acc4 = 0;
  for (int o = 0; o < kBand; ++o) {
    const T cf = d.D1[j * kBand + o];
    if (cf == T(0)) continue;
    const int m = j + o - 4;
    T s = 0;
    for (int q = 0; q < kBand; ++q) {
      const T w = d.D1[m * kBand + q];
      if (w != T(0)) s += w * u(m + q - 4, k);
    }
    acc1 += cf * ((p.mu[(size_t)m * Nz + k] * d.imet[m]) * s);
  }
  for (int o = 0; o < kBand; ++o) {
    const T cf = d.DRT[j * kBand + o];
    if (cf == T(0)) continue;
    const int m = j + o - 4;
    const T w = d.wR[m];
    if (w == T(0)) continue;
    T s = 0;
    for (int q = 0; q < kBand; ++q) {
      const T a = d.DR[m * kBand + q];
      if (a != T(0)) s += a * u(m + q - 4, k);
    }
    const size_t I = (size_t)m * Nz + k;
    const T cm = p.mu[I] * d.imet[m];
    T c3 = cm;
    if (p.order == 4) {
      size_t I1 = I + 1 < N ? I + 1 : N - 1;
      c3 = T(0.5) * (cm + p.mu[I1] * d.imet[(int)(I1 / Nz)]);
    }
    accR += cf * (w * c3 * s);
  }
  if (p.order == 4) {
    for (int o = 0; o < kBand; ++o) {
      const T cf = d.D4T[j * kBand + o];
      if (cf == T(0)) continue;
      const int m = j + o - 4;
      const T w = d.w4[m];
      if (w == T(0)) continue;
      T s = 0;
      for (int q = 0; q < kBand; ++q) {
        const T a = d.D4[m * kBand + q];
        if (a != T(0)) s += a * u(m + q - 4, k);
      }
      acc4 += cf * (w * (p.mu[(size_t)m * Nz + k] * d.imet[m]) * s);
    }
  }
  return d.H[j] * acc1 - (accR + acc4);
}

// Generic 1-D operator along z at (j, k) (no H_y*yq prefactor).
template <typename T, class U>
__device__ T line_z(const DevOp<T>& p, const U& u, int j, int k) {
  const DevDir<T>& d = p.z;
  const int Nz = p.Nz;
  const size_t N = (size_t)p.Ny * Nz;
  const size_t row = (size_t)j * Nz;
  T acc1 = 0, accR = 0, acc4 = 0;
  for (int o = 0; o < kBand; ++o) {
    const T cf = d.D1[k * kBand + o];
    if (cf == T(0)) continue;
    const int m = k + o - 4;
    T s = 0;
    for (int q = 0; q < kBand; ++q) {
      const T w = d.D1[m * kBand + q];
      if (w != T(0)) s += w * u(j, m + q - 4);
    }
    acc1 += cf * ((p.mu[row + m] * d.imet[m]) * s);
  }
  for (int o = 0; o < kBand; ++o) {
    const T cf = d.DRT[k * kBand + o];
    if (cf == T(0)) continue;
    const int m = k + o - 4;
    const T w = d.wR[m];
    if (w == T(0)) continue;
    T s = 0;
    for (int q = 0; q < kBand; ++q) {
      const T a = d.DR[m * kBand + q];
      if (a != T(0)) s += a * u(j, m + q - 4);
    }
    const size_t I = row + m;
    const T cm = p.mu[I] * d.imet[m];
    T c3 = cm;
    if (p.order == 4) {
      size_t I1 = I + 1 < N ? I + 1 : N - 1;
      c3 = T(0.5) * (cm + p.mu[I1] * d.imet[(int)(I1 % Nz)]);
    }
    accR += cf * (w * c3 * s);
  }
  if (p.order == 4) {
    for (int o = 0; o < kBand; ++o) {
      const T cf = d.D4T[k * kBand + o];
      if (cf == T(0)) continue;
      const int m = k + o - 4;
      const T w = d.w4[m];
      if (w == T(0)) continue;
      T s = 0;
      for (int q = 0; q < kBand; ++q) {
        const T a = d.D4[m * kBand + q];
        if (a != T(0)) s += a * u(j, m + q - 4);
      }
      acc4 += cf * (w * (p.mu[row + m] * d.imet[m]) * s);
    }
  }
  return d.H[k] * acc1 - (accR + acc4);
}

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

// SAT terms at (j, k) (sbpOps_m_varGrid.cpp:601-772 constructBC_*).
template <typename T, class U>
__device__ T sat_terms(const DevOp<T>& p, const U& u, int j, int k) {
  T s = 0;
  const int Ny = p.Ny, Nz = p.Nz;
  if (p.doY) {
    const DevDir<T>& d = p.y;
    const T pre = p.z.H[k] * p.z.met[k];
    if (j < kBand) {
      if (p.bcL == SC_DIRICHLET) {
        const T bs = d.BS0[j] + (j == 0 ? d.alphaD : T(0));
        if (bs != T(0)) s += pre * (bs * (p.mu[k] * d.imet[0])) * u(0, k);
      } else if (j == 0) {
        s += pre * (p.mu[k] * (d.imet[0] * d1_row(d, 0, [&](int m) { return u(m, k); })));
      }
    }
    if (j >= Ny - kBand) {
      const size_t IN = (size_t)(Ny - 1) * Nz + k;
      if (p.bcR == SC_DIRICHLET) {
        const T bs = d.BSN[j - (Ny - kBand)] + (j == Ny - 1 ? d.alphaD : T(0));
        if (bs != T(0)) s += pre * (bs * (p.mu[IN] * d.imet[Ny - 1])) * u(Ny - 1, k);
      } else if (j == Ny - 1) {
        s -= pre * (p.mu[IN] * (d.imet[Ny - 1] * d1_row(d, Ny - 1, [&](int m) { return u(m, k); })));
      }
    }
  }
  if (p.doZ) {
    const DevDir<T>& d = p.z;
    const T pre = p.y.H[j] * p.y.met[j];
    const size_t row = (size_t)j * Nz;
    if (k < kBand) {
      if (p.bcT == SC_DIRICHLET) {
        const T bs = d.BS0[k] + (k == 0 ? d.alphaD : T(0));
        if (bs != T(0)) s += pre * (bs * (p.mu[row] * d.imet[0])) * u(j, 0);
      } else if (k == 0) {
        s += pre * (p.mu[row] * (d.imet[0] * d1_row(d, 0, [&](int m) { return u(j, m); })));
      }
    }
    if (k >= Nz - kBand) {
      if (p.bcB == SC_DIRICHLET) {
        const T bs = d.BSN[k - (Nz - kBand)] + (k == Nz - 1 ? d.alphaD : T(0));
        if (bs != T(0)) s += pre * (bs * (p.mu[row + Nz - 1] * d.imet[Nz - 1])) * u(j, Nz - 1);
      } else if (k == Nz - 1) {
        s -= pre * (p.mu[row + Nz - 1] * (d.imet[Nz - 1] * d1_row(d, Nz - 1, [&](int m) { return u(j, m); })));
      }
    }
  }
  return s;
}

template <typename T, class U>
__device__ __forceinline__ T eval_node(const DevOp<T>& p, const U& u, int j, int k) {
  T v = 0;
  if (p.doY) v += (p.z.H[k] * p.z.met[k]) * line_y(p, u, j, k);
  if (p.doZ) v += (p.y.H[j] * p.y.met[j]) * line_z(p, u, j, k);
  return v + sat_terms(p, u, j, k);
}

// Frame node t -> (j, k).  Frame = rows [0,F) U [Ny-F,Ny) (all k) and, for
// the middle rows, columns [0,F) U [Nz-F,Nz).
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
  __device__ void node(long long t, int& j, int& k) const {
    const long long a = (long long)nTop * Nz;
    if (t < a) { j = (int)(t / Nz); k = (int)(t % Nz); return; }
    t -= a;
    const long long b = (long long)nBot * Nz;
    if (t < b) { j = Ny - nBot + (int)(t / Nz); k = (int)(t % Nz); return; }
    t -= b;
    const int w = nL + nR;
    j = nTop + (int)(t / w);
    const int c = (int)(t % w);
    k = c < nL ? c : Nz - nR + (c - nL);
  }
};

template <typename T>
__global__ __launch_bounds__(256) void k_apply_frame(DevOp<T> p, FrameMap fm,
                                                     const T* __restrict__ u, T* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fm.count()) return;
  int j, k;
  fm.node(t, j, k);
  FieldAcc<T> acc{u, p.Nz};
  out[(size_t)j * p.Nz + k] = eval_node(p, acc, j, k);
}

// ------------------------------------------------------ interior kernel ---
// One wave per block.  Lanes cover columns [k0-4, k0+60); lanes 4..59 write
// outputs (56 columns), the outer 4+4 lanes only feed the z-direction
// stencil.  Each lane sweeps JC output rows: the u and mu columns live in
// registers (rows j0-4..j0+JC+3), y-direction fluxes are computed once per
// row and reused by the transposed stencils, z-direction neighbours come from
// cross-lane shuffles.
constexpr int kJC = 8;
constexpr int kLanesOut = 56;

template <typename T, int ORDER>
__global__ __launch_bounds__(64) void k_apply_interior(DevOp<T> p, const T* __restrict__ u,
                                                       T* __restrict__ out) {
  constexpr int JC = kJC;
  const int Ny = p.Ny, Nz = p.Nz;
  const int lane = threadIdx.x;
  const int k = kFrame + blockIdx.x * kLanesOut - 4 + lane;
  const int j0 = kFrame + blockIdx.y * JC;
  const int kc = k < Nz ? k : Nz - 1;  // k >= kFrame-4 >= 0 always
  T uw[JC + 8];
#pragma unroll
  for (int r = 0; r < JC + 8; ++r) {
    const int jj = min(j0 - 4 + r, Ny - 1);
    uw[r] = u[(size_t)jj * Nz + kc];
  }
  T mw[JC + 4];
#pragma unroll
  for (int r = 0; r < JC + 4; ++r) {
    const int jj = min(j0 - 2 + r, Ny - 1);
    mw[r] = p.mu[(size_t)jj * Nz + kc];
  }
  const DevDir<T>& dy = p.y;
  const DevDir<T>& dz = p.z;
  // ---- y-direction fluxes at rows j0-2 .. j0+JC+1
  T g1[JC + 4], g3[JC + 4], g4[JC + 4];
#pragma unroll
  for (int r = 0; r < JC + 4; ++r) {
    const int m = min(j0 - 2 + r, Ny - 1);
    const T qm = dy.imet[m];
    const T cm = mw[r] * qm;
    const T du = dy.d1[0] * uw[r] + dy.d1[1] * uw[r + 1] + dy.d1[3] * uw[r + 3] + dy.d1[4] * uw[r + 4];
    g1[r] = cm * du;
    if (ORDER == 4) {
      const T mnext = __shfl_down(mw[r], 1, 64);
      const T c3 = T(0.5) * (cm + mnext * qm);
      const T d3u = dy.dr[1] * uw[r + 1] + dy.dr[2] * uw[r + 2] + dy.dr[3] * uw[r + 3] + dy.dr[4] * uw[r + 4];
      g3[r] = dy.wr * c3 * d3u;
      const T d4u = dy.d4[0] * uw[r] + dy.d4[1] * uw[r + 1] + dy.d4[2] * uw[r + 2] +
                    dy.d4[3] * uw[r + 3] + dy.d4[4] * uw[r + 4];
      g4[r] = dy.w4c * cm * d4u;
    } else {
      const T d2u = dy.dr[1] * uw[r + 1] + dy.dr[2] * uw[r + 2] + dy.dr[3] * uw[r + 3];
      g3[r] = dy.wr * cm * d2u;
      g4[r] = T(0);
    }
  }
  const T preY = dz.H[kc] * dz.met[kc];
  const T rzk = dz.imet[kc];
  const T rzk1 = dz.imet[kc + 1 < Nz ? kc + 1 : Nz - 1];
  const bool writer = lane >= 4 && lane < 4 + kLanesOut && k < Nz - kFrame;
#pragma unroll
  for (int ro = 0; ro < JC; ++ro) {
    const int j = j0 + ro;
    const int rr = ro + 2;
    T v = 0;
    if (p.doY) {
      T Ly = dy.Hc * (dy.d1[0] * g1[rr - 2] + dy.d1[1] * g1[rr - 1] + dy.d1[3] * g1[rr + 1] + dy.d1[4] * g1[rr + 2]);
      if (ORDER == 4) {
        Ly -= (dy.drt[0] * g3[rr - 2] + dy.drt[1] * g3[rr - 1] + dy.drt[2] * g3[rr] + dy.drt[3] * g3[rr + 1]) +
              (dy.d4t[0] * g4[rr - 2] + dy.d4t[1] * g4[rr - 1] + dy.d4t[2] * g4[rr] +
               dy.d4t[3] * g4[rr + 1] + dy.d4t[4] * g4[rr + 2]);
      } else {
        Ly -= dy.drt[1] * g3[rr - 1] + dy.drt[2] * g3[rr] + dy.drt[3] * g3[rr + 1];
      }
      v = preY * Ly;
    }
    if (p.doZ) {
      const T uc = uw[ro + 4];
      const T um1 = __shfl_up(uc, 1, 64), um2 = __shfl_up(uc, 2, 64);
      const T up1 = __shfl_down(uc, 1, 64), up2 = __shfl_down(uc, 2, 64);
      const T cmz = mw[ro + 2] * rzk;
      const T g1z = cmz * (dz.d1[0] * um2 + dz.d1[1] * um1 + dz.d1[3] * up1 + dz.d1[4] * up2);
      T g3z, g4z;
      if (ORDER == 4) {
        const T mz1 = __shfl_down(mw[ro + 2], 1, 64);
        const T c3z = T(0.5) * (cmz + mz1 * rzk1);
        g3z = dz.wr * c3z * (dz.dr[1] * um1 + dz.dr[2] * uc + dz.dr[3] * up1 + dz.dr[4] * up2);
        g4z = dz.w4c * cmz * (dz.d4[0] * um2 + dz.d4[1] * um1 + dz.d4[2] * uc + dz.d4[3] * up1 + dz.d4[4] * up2);
      } else {
        g3z = dz.wr * cmz * (dz.dr[1] * um1 + dz.dr[2] * uc + dz.dr[3] * up1);
        g4z = T(0);
      }
      // contribution of this lane's fluxes to the lane at offset -t (t = s-2)
      const T w0 = dz.Hc * dz.d1[0] * g1z - (dz.drt[0] * g3z + dz.d4t[0] * g4z);
      const T w1 = dz.Hc * dz.d1[1] * g1z - (dz.drt[1] * g3z + dz.d4t[1] * g4z);
      const T w2 = -(dz.drt[2] * g3z + dz.d4t[2] * g4z);
      const T w3 = dz.Hc * dz.d1[3] * g1z - (dz.drt[3] * g3z + dz.d4t[3] * g4z);
      const T w4 = dz.Hc * dz.d1[4] * g1z - (dz.drt[4] * g3z + dz.d4t[4] * g4z);
      const T Lz = w2 + __shfl_up(w1, 1, 64) + __shfl_up(w0, 2, 64) + __shfl_down(w3, 1, 64) +
                   __shfl_down(w4, 2, 64);
      v += (dy.H[j < Ny ? j : Ny - 1] * dy.met[j < Ny ? j : Ny - 1]) * Lz;
    }
    if (writer && j < Ny - kFrame) out[(size_t)j * Nz + k] = v;
  }
}

template <typename T>
static int launch_apply_t(const Ctx& c, const DevOp<T>& op, const T* in, T* out, hipStream_t s) {
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  FrameMap fm = FrameMap::make(Ny, Nz);
  const long long nf = fm.count();
  if (nf > 0) {
    const unsigned nb = (unsigned)((nf + 255) / 256);
    hipLaunchKernelGGL(k_apply_frame<T>, dim3(nb), dim3(256), 0, s, op, fm, in, out);
  }
  const int niy = Ny - 2 * kFrame, niz = Nz - 2 * kFrame;
  if (niy > 0 && niz > 0) {
    dim3 grid((niz + kLanesOut - 1) / kLanesOut, (niy + kJC - 1) / kJC);
    if (c.d.order == 4)
      hipLaunchKernelGGL((k_apply_interior<T, 4>), grid, dim3(64), 0, s, op, in, out);
    else
      hipLaunchKernelGGL((k_apply_interior<T, 2>), grid, dim3(64), 0, s, op, in, out);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int launch_apply64(const Ctx& c, const double* in, double* out, hipStream_t s) {
  return launch_apply_t<double>(c, c.op64, in, out, s);
}
int launch_apply32(const Ctx& c, const float* in, float* out, hipStream_t s) {
  return launch_apply_t<float>(c, c.op32, in, out, s);
}

// ----------------------------------------------------------------- diag ---
__global__ __launch_bounds__(256) void k_diag(DevOp<double> p, double* __restrict__ diag,
                                              double* __restrict__ dinv, float* __restrict__ dinv32) {
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)p.Ny * p.Nz;
  if (I >= N) return;
  const int j = (int)(I / p.Nz), k = (int)(I % p.Nz);
  DeltaAcc<double> acc{j, k};
  const double d = eval_node(p, acc, j, k);
  diag[I] = d;
  dinv[I] = 1.0 / d;
  dinv32[I] = (float)(1.0 / d);
}

int launch_diag(const Ctx& c, double* out, hipStream_t s) {
  const size_t N = (size_t)c.N;
  hipLaunchKernelGGL(k_diag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, out,
                     c.dinv64, c.dinv32);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// --------------------------------------------------------------- setRhs ---
__global__ __launch_bounds__(256) void k_set_rhs(DevOp<double> p, double* __restrict__ rhs,
                                                 const double* __restrict__ bcL,
                                                 const double* __restrict__ bcR,
                                                 const double* __restrict__ bcT,
                                                 const double* __restrict__ bcB) {
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ny = p.Ny, Nz = p.Nz;
  if (I >= (size_t)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I % Nz);
  double v = 0;
  if (p.doY) {
    const DevDir<double>& d = p.y;
    const double pre = p.z.H[k] * p.z.met[k];
    if (j < kBand) {
      if (p.bcL == SC_DIRICHLET) {
        const double bs = d.BS0[j] + (j == 0 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[k] * d.imet[0])) * bcL[k];
      } else if (j == 0) {
        v += pre * bcL[k];
      }
    }
    if (j >= Ny - kBand) {
      if (p.bcR == SC_DIRICHLET) {
        const double bs = d.BSN[j - (Ny - kBand)] + (j == Ny - 1 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[(size_t)(Ny - 1) * Nz + k] * d.imet[Ny - 1])) * bcR[k];
      } else if (j == Ny - 1) {
        v -= pre * bcR[k];
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
        if (bs != 0.0) v += pre * (bs * (p.mu[row] * d.imet[0])) * bcT[j];
      } else if (k == 0) {
        v += pre * bcT[j];
      }
    }
    if (k >= Nz - kBand) {
      if (p.bcB == SC_DIRICHLET) {
        const double bs = d.BSN[k - (Nz - kBand)] + (k == Nz - 1 ? d.alphaD : 0.0);
        if (bs != 0.0) v += pre * (bs * (p.mu[row + Nz - 1] * d.imet[Nz - 1])) * bcB[j];
      } else if (k == Nz - 1) {
        v -= pre * bcB[j];
      }
    }
  }
  rhs[I] = v;
}

int launch_set_rhs(const Ctx& c, double* rhs, const double* bcL, const double* bcR,
                   const double* bcT, const double* bcB, hipStream_t s) {
  const size_t N = (size_t)c.N;
  hipLaunchKernelGGL(k_set_rhs, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, rhs,
                     bcL, bcR, bcT, bcB);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// ----------------------------------------------------------- derivatives --
__global__ __launch_bounds__(256) void k_deriv(DevOp<double> p, int which,
                                               const double* __restrict__ u,
                                               double* __restrict__ out) {
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ny = p.Ny, Nz = p.Nz;
  if (I >= (size_t)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I % Nz);
  double v;
  switch (which) {
    case 0:
    case 1: {
      v = p.y.imet[j] * d1_row(p.y, j, [&](int m) { return u[(size_t)m * Nz + k]; });
      if (which == 1) v = p.mu[I] * v;
      break;
    }
    case 2:
    case 3: {
      v = p.z.imet[k] * d1_row(p.z, k, [&](int m) { return u[(size_t)j * Nz + m]; });
      if (which == 3) v = p.mu[I] * v;
      break;
    }
    case 4: v = (p.y.H[j] * p.z.H[k]) * u[I]; break;
    default: v = (p.y.Hinv[j] * p.z.Hinv[k]) * u[I]; break;
  }
  out[I] = v;
}

int launch_deriv(const Ctx& c, int which, const double* in, double* out, hipStream_t s) {
  const size_t N = (size_t)c.N;
  hipLaunchKernelGGL(k_deriv, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, which,
                     in, out);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

__global__ __launch_bounds__(256) void k_fault_sxy(DevOp<double> p, const double* __restrict__ u,
                                                   double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int Nz = p.Nz;
  if (k >= Nz) return;
  const double d = d1_row(p.y, 0, [&](int m) { return u[(size_t)m * Nz + k]; });
  out[k] = p.mu[k] * (p.y.imet[0] * d);
}

int launch_fault_sxy(const Ctx& c, const double* u, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_fault_sxy, dim3((c.d.Nz + 255) / 256), dim3(256), 0, s, c.op64, u, out);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc

// ---------------------------------------------------------------- C ABI ---
using sc::fail;

extern "C" int sc_sbp_apply(sc_ctx* h, int prec, const void* in_d, void* out_d) {
  if (!h || !in_d || !out_d) return fail(SC_ERR_ARG, "sc_sbp_apply: null argument");
  if (in_d == out_d) return fail(SC_ERR_ARG, "sc_sbp_apply: in and out must differ");
  if (prec == SC_FP64) return sc::launch_apply64(h->c, (const double*)in_d, (double*)out_d, h->c.stream);
  if (prec == SC_FP32) return sc::launch_apply32(h->c, (const float*)in_d, (float*)out_d, h->c.stream);
  return fail(SC_ERR_ARG, "sc_sbp_apply: prec must be SC_FP64 or SC_FP32");
}

extern "C" int sc_sbp_diag(sc_ctx* h, double* out_d) {
  if (!h || !out_d) return fail(SC_ERR_ARG, "null argument");
  SC_HIP(hipMemcpyAsync(out_d, h->c.diag64, sizeof(double) * h->c.N, hipMemcpyDeviceToDevice, h->c.stream));
  return SC_OK;
}

extern "C" int sc_sbp_set_rhs(sc_ctx* h, double* rhs_d, const double* bcL_d, const double* bcR_d,
                              const double* bcT_d, const double* bcB_d) {
  if (!h || !rhs_d || !bcL_d || !bcR_d || !bcT_d || !bcB_d) return fail(SC_ERR_ARG, "sc_sbp_set_rhs: null argument");
  return sc::launch_set_rhs(h->c, rhs_d, bcL_d, bcR_d, bcT_d, bcB_d, h->c.stream);
}

static int deriv(sc_ctx* h, int which, const double* in, double* out) {
  if (!h || !in || !out) return fail(SC_ERR_ARG, "null argument");
  if (in == out) return fail(SC_ERR_ARG, "in and out must differ");
  return sc::launch_deriv(h->c, which, in, out, h->c.stream);
}
extern "C" int sc_sbp_dy(sc_ctx* h, const double* in, double* out) { return deriv(h, 0, in, out); }
extern "C" int sc_sbp_mux_dy(sc_ctx* h, const double* in, double* out) { return deriv(h, 1, in, out); }
extern "C" int sc_sbp_dz(sc_ctx* h, const double* in, double* out) { return deriv(h, 2, in, out); }
extern "C" int sc_sbp_mux_dz(sc_ctx* h, const double* in, double* out) { return deriv(h, 3, in, out); }
extern "C" int sc_sbp_h(sc_ctx* h, const double* in, double* out) { return deriv(h, 4, in, out); }
extern "C" int sc_sbp_hinv(sc_ctx* h, const double* in, double* out) { return deriv(h, 5, in, out); }

// Times `reps` back-to-back launches on the context stream with HIP events
// (the stream the kernels run on); returns the mean time per launch in ms and
// the number of nodes the timed kernel(s) write.  which: 0 = FP64 apply,
// 1 = FP32 apply, 2 = FP64 interior kernel only, 3 = FP32 interior kernel only.
extern "C" int sc_bench_apply(sc_ctx* h, int which, int reps, double* ms_per_launch,
                              double* nodes) {
  if (!h || reps <= 0 || !ms_per_launch || which < 0 || which > 3)
    return fail(SC_ERR_ARG, "sc_bench_apply: bad argument");
  sc::Ctx& c = h->c;
  const bool f64 = which == 0 || which == 2;
  const size_t N = (size_t)c.N, esz = f64 ? 8 : 4;
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  const int niy = Ny - 2 * sc::kFrame, niz = Nz - 2 * sc::kFrame;
  if (which >= 2 && (niy <= 0 || niz <= 0)) return fail(SC_ERR_ARG, "grid has no interior block");
  if (nodes) *nodes = which >= 2 ? (double)niy * niz : (double)N;
  void *a = nullptr, *b = nullptr;
  SC_HIP(hipMalloc(&a, N * esz));
  SC_HIP(hipMalloc(&b, N * esz));
  SC_HIP(hipMemsetAsync(a, 0, N * esz, c.stream));
  hipEvent_t e0, e1;
  SC_HIP(hipEventCreate(&e0));
  SC_HIP(hipEventCreate(&e1));
  dim3 grid((niz + sc::kLanesOut - 1) / sc::kLanesOut, (niy + sc::kJC - 1) / sc::kJC);
  auto launch = [&]() -> int {
    switch (which) {
      case 0: return sc::launch_apply64(c, (double*)a, (double*)b, c.stream);
      case 1: return sc::launch_apply32(c, (float*)a, (float*)b, c.stream);
      case 2:
        if (c.d.order == 4) hipLaunchKernelGGL((sc::k_apply_interior<double, 4>), grid, dim3(64), 0, c.stream, c.op64, (const double*)a, (double*)b);
        else hipLaunchKernelGGL((sc::k_apply_interior<double, 2>), grid, dim3(64), 0, c.stream, c.op64, (const double*)a, (double*)b);
        return 0;
      default:
        if (c.d.order == 4) hipLaunchKernelGGL((sc::k_apply_interior<float, 4>), grid, dim3(64), 0, c.stream, c.op32, (const float*)a, (float*)b);
        else hipLaunchKernelGGL((sc::k_apply_interior<float, 2>), grid, dim3(64), 0, c.stream, c.op32, (const float*)a, (float*)b);
        return 0;
    }
  };
  int err = 0;
  for (int w = 0; w < 3 && !err; ++w) err = launch();
  SC_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps && !err; ++r) err = launch();
  SC_HIP(hipEventRecord(e1, c.stream));
  SC_HIP(hipEventSynchronize(e1));
  SC_HIP(hipGetLastError());
  float ms = 0;
  SC_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_launch = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return err;
}
