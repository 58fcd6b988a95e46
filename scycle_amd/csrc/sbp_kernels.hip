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
// One launch does the work: interior blocks (rows/columns at least kFrame
// from every boundary, where all stencils are the constant interior ones)
// and frame blocks, one thread per node within kFrame of a boundary, which
// apply the frame operator assembled once per operator (34 coefficients a
// node: the boundary closures, prefactors and SAT terms folded in).  Grids
// with a side of at most 2*kFrame nodes use a per-node generic kernel
// instead.
#include "fault_dev.hpp"
#include "sbp_dev.hpp"

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
__device__ __forceinline__ T line_y(const DevOp<T>& p, const U& u, int j, int k) {
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
__device__ __forceinline__ T line_z(const DevOp<T>& p, const U& u, int j, int k) {
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

// SAT terms at (j, k) (sbpOps_m_varGrid.cpp:601-772 constructBC_*).
template <typename T, class U>
__device__ __forceinline__ T sat_terms(const DevOp<T>& p, const U& u, int j, int k) {
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

// Raw buffer access (stride 0, the range in bytes) for the tile kernel: a
// lane's column is the vector offset, a row's start the scalar one.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, T x, int voff, int soff) {
  if constexpr (sizeof(T) == 8) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), r, voff, soff, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, 0);
  }
}

// ------------------------------------------------------------ epilogues --
// What the apply kernels do with (A u) at node (j, k), index I: `uc` is u at
// the node.  operator() serves the frame nodes (any node, the boundary's
// terms included); the interior tile kernel instead fetches each output
// row's operands (pre) two rows before it stores that row (interior), all
// through buffer resources made once per wave (rows), so that no row's store
// waits for a load issued just before it.
template <typename T>
struct StoreAu {  // out = A u
  T* __restrict__ out;
  __device__ __forceinline__ void operator()(int, int, size_t I, T Au, T, bool) const { out[I] = Au; }
  struct Rows {
    __amdgpu_buffer_rsrc_t o;
  };
  struct Pre {};
  __device__ __forceinline__ Rows rows(int bytes) const { return {buf_rsrc(out, bytes)}; }
  __device__ __forceinline__ Pre pre(const Rows&, int, int) const { return {}; }
  __device__ __forceinline__ void interior(const Rows& r, int, int, int voff, int soff, T Au, T,
                                           const Pre&) const {
    buf_st<T>(r.o, Au, voff, soff);
  }
};

// propagateWaves (strikeSlip_linearElastic_fd.cpp:445-516):
//   D2u = Jinv * Hinv * (A u),  Jinv = qy*rz, Hinv = Hy^-1 (x) Hz^-1
//   uNext = (dt^2/rho * D2u + 2u + (dt*ay - 1)*uPrev) / (dt*ay + 1)
// a_y is zero off the domain boundary, where the update reduces exactly to
// dt^2/rho*D2u + 2u - uPrev (every interior tile node).  D2u on the fault
// (j = 0) is kept for the fault update.  C1: `rho` holds the whole per-node
// scale s = dt^2/rho * Hy^-1 Hz^-1 (* qy rz on a variable grid), formed once
// for a run at constant dt (k_wave_c1), so the update is s (A u) + 2u - uPrev.
template <typename T, bool C1>
struct WaveUpdate {
  T* __restrict__ uNext;
  const T* __restrict__ uPrev;
  const T* __restrict__ rho;  // C1: the per-node scale s
  const T* __restrict__ ay;
  const T *hinvY, *hinvZ, *qy, *rz;
  T* __restrict__ d2uFault;
  T dt;
  int varGrid;
  __device__ __forceinline__ T d2u_of(int j, int k, T Au) const {
    T d2u = (hinvY[j] * hinvZ[k]) * Au;
    if (varGrid) d2u = (qy[j] * rz[k]) * d2u;
    return d2u;
  }
  // the update's (dt^2/rho) D2u term at node I, given rho or s there
  __device__ __forceinline__ T scaled(int j, int k, T r, T Au) const {
    if constexpr (C1) return r * Au;
    else return (dt * dt / r) * d2u_of(j, k, Au);
  }
  __device__ __forceinline__ void operator()(int j, int k, size_t I, T Au, T uc, bool edge) const {
    if (j == 0) d2uFault[k] = d2u_of(j, k, Au);
    const T sAu = scaled(j, k, rho[I], Au);
    if (edge) {
      const T a = ay[I];
      const T c2 = dt * a - T(1), c3 = dt * a + T(1);
      uNext[I] = (sAu + T(2) * uc + c2 * uPrev[I]) / c3;
    } else {
      uNext[I] = sAu + T(2) * uc + T(-1) * uPrev[I];
    }
  }
  struct Rows {
    __amdgpu_buffer_rsrc_t n, p, r;
  };
  struct Pre {
    T r, up;
  };
  __device__ __forceinline__ Rows rows(int bytes) const {
    return {buf_rsrc(uNext, bytes), buf_rsrc(uPrev, bytes), buf_rsrc(rho, bytes)};
  }
  __device__ __forceinline__ Pre pre(const Rows& b, int voff, int soff) const {
    return {buf_ld<T>(b.r, voff, soff), buf_ld<T>(b.p, voff, soff)};
  }
  __device__ __forceinline__ void interior(const Rows& b, int j, int k, int voff, int soff, T Au, T uc,
                                           const Pre& q) const {
    buf_st<T>(b.n, scaled(j, k, q.r, Au) + T(2) * uc + T(-1) * q.up, voff, soff);
  }
};

// The wave update with the fault's step fused in (Fault_fd::d_dt after
// propagateWaves, strikeSlip_linearElastic_fd.cpp:421-445): at a fault node
// (j = 0, a frame node) the thread that forms D2u there runs the node's
// slip-velocity and state solves (fault_fd_node) and writes the fault
// displacement into uNext, so the fault's root solves run beside the
// interior tiles of the same launch instead of in a launch of their own.
// The arithmetic is the two launches' own, so the results are the same bit
// for bit.
template <typename T>
struct WaveFault {
  WaveUpdate<T, true> w;
  sc_fault_fd_args f;
  __device__ __forceinline__ void operator()(int j, int k, size_t I, T Au, T uc, bool edge) const {
    if (j != 0) {
      w(j, k, I, Au, uc, edge);
      return;
    }
    const T d2u = w.d2u_of(j, k, Au);
    w.d2uFault[k] = d2u;
    w.uNext[I] = (T)fault_fd_node(f, k, (double)d2u);
  }
  using Rows = typename WaveUpdate<T, true>::Rows;
  using Pre = typename WaveUpdate<T, true>::Pre;
  __device__ __forceinline__ Rows rows(int bytes) const { return w.rows(bytes); }
  __device__ __forceinline__ Pre pre(const Rows& b, int voff, int soff) const { return w.pre(b, voff, soff); }
  __device__ __forceinline__ void interior(const Rows& b, int j, int k, int voff, int soff, T Au, T uc,
                                           const Pre& q) const {
    w.interior(b, j, k, voff, soff, Au, uc, q);
  }
};

// Frame node t -> (j, k).  Frame = rows [0,F) U [Ny-F,Ny) (all k) and, for
// the middle rows, columns [0,F) U [Nz-F,Nz).

template <typename T, class EP>
__global__ __launch_bounds__(256) void k_apply_frame(DevOp<T> p, FrameMap fm,
                                                     const T* __restrict__ u, EP ep) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fm.count()) return;
  int j, k;
  fm.node(t, j, k);
  FieldAcc<T> acc{u, p.Nz};
  const size_t I = (size_t)j * p.Nz + k;
  ep(j, k, I, eval_node(p, acc, j, k), u[I], true);
}

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1 of the GFX9
// family): lane i receives lane i-1 (up) or lane i+1 (down), as
// __shfl_up/__shfl_down(x, 1, 64) but as a VALU operand modifier rather than
// an LDS permute.  The end lane receives zero (bound_ctrl, which also spares
// the copy of the old value into the destination); the kernels never use the
// values shifted in at the wave's edges.
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float x) {
  const int v = __float_as_int(x);
  return __int_as_float(__builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  const int lo2 = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  const int hi2 = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi2 << 32) | (unsigned int)lo2);
}
// a quad permutation (every lane of the quad has a source)
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_quad(T x) {
  return dpp_shift<CTRL>(x);
}
template <typename T>
__device__ __forceinline__ T lane_up1(T x) { return dpp_shift<kDppWaveShr1>(x); }
template <typename T>
__device__ __forceinline__ T lane_down1(T x) { return dpp_shift<kDppWaveShl1>(x); }

// ------------------------------------------------------ interior kernel ---
// One wave per block.  Lanes cover columns [k0-4, k0+60); lanes 4..59 write
// outputs (56 columns), the outer 4+4 lanes only feed the z-direction
// stencil.  Each lane sweeps JC output rows: the u and mu columns live in
// registers (rows j0-4..j0+JC+3), y-direction fluxes are computed once per
// row and reused by the transposed stencils, z-direction neighbours come from
// cross-lane shuffles.
//
// Tiles are numbered down the column blocks (tile t: column block t / nby,
// row block t % nby) so that vertically adjacent tiles, which share eight halo
// rows of u, are neighbours.  With XCD = 1 the linear block id is remapped
// so that each of the eight XCDs (which receive blocks round-robin) owns a
// contiguous range of tiles and the halo rows stay in its own L2.
// ------------------------------------------- assembled frame operator ----
// At a frame node the operator (both directions, their prefactors and the
// SAT terms) is a cross of two 17-point stencils: cy[r] on u(j-8+r, k) and
// cz[r] on u(j, k-8+r), zero outside the grid.  Its coefficients depend only
// on mu, the grid and the BC types, so they are formed once per operator
// (k_frame_coefs, the same algebra as the closure and SAT terms with the sums
// taken per u entry) and the frame apply is a 34-term product per node,
// coefficients stored coef[r * nFrame + t] for frame node t (FrameMap order).
constexpr int kCross = 17;
constexpr int kFrameCoefs = 2 * kCross;

// Four lanes per frame node: lane s of the quad takes the y-taps 0..8 (s = 0),
// 9..16 (s = 1) or the z-taps 0..8 (s = 2), 9..16 (s = 3) of the node's two
// 17-point stencils (18 loads a lane instead of 68 for one lane), and a quad
// sum over two lane swaps gives lane 0 the node's value.
constexpr int kFrameQuad = 4;
constexpr int kDppQuadSwap1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4E;  // quad_perm [2,3,0,1]

template <typename T, class EP>
__device__ __forceinline__ void frame_cf_quad(const DevOp<T>& p, const FrameMap& fm, const T* __restrict__ coef,
                                              const T* __restrict__ u, const EP& ep, long long t, int sub) {
  const long long nf = fm.count();
  const bool valid = t < nf;
  const long long tc = valid ? t : nf - 1;  // every lane takes part in the quad sums
  int j, k;
  fm.node(tc, j, k);
  const int Ny = p.Ny, Nz = p.Nz;
  const bool ydir = sub < 2;
  const int r0 = (sub & 1) ? 9 : 0;
  const size_t cbase = (size_t)(ydir ? 0 : kCross) * (size_t)nf + (size_t)tc;
  T v = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {  // clamped loads, then selects: all in flight together
    const int r = r0 + i;
    const bool tap = r < kCross;
    const int rc = tap ? r : kCross - 1;
    const int jj = ydir ? j - 8 + rc : j;
    const int kk = ydir ? k : k - 8 + rc;
    const bool in = tap && jj >= 0 && jj < Ny && kk >= 0 && kk < Nz;
    const T x = u[(size_t)min(max(jj, 0), Ny - 1) * Nz + min(max(kk, 0), Nz - 1)];
    const T w = coef[cbase + (size_t)rc * (size_t)nf];
    v += w * (in ? x : T(0));
  }
  v += dpp_quad<kDppQuadSwap1>(v);
  v += dpp_quad<kDppQuadSwap2>(v);
  if (valid && sub == 0) {
    const size_t I = (size_t)j * Nz + k;
    ep(j, k, I, v, u[I], true);
  }
}

// output rows of an interior tile, and how many output rows ahead of the
// one it computes a tile's wave issues its loads: u and mu (In; kJC: every
// row at the start, measured faster than streaming them in) and the
// epilogue's operands (Out; two rows, so that no row's store waits on a load
// issued just before it)
constexpr int kJC = 8;
constexpr int kAheadIn = kJC;
constexpr int kAheadOut = 2;
constexpr int kLanesOut = 56;
constexpr int kXcd = 8;

__device__ __forceinline__ int tile_of_block(int b, int total, int xcd) {
  if (!xcd) return b;
  const int q = total / kXcd, r = total % kXcd;
  const int x = b % kXcd, slot = b / kXcd;
  return x * q + min(x, r) + slot;
}

//
// The first nfb blocks (a multiple of kXcd, so the remap above is unchanged
// for the rest) apply the assembled frame operator to 64 frame nodes each:
// one launch covers the whole grid, and the frame's latency-bound loads
// overlap the interior's streaming.
template <typename T, int ORDER, class EP, int XCD = 1>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_apply_interior(
                                                       DevOp<T> p, const T* __restrict__ u, EP ep,
                                                       int nbx, int nby, FrameMap fm,
                                                       const T* __restrict__ fcoef, int nfb) {
  constexpr int JC = kJC;
  const int Ny = p.Ny, Nz = p.Nz;
  const int lane = threadIdx.x;
  if ((int)blockIdx.x < nfb) {
    frame_cf_quad(p, fm, fcoef, u, ep, (long long)blockIdx.x * (64 / kFrameQuad) + lane / kFrameQuad,
                  lane % kFrameQuad);
    return;
  }
  const int tile = tile_of_block((int)blockIdx.x - nfb, nbx * nby, XCD);
  const int bx = tile / nby, by = tile - bx * nby;
  const int k = kFrame + bx * kLanesOut - 4 + lane;
  const int j0 = kFrame + by * JC;
  const int kc = k < Nz ? k : Nz - 1;  // k >= kFrame-4 >= 0 always
  // u and mu through buffer loads: the lane's column is the one vector
  // offset, each row's start a scalar offset (rows clamped to the grid on the
  // scalar side), so the loads cost no vector address arithmetic
  const int voff = kc * (int)sizeof(T);
  const int rowB = Nz * (int)sizeof(T);
  const int nB = Ny * rowB;
  const __amdgpu_buffer_rsrc_t ru = buf_rsrc(u, nB), rm = buf_rsrc(p.mu, nB);
  const DevDir<T>& dy = p.y;
  const DevDir<T>& dz = p.z;
  // the lane's column weights.  The laplace type's direction switches are
  // zero factors: no load sits in a branch (a branch's loads are waited for
  // inside it).
  const T selY = p.doY ? T(1) : T(0), selZ = p.doZ ? T(1) : T(0);
  const T hzk = dz.H[kc], mzk = dz.met[kc];
  const T rzk = dz.imet[kc];
  const T rzk1 = dz.imet[kc + 1 < Nz ? kc + 1 : Nz - 1];
  // The tile's rows: u rows j0-4 .. j0+JC+3 (uw), mu and the y-metric at the
  // flux rows j0-2 .. j0+JC+1 (mw, qmr), the output rows' z-weights (myr) and
  // the epilogue's operands (pq), each loaded kAheadIn or kAheadOut output
  // rows before the row that first reads it.
  // (the per-row factors qmr, myr are uniform: scalar loads where no
  // store of the tile precedes them, vector loads after)
  T uw[JC + 8], mw[JC + 4], qmr[JC + 4], myr[JC];
  auto ld_u = [&](int r) {
    if (r < JC + 8) uw[r] = buf_ld<T>(ru, voff, min(j0 - 4 + r, Ny - 1) * rowB);
  };
  auto ld_m = [&](int r) {
    if (r < JC + 4) {
      const int jm = min(j0 - 2 + r, Ny - 1);
      mw[r] = buf_ld<T>(rm, voff, jm * rowB);
      qmr[r] = dy.imet[jm];
    }
  };
  const typename EP::Rows er = ep.rows(nB);
  typename EP::Pre pq[JC];
  auto ld_out = [&](int ro) {
    if (ro < JC) {
      const int jc = min(j0 + ro, Ny - 1);
      myr[ro] = dy.met[jc];  // (H is the interior weight Hc on these rows)
      pq[ro] = ep.pre(er, voff, jc * rowB);
    }
  };
  // what the flux prologue and output row 0 read, and the rows ahead
#pragma unroll
  for (int r = 0; r < 8 + kAheadIn; ++r) ld_u(r);
#pragma unroll
  for (int r = 0; r < 4 + kAheadIn; ++r) ld_m(r);
#pragma unroll
  for (int ro = 0; ro < kAheadOut; ++ro) ld_out(ro);
  // (the loads stay where they are written: the scheduler would otherwise
  // sink them towards their uses and wait for each group in turn)
  __builtin_amdgcn_sched_barrier(0);
  const T preY = selY * (hzk * mzk);
  // ---- y-direction fluxes at rows j0-2 .. j0+JC+1 (factored interior rows:
  // D1 u = ia (u[+1]-u[-1]) + ib (u[+2]-u[-2]), DR = [-1 3 -3 1] or the
  // three-point D2, D4 = [1 -4 6 -4 1])
  T g1[JC + 4], g3[JC + 4], g4[JC + 4];
  // flux row r (grid row j0-2+r); each is formed just before the first
  // output row that needs it, so that only five rows of fluxes are live
  auto flux = [&](int r) {
    const T qm = qmr[r];
    const T cm = mw[r] * qm;
    T du = dy.ia * (uw[r + 3] - uw[r + 1]);
    if (ORDER == 4) du += dy.ib * (uw[r + 4] - uw[r]);
    g1[r] = cm * du;
    if (ORDER == 4) {
      const T mnext = lane_down1(mw[r]);
      const T c3 = mnext * qm + cm;  // twice the two-point average (the 1/2 is in g3s)
      const T d3u = (uw[r + 4] - uw[r + 1]) + T(3) * (uw[r + 2] - uw[r + 3]);
      g3[r] = (dy.g3s * c3) * d3u;
      const T d4u = (uw[r] + uw[r + 4]) + T(6) * uw[r + 2] - T(4) * (uw[r + 1] + uw[r + 3]);
      g4[r] = (dy.g4s * cm) * d4u;
    } else {
      const T d2u = (uw[r + 1] + uw[r + 3]) - T(2) * uw[r + 2];
      g3[r] = (dy.g3s * cm) * d2u;
      g4[r] = T(0);
    }
  };
  const bool writer = lane >= 4 && lane < 4 + kLanesOut && k < Nz - kFrame;
#pragma unroll
  for (int r = 0; r < 4; ++r) flux(r);
#pragma unroll
  for (int ro = 0; ro < JC; ++ro) {
    const int j = j0 + ro;
    const int rr = ro + 2;
    ld_u(ro + 8 + kAheadIn);
    ld_m(ro + 4 + kAheadIn);
    ld_out(ro + kAheadOut);
    flux(ro + 4);
    T v;
    {
      // transposed combination: H D1^T on g1, DR^T on g3, D4^T on g4
      T Ly = dy.Ha * (g1[rr + 1] - g1[rr - 1]);
      if (ORDER == 4) {
        Ly += dy.Hb * (g1[rr + 2] - g1[rr - 2]);
        const T r3 = (g3[rr - 2] - g3[rr + 1]) + T(3) * (g3[rr] - g3[rr - 1]);
        const T r4 = (g4[rr - 2] + g4[rr + 2]) + T(6) * g4[rr] - T(4) * (g4[rr - 1] + g4[rr + 1]);
        Ly -= r3 + r4;
      } else {
        Ly -= dy.rts * ((g3[rr - 1] + g3[rr + 1]) - T(2) * g3[rr]);
      }
      v = preY * Ly;
    }
    {
      const T uc = uw[ro + 4];
      const T um1 = lane_up1(uc), um2 = lane_up1(um1);
      const T up1 = lane_down1(uc), up2 = lane_down1(up1);
      const T cmz = mw[ro + 2] * rzk;
      T duz = dz.ia * (up1 - um1);
      if (ORDER == 4) duz += dz.ib * (up2 - um2);
      const T g1z = cmz * duz;
      // w_s: this lane's fluxes' share of the lane at offset 2 - s
      T Lz;
      if (ORDER == 4) {
        const T mz1 = lane_down1(mw[ro + 2]);
        const T g3z = (dz.g3s * (mz1 * rzk1 + cmz)) * ((up2 - um1) + T(3) * (uc - up1));
        const T g4z = (dz.g4s * cmz) * ((um2 + up2) + T(6) * uc - T(4) * (um1 + up1));
        const T t3 = T(3) * g3z, t4 = T(4) * g4z;
        const T w0 = -(dz.Hb * g1z) - (g3z + g4z);
        const T w1 = (t3 + t4) - dz.Ha * g1z;
        const T w2 = -(t3 + T(6) * g4z);
        const T w3 = dz.Ha * g1z + (g3z + t4);
        const T w4 = dz.Hb * g1z - g4z;
        Lz = w2 + lane_up1(w1 + lane_up1(w0)) + lane_down1(w3 + lane_down1(w4));
      } else {
        const T g3z = (dz.g3s * cmz) * ((um1 + up1) - T(2) * uc);
        const T hr = dz.rts * g3z;
        const T w1 = -(dz.Ha * g1z) - hr;
        const T w3 = dz.Ha * g1z - hr;
        Lz = T(2) * hr + lane_up1(w1) + lane_down1(w3);
      }
      v += (selZ * (dy.Hc * myr[ro])) * Lz;
    }
    // (v formed by every lane: a value only the store reads would be sunk
    // into the store's branch, its loads with it)
    asm volatile("" : "+v"(v));
    if (writer && j < Ny - kFrame) ep.interior(er, j, k, voff, j * rowB, v, uw[ro + 4], pq[ro]);
  }
}

// Copies the four packed closure tables (ClosureTab, built at context
// creation) into LDS: every thread issues all of its 16-byte loads before its
// first LDS store, so the copy costs one memory latency.
constexpr int kFrameThreads = 128;

template <typename T>
__device__ __forceinline__ void stage_closures(ClosureTab<T> (&tab)[4], const ClosureTab<T>* src) {
  constexpr int n16 = (int)(4 * sizeof(ClosureTab<T>) / 16);
  constexpr int per = (n16 + kFrameThreads - 1) / kFrameThreads;
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(&tab[0]);
  uint4 v[per];
#pragma unroll
  for (int i = 0; i < per; ++i) {  // unconditional (clamped) loads, all in flight together
    const int e = threadIdx.x + i * kFrameThreads;
    v[i] = s[e < n16 ? e : n16 - 1];
  }
#pragma unroll
  for (int i = 0; i < per; ++i) {
    const int e = threadIdx.x + i * kFrameThreads;
    if (e < n16) d[e] = v[i];
  }
}

// ---------------------------------------- frame coefficient assembly ----
// 1-D operator at line index i within kFrame of one end of a line of length n
// (n > 2*kFrame), in flux form with the boundary closures of the staged table
// t, as coefficients on u at line index i-8+r (cf[r], scaled by scale):
//   L(u)[i] = H[i] sum_o oD1[i][o] c_m (fD1 u)_m - oDRT[i][o] wR_m c3_m (fDR u)_m
//             - oD4T[i][o] w4_m c_m (fD4 u)_m,   m = i-4+o
// with cw[o] = c and c3w[o] = its two-point average at flux row m.
template <typename T, int ORDER>
__device__ __forceinline__ void closure_coefs(const ClosureTab<T>& t, int n, int i, const T (&cw)[kBand],
                                              const T (&c3w)[kBand], T scale, T (&cf)[kCross]) {
  const bool hi = i >= kFrame;
  const int mb = hi ? n - 12 : 0, jb = hi ? n - kFrame : 0;
  const int io = i - jb;
  const T h = t.H[io];
  for (int o = 0; o < kBand; ++o) {
    const int m = i + o - 4;
    if (m < 0 || m >= n) continue;
    const int fm = m - mb;
    const T cm = cw[o];
    const T a1 = scale * (h * (t.oD1[io * kBand + o] * cm));
    const T wr = t.wR[fm];
    const T aR = wr != T(0) ? scale * (t.oDRT[io * kBand + o] * (wr * (ORDER == 4 ? c3w[o] : cm))) : T(0);
    const T w4 = ORDER == 4 ? t.w4[fm] : T(0);
    const T a4 = w4 != T(0) ? scale * (t.oD4T[io * kBand + o] * (w4 * cm)) : T(0);
    for (int q = 0; q < kBand; ++q) {
      T v = a1 * t.fD1[fm * kBand + q] - aR * t.fDR[fm * kBand + q];
      if (ORDER == 4) v -= a4 * t.fD4[fm * kBand + q];
      cf[o + q] += v;
    }
  }
}

// The constant interior 1-D operator (a node at least kFrame from both ends
// of its line) as coefficients on u at offsets -4..4 (cf[4..12]); cs[s],
// c3s[s] are c and its two-point average at the five flux rows offset s-2.
template <typename T, int ORDER>
__device__ __forceinline__ void interior_coefs(const DevDir<T>& d, const T (&cs)[5], const T (&c3s)[5], T scale,
                                               T (&cf)[kCross]) {
  for (int sf = 0; sf < 5; ++sf) {
    const T a1 = (sf == 2) ? T(0) : scale * (d.Hc * d.d1[sf] * cs[sf]);
    const T aR = scale * (d.drt[sf] * d.wr * (ORDER == 4 ? c3s[sf] : cs[sf]));
    const T a4 = ORDER == 4 ? scale * (d.d4t[sf] * d.w4c * cs[sf]) : T(0);
    for (int q = 0; q < 5; ++q) {
      T v = (q == 2 ? T(0) : a1 * d.d1[q]);
      if (ORDER == 4) {
        if (q >= 1) v -= aR * d.dr[q];
        v -= a4 * d.d4[q];
      } else if (q >= 1 && q <= 3) {
        v -= aR * d.dr[q];
      }
      cf[4 + sf + q] += v;
    }
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kFrameThreads) void k_frame_coefs(DevOp<T> p, FrameMap fm, T* __restrict__ coef) {
  __shared__ __align__(16) ClosureTab<T> tab[4];
  stage_closures(tab, p.ctab);
  __syncthreads();
  const long long nf = fm.count();
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  int j, k;
  fm.node(t, j, k);
  const int Ny = p.Ny, Nz = p.Nz;
  const T* mu = p.mu;
  T cy[kCross], cz[kCross];
  for (int r = 0; r < kCross; ++r) cy[r] = cz[r] = T(0);
  if (p.doY) {
    const T pre = p.z.H[k] * p.z.met[k];
    const T* im = p.y.imet;
    if (j < kFrame || j >= Ny - kFrame) {
      T cw[kBand], c3w[kBand];
      // c = mu*qy at the nine flux rows m = j-4+o (clamped: rows outside the
      // line get zero coefficients); c3 = average with global index I+1: (m, k+1),
      // the first node of row m+1 after the last column, or I at the last node
      for (int o = 0; o < kBand; ++o) {
        const int m = min(max(j + o - 4, 0), Ny - 1);
        const size_t I = (size_t)m * Nz + k;
        const bool inRow = k + 1 < Nz, lastRow = m + 1 >= Ny;
        const size_t In = inRow ? I + 1 : (lastRow ? I : (size_t)(m + 1) * Nz);
        const int mn = (inRow || lastRow) ? m : m + 1;
        const T cm = mu[I] * im[m];
        cw[o] = cm;
        c3w[o] = T(0.5) * (cm + mu[In] * im[mn]);
      }
      closure_coefs<T, ORDER>(tab[j >= kFrame ? 1 : 0], Ny, j, cw, c3w, pre, cy);
    } else {
      T cs[5], c3s[5];
      for (int sf = 0; sf < 5; ++sf) {  // flux rows j-2..j+2
        const int m = j - 2 + sf;
        const T cm = mu[(size_t)m * Nz + k] * im[m];
        const bool inRow = k + 1 < Nz;
        const T mn = mu[inRow ? (size_t)m * Nz + k + 1 : (size_t)(m + 1) * Nz];
        const T qn = im[inRow ? m : m + 1];
        cs[sf] = cm;
        c3s[sf] = T(0.5) * (cm + mn * qn);
      }
      interior_coefs<T, ORDER>(p.y, cs, c3s, pre, cy);
    }
    // SAT terms of the y-sides (constructBC_*, sbpOps_m_varGrid.cpp:601-772): BS
    // rows on the boundary value, or the boundary row of D1 (Neumann)
    const DevDir<T>& d = p.y;
    if (j < kFrame) {
      if (p.bcL == SC_DIRICHLET) {
        const T bs = tab[0].BS[j] + (j == 0 ? d.alphaD : T(0));
        if (bs != T(0)) cy[8 - j] += pre * (bs * (mu[k] * d.imet[0]));
      } else if (j == 0) {
        for (int q = 4; q < kBand; ++q) cy[8 + q - 4] += pre * (mu[k] * (d.imet[0] * tab[0].fD1[q]));
      }
    }
    if (j >= Ny - kFrame) {
      const size_t IN = (size_t)(Ny - 1) * Nz + k;
      if (p.bcR == SC_DIRICHLET) {
        const T bs = tab[1].BS[j - (Ny - kBand)] + (j == Ny - 1 ? d.alphaD : T(0));
        if (bs != T(0)) cy[8 + (Ny - 1 - j)] += pre * (bs * (mu[IN] * d.imet[Ny - 1]));
      } else if (j == Ny - 1) {
        for (int q = 0; q < 5; ++q)
          cy[8 + q - 4] -= pre * (mu[IN] * (d.imet[Ny - 1] * tab[1].fD1[11 * kBand + q]));
      }
    }
  }
  if (p.doZ) {
    const T pre = p.y.H[j] * p.y.met[j];
    const T* im = p.z.imet;
    const size_t row = (size_t)j * Nz;
    if (k < kFrame || k >= Nz - kFrame) {
      const bool lastRow = j + 1 >= Ny;
      T cw[kBand], c3w[kBand];
      for (int o = 0; o < kBand; ++o) {  // as along y, within row j
        const int m = min(max(k + o - 4, 0), Nz - 1);
        const bool inRow = m + 1 < Nz;
        const size_t In = inRow ? row + m + 1 : (lastRow ? row + m : row + Nz);
        const int kn = inRow ? m + 1 : (lastRow ? m : 0);
        const T cm = mu[row + m] * im[m];
        cw[o] = cm;
        c3w[o] = T(0.5) * (cm + mu[In] * im[kn]);
      }
      closure_coefs<T, ORDER>(tab[k >= kFrame ? 3 : 2], Nz, k, cw, c3w, pre, cz);
    } else {
      T cs[5], c3s[5];
      for (int sf = 0; sf < 5; ++sf) {  // flux columns k-2..k+2
        const int m = k - 2 + sf;
        const T cm = mu[row + m] * im[m];
        cs[sf] = cm;
        c3s[sf] = T(0.5) * (cm + mu[row + m + 1] * im[m + 1]);
      }
      interior_coefs<T, ORDER>(p.z, cs, c3s, pre, cz);
    }
    const DevDir<T>& d = p.z;
    if (k < kFrame) {
      if (p.bcT == SC_DIRICHLET) {
        const T bs = tab[2].BS[k] + (k == 0 ? d.alphaD : T(0));
        if (bs != T(0)) cz[8 - k] += pre * (bs * (mu[row] * d.imet[0]));
      } else if (k == 0) {
        for (int q = 4; q < kBand; ++q) cz[8 + q - 4] += pre * (mu[row] * (d.imet[0] * tab[2].fD1[q]));
      }
    }
    if (k >= Nz - kFrame) {
      if (p.bcB == SC_DIRICHLET) {
        const T bs = tab[3].BS[k - (Nz - kBand)] + (k == Nz - 1 ? d.alphaD : T(0));
        if (bs != T(0)) cz[8 + (Nz - 1 - k)] += pre * (bs * (mu[row + Nz - 1] * d.imet[Nz - 1]));
      } else if (k == Nz - 1) {
        for (int q = 0; q < 5; ++q)
          cz[8 + q - 4] -= pre * (mu[row + Nz - 1] * (d.imet[Nz - 1] * tab[3].fD1[11 * kBand + q]));
      }
    }
  }
  for (int r = 0; r < kCross; ++r) {
    coef[(size_t)r * nf + t] = cy[r];
    coef[(size_t)(kCross + r) * nf + t] = cz[r];
  }
}

template <typename T, class EP>
__global__ __launch_bounds__(256) void k_apply_frame_cf(DevOp<T> p, FrameMap fm, const T* __restrict__ coef,
                                                       const T* __restrict__ u, EP ep) {
  frame_cf_quad(p, fm, coef, u, ep, ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kFrameQuad,
                (int)(threadIdx.x % kFrameQuad));
}

template <typename T>
static int build_frame_coefs(const Ctx& c, const DevOp<T>& op, void*& buf) {
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  if (!(Ny > 2 * kFrame && Nz > 2 * kFrame)) return SC_OK;
  const FrameMap fm = FrameMap::make(Ny, Nz);
  const long long nf = fm.count();
  if (!buf) SC_HIP(hipMalloc(&buf, (size_t)kFrameCoefs * nf * sizeof(T)));
  const unsigned nb = (unsigned)((nf + kFrameThreads - 1) / kFrameThreads);
  if (c.d.order == 4)
    hipLaunchKernelGGL((k_frame_coefs<T, 4>), dim3(nb), dim3(kFrameThreads), 0, c.stream, op, fm, (T*)buf);
  else
    hipLaunchKernelGGL((k_frame_coefs<T, 2>), dim3(nb), dim3(kFrameThreads), 0, c.stream, op, fm, (T*)buf);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

__global__ __launch_bounds__(256) void k_round_coefs(const double* __restrict__ a, float* __restrict__ b,
                                                     long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// The FP32 coefficients are the FP64 ones rounded (closer to the operator
// than the same algebra in FP32 would be).
int ctx_build_frame(Ctx& c) {
  int e = build_frame_coefs<double>(c, c.op64, c.fcoef64);
  if (e || !c.fcoef64) return e;
  const long long n = (long long)kFrameCoefs * FrameMap::make(c.d.Ny, c.d.Nz).count();
  if (!c.fcoef32) SC_HIP(hipMalloc(&c.fcoef32, (size_t)n * sizeof(float)));
  hipLaunchKernelGGL(k_round_coefs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream,
                     (const double*)c.fcoef64, (float*)c.fcoef32, n);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

template <typename T>
static const T* frame_coefs(const Ctx& c) {
  if constexpr (sizeof(T) == 8) return (const T*)c.fcoef64;
  else return (const T*)c.fcoef32;
}

template <typename T, class EP>
static int launch_apply_t(const Ctx& c, const DevOp<T>& op, const T* in, const EP& ep, hipStream_t s) {
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  FrameMap fm = FrameMap::make(Ny, Nz);
  const long long nf = fm.count();
  if (Ny > 2 * kFrame && Nz > 2 * kFrame) {  // frame and interior in one launch
    const int nbx = (Nz - 2 * kFrame + kLanesOut - 1) / kLanesOut, nby = (Ny - 2 * kFrame + kJC - 1) / kJC;
    const int npb = 64 / kFrameQuad;  // frame nodes per one-wave block
    const int nfb = (int)((nf + npb * kXcd - 1) / (npb * kXcd)) * kXcd;
    const T* cf = frame_coefs<T>(c);
    const dim3 g((unsigned)(nfb + nbx * nby));
    if (c.d.order == 4)
      hipLaunchKernelGGL((k_apply_interior<T, 4, EP>), g, dim3(64), 0, s, op, in, ep, nbx, nby, fm, cf, nfb);
    else
      hipLaunchKernelGGL((k_apply_interior<T, 2, EP>), g, dim3(64), 0, s, op, in, ep, nbx, nby, fm, cf, nfb);
  } else if (nf > 0) {  // no interior block: every node is a frame node
    const unsigned nb = (unsigned)((nf + 255) / 256);
    hipLaunchKernelGGL((k_apply_frame<T, EP>), dim3(nb), dim3(256), 0, s, op, fm, in, ep);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int launch_apply64(const Ctx& c, const double* in, double* out, hipStream_t s) {
  return launch_apply_t(c, c.op64, in, StoreAu<double>{out}, s);
}
int launch_apply32(const Ctx& c, const float* in, float* out, hipStream_t s) {
  return launch_apply_t(c, c.op32, in, StoreAu<float>{out}, s);
}

template <typename T>
static int launch_wave_t(const Ctx& c, const DevOp<T>& op, const WaveArgs<T>& a, hipStream_t s) {
  if (a.rhoIsC1) {
    WaveUpdate<T, true> ep{a.uNext, a.uPrev, a.rho, a.ay, op.y.Hinv, op.z.Hinv, op.y.imet, op.z.imet,
                           a.d2uFault, a.dt, c.d.var_grid};
    if (a.fault) {
      if (a.fault->N != c.d.Nz) return fail(SC_ERR_ARG, "wave step: the fault has Nz nodes");
      return launch_apply_t(c, op, a.u, WaveFault<T>{ep, *a.fault}, s);
    }
    return launch_apply_t(c, op, a.u, ep, s);
  }
  if (a.fault) return fail(SC_ERR_ARG, "wave step: the fused fault step needs the per-node scale");
  WaveUpdate<T, false> ep{a.uNext, a.uPrev, a.rho, a.ay, op.y.Hinv, op.z.Hinv, op.y.imet, op.z.imet,
                          a.d2uFault, a.dt, c.d.var_grid};
  return launch_apply_t(c, op, a.u, ep, s);
}

// the wave update's per-node scale s = dt^2/rho * Hy^-1 Hz^-1 (* qy rz on a
// variable grid): the same factors WaveUpdate forms from rho, multiplied out once
template <typename T>
__global__ __launch_bounds__(256) void k_wave_c1(DevOp<T> p, const T* __restrict__ rho, T* __restrict__ c1,
                                                 T dt, int varGrid) {
  const size_t n = (size_t)p.Ny * p.Nz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(i / (size_t)p.Nz), k = (int)(i - (size_t)j * p.Nz);
    T h = p.y.Hinv[j] * p.z.Hinv[k];
    if (varGrid) h = h * (p.y.imet[j] * p.z.imet[k]);
    c1[i] = (dt * dt / rho[i]) * h;
  }
}
template <typename T>
static int launch_wave_c1_t(const Ctx& c, const DevOp<T>& op, const T* rho, T* c1, T dt, hipStream_t s) {
  const size_t b = ((size_t)c.N + 255) / 256;
  hipLaunchKernelGGL(k_wave_c1<T>, dim3((unsigned)(b < 4096 ? b : 4096)), dim3(256), 0, s, op, rho, c1, dt,
                     c.d.var_grid);
  SC_HIP(hipGetLastError());
  return SC_OK;
}
int launch_wave_c1(const Ctx& c, const double* rho, double* c1, double dt, hipStream_t s) {
  return launch_wave_c1_t(c, c.op64, rho, c1, dt, s);
}
int launch_wave_c1(const Ctx& c, const float* rho, float* c1, float dt, hipStream_t s) {
  return launch_wave_c1_t(c, c.op32, rho, c1, dt, s);
}
int launch_wave64(const Ctx& c, const WaveArgs<double>& a, hipStream_t s) {
  return launch_wave_t(c, c.op64, a, s);
}
int launch_wave32(const Ctx& c, const WaveArgs<float>& a, hipStream_t s) {
  return launch_wave_t(c, c.op32, a, s);
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

// whole vector: row j = blockIdx.y
__global__ __launch_bounds__(256) void k_set_rhs(DevOp<double> p, double* __restrict__ rhs,
                                                 const double* __restrict__ bcL,
                                                 const double* __restrict__ bcR,
                                                 const double* __restrict__ bcT,
                                                 const double* __restrict__ bcB) {
  const int j = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.Nz) return;
  rhs[(size_t)j * p.Nz + k] = rhs_node(p, j, k, [&](int side, int i) {
    return side == 0 ? bcL[i] : side == 1 ? bcR[i] : side == 2 ? bcT[i] : bcB[i];
  });
}

// the boundary frame only (the rest of rhs is known to be zero)
__global__ __launch_bounds__(256) void k_set_rhs_frame(DevOp<double> p, FrameMap fm, double* __restrict__ rhs,
                                                       const double* __restrict__ bcL,
                                                       const double* __restrict__ bcR,
                                                       const double* __restrict__ bcT,
                                                       const double* __restrict__ bcB) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fm.count()) return;
  int j, k;
  fm.node(t, j, k);
  rhs[(size_t)j * p.Nz + k] = rhs_node(p, j, k, [&](int side, int i) {
    return side == 0 ? bcL[i] : side == 1 ? bcR[i] : side == 2 ? bcT[i] : bcB[i];
  });
}

int launch_set_rhs(const Ctx& c, double* rhs, const double* bcL, const double* bcR,
                   const double* bcT, const double* bcB, hipStream_t s, bool frame_only) {
  if (frame_only) {
    const FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
    const long long n = fm.count();
    hipLaunchKernelGGL(k_set_rhs_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c.op64, fm,
                       rhs, bcL, bcR, bcT, bcB);
  } else {
    hipLaunchKernelGGL(k_set_rhs, dim3((unsigned)((c.d.Nz + 255) / 256), (unsigned)c.d.Ny), dim3(256), 0,
                       s, c.op64, rhs, bcL, bcR, bcT, bcB);
  }
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
    case 5: v = (p.y.Hinv[j] * p.z.Hinv[k]) * u[I]; break;
    // Dy (mu u), Dz (mu u): the product formed at each stencil point
    case 6: {
      v = p.y.imet[j] * d1_row(p.y, j, [&](int m) {
            const size_t i = (size_t)m * Nz + k;
            return p.mu[i] * u[i];
          });
      break;
    }
    default: {
      v = p.z.imet[k] * d1_row(p.z, k, [&](int m) {
            const size_t i = (size_t)j * Nz + m;
            return p.mu[i] * u[i];
          });
      break;
    }
  }
  out[I] = v;
}

// Hy^-1 e0y etc.: the 1D inverse norm is diagonal, so each operator keeps one
// face of the grid, scaled by that face's inverse weight, and zeroes the rest
__global__ __launch_bounds__(256) void k_bnd(DevOp<double> p, int which, const double* __restrict__ in,
                                             double* __restrict__ out) {
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ny = p.Ny, Nz = p.Nz;
  if (I >= (size_t)Ny * Nz) return;
  const int j = (int)(I / Nz), k = (int)(I % Nz);
  double v = 0.0;
  switch (which) {
    case 0: if (j == 0) v = p.y.Hinv[0] * in[k]; break;
    case 1: if (j == Ny - 1) v = p.y.Hinv[Ny - 1] * in[k]; break;
    case 2: if (j == 0) v = p.y.Hinv[0] * in[I]; break;
    case 3: if (j == Ny - 1) v = p.y.Hinv[Ny - 1] * in[I]; break;
    case 4: if (k == 0) v = p.z.Hinv[0] * in[I]; break;
    default: if (k == Nz - 1) v = p.z.Hinv[Nz - 1] * in[I]; break;
  }
  out[I] = v;
}

int launch_bnd(const Ctx& c, int which, const double* in, double* out, hipStream_t s) {
  const size_t N = (size_t)c.N;
  hipLaunchKernelGGL(k_bnd, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, c.op64, which, in, out);
  SC_HIP(hipGetLastError());
  return SC_OK;
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
  out[k] = p.mu[k] * dy_row0(p, u, k);
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
extern "C" int sc_sbp_dyxmu(sc_ctx* h, const double* in, double* out) { return deriv(h, 6, in, out); }
extern "C" int sc_sbp_dzxmu(sc_ctx* h, const double* in, double* out) { return deriv(h, 7, in, out); }

static int bnd(sc_ctx* h, int which, const double* in, double* out) {
  if (!h || !in || !out) return fail(SC_ERR_ARG, "null argument");
  if (in == out) return fail(SC_ERR_ARG, "in and out must differ");
  return sc::launch_bnd(h->c, which, in, out, h->c.stream);
}
extern "C" int sc_sbp_hyinvx_e0y(sc_ctx* h, const double* in, double* out) { return bnd(h, 0, in, out); }
extern "C" int sc_sbp_hyinvx_eNy(sc_ctx* h, const double* in, double* out) { return bnd(h, 1, in, out); }
extern "C" int sc_sbp_hyinvx_E0y(sc_ctx* h, const double* in, double* out) { return bnd(h, 2, in, out); }
extern "C" int sc_sbp_hyinvx_ENy(sc_ctx* h, const double* in, double* out) { return bnd(h, 3, in, out); }
extern "C" int sc_sbp_hzinvx_E0z(sc_ctx* h, const double* in, double* out) { return bnd(h, 4, in, out); }
extern "C" int sc_sbp_hzinvx_ENz(sc_ctx* h, const double* in, double* out) { return bnd(h, 5, in, out); }

extern "C" int sc_wave_step(sc_ctx* h, int prec, double dt, const void* u, const void* uPrev,
                            void* uNext, const void* rho, const void* ay, void* d2uFault) {
  if (!h || !u || !uPrev || !uNext || !rho || !ay || !d2uFault)
    return fail(SC_ERR_ARG, "sc_wave_step: null argument");
  if (u == uNext || uPrev == uNext) return fail(SC_ERR_ARG, "sc_wave_step: uNext must be a separate buffer");
  if (!(dt > 0)) return fail(SC_ERR_ARG, "sc_wave_step: dt must be > 0");
  const sc::Ctx& c = h->c;
  if (c.d.laplace != 0) return fail(SC_ERR_ARG, "sc_wave_step: needs laplace type yz");
  if (prec == SC_FP64)
    return sc::launch_wave64(c, {(const double*)u, (const double*)uPrev, (double*)uNext, (const double*)rho,
                                 (const double*)ay, (double*)d2uFault, dt}, c.stream);
  if (prec == SC_FP32)
    return sc::launch_wave32(c, {(const float*)u, (const float*)uPrev, (float*)uNext, (const float*)rho,
                                 (const float*)ay, (float*)d2uFault, (float)dt}, c.stream);
  return fail(SC_ERR_ARG, "sc_wave_step: prec must be SC_FP64 or SC_FP32");
}

namespace sc {
// single-kernel launches for timing: interior kernel with a given tile map, or
// the frame kernel alone
template <typename T, int XCD>
static void bench_interior(const Ctx& c, const DevOp<T>& op, const T* a, T* b) {
  const int niy = c.d.Ny - 2 * kFrame, niz = c.d.Nz - 2 * kFrame;
  const int nbx = (niz + kLanesOut - 1) / kLanesOut, nby = (niy + kJC - 1) / kJC;
  if (c.d.order == 4)
    hipLaunchKernelGGL((k_apply_interior<T, 4, StoreAu<T>, XCD>), dim3(nbx * nby), dim3(64), 0, c.stream,
                       op, a, StoreAu<T>{b}, nbx, nby, FrameMap{}, (const T*)nullptr, 0);
  else
    hipLaunchKernelGGL((k_apply_interior<T, 2, StoreAu<T>, XCD>), dim3(nbx * nby), dim3(64), 0, c.stream,
                       op, a, StoreAu<T>{b}, nbx, nby, FrameMap{}, (const T*)nullptr, 0);
}
template <typename T>
static void bench_frame(const Ctx& c, const DevOp<T>& op, const T* a, T* b) {
  FrameMap fm = FrameMap::make(c.d.Ny, c.d.Nz);
  const unsigned nb = (unsigned)((fm.count() * kFrameQuad + 255) / 256);
  hipLaunchKernelGGL((k_apply_frame_cf<T, StoreAu<T>>), dim3(nb), dim3(256), 0, c.stream, op, fm,
                     frame_coefs<T>(c), a, StoreAu<T>{b});
}
}  // namespace sc

// Times `reps` back-to-back launches on the context stream with HIP events
// (the stream the kernels run on); returns the mean time per launch in ms and
// the number of nodes the timed kernel(s) write.  which: 0 = FP64 apply,
// 1 = FP32 apply, 2 / 3 = FP64 / FP32 interior kernel only, 4 / 5 = the same
// with the plain (not XCD-grouped) tile order, 6 / 7 = FP64 / FP32 frame
// kernel only.
extern "C" int sc_bench_apply(sc_ctx* h, int which, int reps, double* ms_per_launch,
                              double* nodes) {
  if (!h || reps <= 0 || !ms_per_launch || which < 0 || which > 13)
    return fail(SC_ERR_ARG, "sc_bench_apply: bad argument");
  if (which >= 9 && h->c.d.laplace != 0) return fail(SC_ERR_ARG, "sc_bench_apply: the wave step needs laplace type yz");
  if (which == 8) {
    double words = 0;
    int e = sc::bench_proj_guess(h->c, reps, ms_per_launch, &words);
    if (nodes) *nodes = words;
    return e;
  }
  sc::Ctx& c = h->c;
  if (which >= 11) {  // the fast-diagonalisation preconditioner, r -> z
    const int prec = which == 11 ? SC_PREC_FDIAG64 : (which == 12 ? SC_PREC_FDIAG32 : SC_PREC_FDIAG_BF16);
    const size_t N = (size_t)c.N;
    double *r = nullptr, *z = nullptr;
    SC_HIP(hipMalloc(&r, N * sizeof(double)));
    SC_HIP(hipMalloc(&z, N * sizeof(double)));
    std::vector<double> rh(N);
    for (size_t i = 0; i < N; ++i) rh[i] = 1.0 + 1e-3 * (double)(i % 97);
    SC_HIP(hipMemcpy(r, rh.data(), N * sizeof(double), hipMemcpyHostToDevice));
    int err = 0;
    for (int w = 0; w < 3 && !err; ++w) err = sc::precond_apply(c, prec, r, z);
    hipEvent_t e0, e1;
    SC_HIP(hipEventCreate(&e0));
    SC_HIP(hipEventCreate(&e1));
    SC_HIP(hipEventRecord(e0, c.stream));
    for (int q = 0; q < reps && !err; ++q) err = sc::precond_apply(c, prec, r, z);
    SC_HIP(hipEventRecord(e1, c.stream));
    SC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SC_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_launch = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(r);
    (void)hipFree(z);
    // four GEMMs: two Nz x Nz x Ny, two Nz x Ny x Ny
    if (nodes) *nodes = 4.0 * (double)c.d.Nz * c.d.Ny * ((double)c.d.Ny + c.d.Nz);
    return err;
  }
  const bool f64 = which == 9 || (which < 9 && which % 2 == 0);
  const size_t N = (size_t)c.N, esz = f64 ? 8 : 4;
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  const int niy = Ny - 2 * sc::kFrame, niz = Nz - 2 * sc::kFrame;
  if (which >= 2 && (niy <= 0 || niz <= 0)) return fail(SC_ERR_ARG, "grid has no interior block");
  if (nodes) {
    if (which < 2 || which >= 9) *nodes = (double)N;
    else if (which < 6) *nodes = (double)niy * niz;
    else *nodes = (double)N - (double)niy * niz;
  }
  // apply: a -> b; wave step (9 FP64, 10 FP32): u = a, uPrev = b, uNext = w,
  // rho = 1 in r, ay = 0 in y (zero u: the timing does not depend on values)
  void *a = nullptr, *b = nullptr, *w = nullptr, *r = nullptr, *y = nullptr, *d2 = nullptr;
  SC_HIP(hipMalloc(&a, N * esz));
  SC_HIP(hipMalloc(&b, N * esz));
  SC_HIP(hipMemsetAsync(a, 0, N * esz, c.stream));
  if (which >= 9) {
    SC_HIP(hipMalloc(&w, N * esz));
    SC_HIP(hipMalloc(&r, N * esz));
    SC_HIP(hipMalloc(&y, N * esz));
    SC_HIP(hipMalloc(&d2, (size_t)Nz * esz));
    SC_HIP(hipMemsetAsync(b, 0, N * esz, c.stream));
    SC_HIP(hipMemsetAsync(y, 0, N * esz, c.stream));
    if (f64) {
      std::vector<double> one(N, 1.0);
      SC_HIP(hipMemcpy(r, one.data(), N * esz, hipMemcpyHostToDevice));
    } else {
      std::vector<float> one(N, 1.0f);
      SC_HIP(hipMemcpy(r, one.data(), N * esz, hipMemcpyHostToDevice));
    }
  }
  hipEvent_t e0, e1;
  SC_HIP(hipEventCreate(&e0));
  SC_HIP(hipEventCreate(&e1));
  auto launch = [&]() -> int {
    const double *a64 = (const double*)a;
    const float* a32 = (const float*)a;
    switch (which) {
      case 0: return sc::launch_apply64(c, a64, (double*)b, c.stream);
      case 1: return sc::launch_apply32(c, a32, (float*)b, c.stream);
      case 2: sc::bench_interior<double, 1>(c, c.op64, a64, (double*)b); return 0;
      case 3: sc::bench_interior<float, 1>(c, c.op32, a32, (float*)b); return 0;
      case 4: sc::bench_interior<double, 0>(c, c.op64, a64, (double*)b); return 0;
      case 5: sc::bench_interior<float, 0>(c, c.op32, a32, (float*)b); return 0;
      case 6: sc::bench_frame<double>(c, c.op64, a64, (double*)b); return 0;
      case 7: sc::bench_frame<float>(c, c.op32, a32, (float*)b); return 0;
      case 9:  // the FD driver's form: r holds the per-node scale
        return sc::launch_wave64(c, {a64, (const double*)b, (double*)w, (const double*)r, (const double*)y,
                                     (double*)d2, 1e-3, 1}, c.stream);
      default:
        return sc::launch_wave32(c, {a32, (const float*)b, (float*)w, (const float*)r, (const float*)y,
                                     (float*)d2, 1e-3f, 1}, c.stream);
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
  for (void* p : {a, b, w, r, y, d2})
    if (p) (void)hipFree(p);
  return err;
}
