// Fast-diagonalisation preconditioner: set-up (1-D operators, generalised
// eigen-decompositions with rocSOLVER) and application (four GEMMs on the
// matrix cores, this file's own MFMA kernels).
#include <hip/hip_bf16.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "fdiag.hpp"

namespace sc {

namespace {

inline double band(const std::vector<double>& b, int n, int i, int c) {
  const int o = c - i + 4;
  if (i < 0 || i >= n || c < 0 || c >= n || o < 0 || o >= kBand) return 0.0;
  return b[(size_t)i * kBand + o];
}

// Dense matrix (row-major n x n) of the 1-D SBP operator along one direction
// with coefficient c (and its two-point average c3), including the SAT terms
// of the low and high sides, without the prefactor of the other direction.
// Same algebra as the kernels' flux form (line_y / line_z, sat_terms).
std::vector<double> dense_line(const Dir1D& d, const std::vector<double>& c,
                               const std::vector<double>& c3, int bcLow, int bcHigh) {
  const int n = d.n;
  std::vector<double> M((size_t)n * n, 0.0);
  std::vector<double> g1(n), g3(n), g4(n);
  for (int jc = 0; jc < n; ++jc) {
    const int m0 = std::max(0, jc - 4), m1 = std::min(n - 1, jc + 4);
    for (int m = m0; m <= m1; ++m) {
      g1[m] = c[m] * band(d.D1, n, m, jc);
      g3[m] = d.wR[m] * c3[m] * band(d.DR, n, m, jc);
      g4[m] = d.w4[m] * c[m] * band(d.D4, n, m, jc);
    }
    const int i0 = std::max(0, jc - 8), i1 = std::min(n - 1, jc + 8);
    for (int i = i0; i <= i1; ++i) {
      double a1 = 0, aR = 0, a4 = 0;
      for (int m = std::max(m0, i - 4); m <= std::min(m1, i + 4); ++m) {
        a1 += band(d.D1, n, i, m) * g1[m];
        aR += band(d.DR, n, m, i) * g3[m];
        a4 += band(d.D4, n, m, i) * g4[m];
      }
      M[(size_t)i * n + jc] = d.H[i] * a1 - (aR + a4);
    }
  }
  if (bcLow == SC_DIRICHLET) {
    for (int i = 0; i < std::min(n, kBand); ++i)
      M[(size_t)i * n + 0] += (d.BS0[i] + (i == 0 ? d.alphaD : 0.0)) * c[0];
  } else {
    for (int m = 0; m < std::min(n, kBand); ++m) M[m] += c[0] * band(d.D1, n, 0, m);
  }
  if (bcHigh == SC_DIRICHLET) {
    for (int i = std::max(0, n - kBand); i < n; ++i)
      M[(size_t)i * n + (n - 1)] += (d.BSN[i - (n - kBand)] + (i == n - 1 ? d.alphaD : 0.0)) * c[n - 1];
  } else {
    for (int m = std::max(0, n - kBand); m < n; ++m)
      M[(size_t)(n - 1) * n + m] -= c[n - 1] * band(d.D1, n, n - 1, m);
  }
  return M;
}

// Symmetric eigen-decomposition S = Q diag(w) Q^T on the device with
// rocSOLVER's divide-and-conquer driver (dsyevd), once per operator and BC
// set.  On return w is ascending and vt[i*n + k] is component k of
// eigenvector i (rocSOLVER's column-major Q read row-wise).
int sym_eig(rocblas_handle hb, hipStream_t s, int n, const std::vector<double>& S,
            std::vector<double>& w, std::vector<double>& vt) {
  const size_t nn = (size_t)n * n;
  double *a_d = nullptr, *d_d = nullptr, *e_d = nullptr;
  rocblas_int* info_d = nullptr;
  int rc = SC_OK;
  rocblas_int info = 0;
  rocblas_set_stream(hb, s);
  if (hipMalloc(&a_d, nn * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_d, (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc(&e_d, (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc(&info_d, sizeof(rocblas_int)) != hipSuccess) {
    rc = fail(SC_ERR_HIP, "eigensolver workspace allocation failed");
  } else if (hipMemcpyAsync(a_d, S.data(), nn * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) {
    rc = fail(SC_ERR_HIP, "eigensolver upload failed");
  } else {
    const rocblas_status st = rocsolver_dsyevd(hb, rocblas_evect_original, rocblas_fill_upper, n, a_d,
                                               n, d_d, e_d, info_d);
    if (st != rocblas_status_success) {
      rc = fail(SC_ERR_HIP, std::string("rocsolver_dsyevd -> ") + rocblas_status_to_string(st));
    } else {
      w.resize(n);
      vt.resize(nn);
      if (hipMemcpyAsync(w.data(), d_d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(vt.data(), a_d, nn * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        rc = fail(SC_ERR_HIP, "eigensolver read-back failed");
      else if (info != 0)
        rc = fail(SC_ERR_NOCONV, "rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")");
    }
  }
  for (void* p : {(void*)a_d, (void*)d_d, (void*)e_d, (void*)info_d})
    if (p) (void)hipFree(p);
  return rc;
}

// Generalised symmetric eigenproblem L V = D V Lambda with V^T D V = I
// (D diagonal, positive), through S = D^-1/2 L D^-1/2 = Q Lambda Q^T and
// V = D^-1/2 Q.  Returns the column-major V (n x n) in vcm and eigenvalues in w.
int gen_eig(rocblas_handle hb, hipStream_t s, int n, std::vector<double> L,
            const std::vector<double>& D, std::vector<double>& w, std::vector<double>& vcm,
            double* asym) {
  double amax = 0, dmax = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      amax = std::max(amax, std::fabs(L[(size_t)i * n + j]));
      dmax = std::max(dmax, std::fabs(L[(size_t)i * n + j] - L[(size_t)j * n + i]));
    }
  *asym = amax > 0 ? dmax / amax : 0;
  std::vector<double> dis(n);
  for (int i = 0; i < n; ++i) dis[i] = 1.0 / std::sqrt(D[i]);
  std::vector<double> S((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      S[(size_t)i * n + j] = 0.5 * (L[(size_t)i * n + j] + L[(size_t)j * n + i]) * dis[i] * dis[j];
  std::vector<double> vt;
  if (int e = sym_eig(hb, s, n, S, w, vt)) return e;
  // V = D^{-1/2} Q; column i of V is row i of vt scaled
  vcm.resize((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) vcm[(size_t)i * n + k] = vt[(size_t)i * n + k] * dis[k];
  return 0;
}

// ----------------------------------------------------------- MFMA GEMMs --
// C = A B on the matrix cores, column-major operands (the preconditioner's
// vectors are z-fast, i.e. column-major Nz x Ny):
//   A  M x K (lda), zero-padded factors or intermediates: read unchecked;
//   B  K x N (ldb): rows >= kB or columns >= nB read as zero (the caller's
//      unpadded vector), converted to the GEMM's input type on load;
//   C  M x N (ldc): rows >= mC or columns >= nC are not stored (the caller's
//      unpadded vector), converted from the accumulator on store;
//   S  optional M x N scale in the accumulator type: C = (A B) .* S.
// M, N, K are multiples of kFdTile.  FP64 (v_mfma_f64_16x16x4f64) and FP32
// (v_mfma_f32_16x16x4f32): a block owns a BM x BN tile of C; its waves form
// KS groups of WM x WN waves, each group multiplying its share of every K
// slab (split K inside the block, summed through LDS at the end), so that
// a grid with one tile per CU still runs two or more waves per SIMD.  K
// runs in slabs of BK through two LDS buffers: after each barrier the next
// slab is stored and the one after it loaded before the current slab is
// multiplied.  LDS layout: A as [k][m] (rows padded so the four 16-lane
// row groups of a fragment read fall on different banks), B as [n][k] in
// its global k-fast order (pitch BK + 16 bytes: conflict-free fragment
// reads and aligned 16-byte stores); 16-byte global loads wherever the
// operand is padded.  bf16 (v_mfma_f32_16x16x16bf16_1k, four consecutive k
// a lane): 64 x 64 tiles, A and B both k-fast in LDS.
typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

template <int KIND>
struct Mfma;
template <>
struct Mfma<0> {
  using T = double;
  using Acc = double;
  using V = d4v;
  __device__ static V op(T a, T b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // row of accumulator r of lane group g inside a 16 x 16 tile (interleaved
  // for the FP64 instruction)
  __device__ static int row(int g, int r) { return g + 4 * r; }
};
template <>
struct Mfma<1> {
  using T = float;
  using Acc = float;
  using V = f4v;
  __device__ static V op(T a, T b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int r) { return 4 * g + r; }
};
template <>
struct Mfma<2> {
  using T = unsigned short;  // bf16 bits
  using Acc = float;
  using V = f4v;
  __device__ static int row(int g, int r) { return 4 * g + r; }
};

// round to nearest even FP32 -> bf16 (NaN kept quiet)
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

template <int KIND, typename TB>
__device__ __forceinline__ typename Mfma<KIND>::T to_op(TB x) {
  if constexpr (KIND == 0) return (double)x;
  else if constexpr (KIND == 1) return (float)x;
  else if constexpr (std::is_same<TB, unsigned short>::value) return x;
  else return bf16_bits((float)x);
}
template <typename TC, typename Acc>
__device__ __forceinline__ TC from_acc(Acc x) {
  if constexpr (std::is_same<TC, unsigned short>::value) return bf16_bits((float)x);
  else return (TC)x;
}

template <int KIND, typename TB, typename TC>
struct GemmArgs {
  const typename Mfma<KIND>::T* A;
  int lda;
  const TB* B;
  int ldb, kB, nB;
  TC* C;
  int ldc, mC, nC;
  const typename Mfma<KIND>::Acc* S;
  int lds;
  int M, N, K;
};

// the tile configurations (measured at 448^2 and 1024^2 on MI355X)
template <int BM_, int BN_, int BK_, int WM_, int WN_, int KS_>
struct GemmCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_, KS = KS_;
  static constexpr int NT = 64 * WM * WN * KS;
};
using CfgF64Large = GemmCfg<64, 64, 32, 2, 4, 1>;
using CfgF64Small = GemmCfg<32, 32, 32, 2, 2, 1>;
using CfgF32Large = GemmCfg<64, 64, 32, 2, 2, 2>;
using CfgF32Small = GemmCfg<32, 32, 32, 2, 2, 2>;

// the block's tile: each XCD (block id mod 8) takes a contiguous run of
// tiles, M fastest
__device__ __forceinline__ void gemm_tile(int M, int N, int BM, int BN, int* m0, int* n0) {
  const int tilesM = M / BM, nt = tilesM * (N / BN);
  int b = blockIdx.x;
  if ((nt & 7) == 0) b = (b & 7) * (nt >> 3) + (b >> 3);
  *m0 = (b % tilesM) * BM;
  *n0 = (b / tilesM) * BN;
}

// FP64 / FP32.  VECB: 0 for the caller's unpadded vector as B (element
// loads, bounds-checked), 1 for a padded B (loads of up to 16 bytes)
template <int KIND, class CF, int VECB, typename TB, typename TC>
__global__ __launch_bounds__(CF::NT) void k_fd_gemm(GemmArgs<KIND, TB, TC> g) {
  using P = Mfma<KIND>;
  using T = typename P::T;
  using V = typename P::V;
  using Acc = typename P::Acc;
  constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, WM = CF::WM, WN = CF::WN, KS = CF::KS, NT = CF::NT;
  constexpr int V16 = 16 / (int)sizeof(T);    // elements of a 16-byte load
  // elements a thread loads at once: up to 16 bytes, at most its share of a slab
  constexpr int VE = BM * BK / NT < V16 ? BM * BK / NT : V16;
  constexpr int VB = VECB ? (BN * BK / NT < V16 ? BN * BK / NT : V16) : 1;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int PMN = BM + 16;                 // [k][m] rows of A
  constexpr int PNK = BK + V16;                // [n][k] rows of B
  constexpr int ASZ = BK * PMN, BSZ = BN * PNK;
  static_assert(BM * BN <= 2 * (ASZ + BSZ), "the k-groups' reduction fits in the slab buffers");
  __shared__ __attribute__((aligned(16))) T smem[2 * (ASZ + BSZ)];
  T* const As0 = smem;
  T* const Bs0 = smem + 2 * ASZ;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int m0, n0;
  gemm_tile(g.M, g.N, BM, BN, &m0, &n0);
  // staging: A VE consecutive m a thread; B VB consecutive k a thread
  constexpr int LA = BM * BK / NT / VE, LB = BN * BK / NT / VB;
  static_assert(LA >= 1 && LB >= 1 && BM * BK % (NT * VE) == 0 && BN * BK % (NT * VB) == 0, "staging");
  constexpr int AT = BM / VE, BT = BK / VB;
  typedef T TA __attribute__((ext_vector_type(VE)));
  typedef T TBV __attribute__((ext_vector_type(VB)));
  const int am = (tid % AT) * VE, ak = tid / AT;
  const int bk = (tid % BT) * VB, bn = tid / BT;
  TA ra[2][LA];
  TBV rb[2][LB];
  const T* pa = g.A + (size_t)ak * g.lda + m0 + am;
  auto gload = [&](TA (&xa)[LA], TBV (&xb)[LB], int k0) {
    const T* qa = pa + (size_t)k0 * g.lda;
#pragma unroll
    for (int i = 0; i < LA; ++i) xa[i] = *reinterpret_cast<const TA*>(qa + (size_t)i * (NT / AT) * g.lda);
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int n = n0 + bn + (NT / BT) * i, k = k0 + bk;
      if constexpr (!VECB) {
        const bool ok = k < g.kB && n < g.nB;
        const TB v = g.B[(size_t)(ok ? n : 0) * g.ldb + (ok ? k : 0)];  // unconditional, clamped
        xb[i] = ok ? to_op<KIND>(v) : to_op<KIND>((TB)0);
      } else {
        static_assert(std::is_same<TB, T>::value, "16-byte B loads take the operand type");
        xb[i] = *reinterpret_cast<const TBV*>(g.B + (size_t)n * g.ldb + k);
      }
    }
  };
  auto sstore = [&](const TA (&xa)[LA], const TBV (&xb)[LB], int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i)
      *reinterpret_cast<TA*>(As0 + buf * ASZ + (ak + (NT / AT) * i) * PMN + am) = xa[i];
#pragma unroll
    for (int i = 0; i < LB; ++i)
      *reinterpret_cast<TBV*>(Bs0 + buf * BSZ + (bn + (NT / BT) * i) * PNK + bk) = xb[i];
  };
  V acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = V{0, 0, 0, 0};
  const int kg = wid / (WM * WN), wi = wid % (WM * WN);
  const int wm = (wi % WM) * (BM / WM), wn = (wi / WM) * (BN / WN);
  const int lr = lane & 15, lk = lane >> 4;
  constexpr int KPG = BK / KS;
  auto compute = [&](int buf) {
    const T* as = As0 + buf * ASZ;
    const T* bs = Bs0 + buf * BSZ;
#pragma unroll
    for (int kk = 0; kk < KPG; kk += 4) {
      const int k = kg * KPG + kk + lk;
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = as[k * PMN + wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = bs[(wn + 16 * j + lr) * PNK + k];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = P::op(a[i], b[j], acc[i][j]);
    }
  };
  const int S = g.K / BK;
  gload(ra[0], rb[0], 0);
  if (S > 1) gload(ra[1], rb[1], BK);
  sstore(ra[0], rb[0], 0);
  for (int s0 = 0; s0 < S; s0 += 2) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int s = s0 + p;
      if (s < S) {
        __syncthreads();
        if (s + 1 < S) sstore(ra[p ^ 1], rb[p ^ 1], (s + 1) & 1);
        if (s + 2 < S) gload(ra[p], rb[p], (s + 2) * BK);
        compute(s & 1);
      }
    }
  }
  // the k-groups' partial tiles summed through LDS into group 0
  if constexpr (KS > 1) {
    Acc* red = reinterpret_cast<Acc*>(smem);
    for (int q = 1; q < KS; ++q) {
      __syncthreads();
      if (kg == q) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              red[(wn + 16 * j + lr) * BM + wm + 16 * i + P::row(lk, r)] = acc[i][j][r];
      }
      __syncthreads();
      if (kg == 0) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] += red[(wn + 16 * j + lr) * BM + wm + 16 * i + P::row(lk, r)];
      }
    }
    if (kg != 0) return;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn + 16 * j + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + P::row(lk, r);
        Acc v = acc[i][j][r];
        if (g.S) v *= g.S[(size_t)n * g.lds + m];
        if (m < g.mC && n < g.nC) g.C[(size_t)n * g.ldc + m] = from_acc<TC>(v);
      }
    }
}

// bf16 inputs, FP32 accumulation.  Both operands arrive k-fast (A as the
// transpose of its column-major form, lda its row length; B column-major),
// so both go to LDS in their global order with 16-byte loads and stores:
// A as [m][k], B as [n][k] (rows of BK + 8, conflict-free 8-byte fragment
// reads).  64 x 64 tiles, four waves of 32 x 32, K in slabs of 64.  TRC:
// C is stored transposed (row-major, ldc its row length), as the k-fast A
// of the next GEMM.
template <bool TRC, typename TC>
__global__ __launch_bounds__(256) void k_fd_gemm_bf16(GemmArgs<2, unsigned short, TC> g) {
  using T = unsigned short;
  using V = f4v;
  constexpr int BM = 64, BN = 64, BK = 64, PK = BK + 8, ASZ = BM * PK;
  typedef T T8 __attribute__((ext_vector_type(8)));
  __shared__ __attribute__((aligned(16))) T As[2][ASZ];
  __shared__ __attribute__((aligned(16))) T Bs[2][ASZ];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int m0, n0;
  gemm_tile(g.M, g.N, BM, BN, &m0, &n0);
  // a thread moves 8 consecutive k of two rows of A and two columns of B
  const int kq = (tid & 7) * 8, rq = tid >> 3;  // rq in [0, 32)
  T8 ra[2][2], rb[2][2];
  auto gload = [&](T8 (&xa)[2], T8 (&xb)[2], int k0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      xa[i] = *reinterpret_cast<const T8*>(g.A + (size_t)(m0 + rq + 32 * i) * g.lda + k0 + kq);
      xb[i] = *reinterpret_cast<const T8*>(g.B + (size_t)(n0 + rq + 32 * i) * g.ldb + k0 + kq);
    }
  };
  auto sstore = [&](const T8 (&xa)[2], const T8 (&xb)[2], int buf) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      *reinterpret_cast<T8*>(&As[buf][(rq + 32 * i) * PK + kq]) = xa[i];
      *reinterpret_cast<T8*>(&Bs[buf][(rq + 32 * i) * PK + kq]) = xb[i];
    }
  };
  V acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = V{0, 0, 0, 0};
  const int wm = (wid & 1) * 32, wn = (wid >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int k = kk + 4 * lk;
      const s4v a0 = *reinterpret_cast<const s4v*>(&As[buf][(wm + lr) * PK + k]);
      const s4v a1 = *reinterpret_cast<const s4v*>(&As[buf][(wm + 16 + lr) * PK + k]);
      const s4v b0 = *reinterpret_cast<const s4v*>(&Bs[buf][(wn + lr) * PK + k]);
      const s4v b1 = *reinterpret_cast<const s4v*>(&Bs[buf][(wn + 16 + lr) * PK + k]);
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc[1][1], 0, 0, 0);
    }
  };
  const int S = g.K / BK;
  gload(ra[0], rb[0], 0);
  if (S > 1) gload(ra[1], rb[1], BK);
  sstore(ra[0], rb[0], 0);
  for (int s0 = 0; s0 < S; s0 += 2) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int s = s0 + p;
      if (s < S) {
        __syncthreads();
        if (s + 1 < S) sstore(ra[p ^ 1], rb[p ^ 1], (s + 1) & 1);
        if (s + 2 < S) gload(ra[p], rb[p], (s + 2) * BK);
        compute(s & 1);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + Mfma<2>::row(lk, r);
        float v = acc[i][j][r];
        if (g.S) v *= g.S[(size_t)n * g.lds + m];
        if (m < g.mC && n < g.nC) g.C[TRC ? (size_t)m * g.ldc + n : (size_t)n * g.ldc + m] = from_acc<TC>(v);
      }
    }
}

// the vector r (Nz x Ny, z fast) into the padded Pz x Py operand layout of
// the first GEMM, converted to its type, zero in the padding
template <int KIND, typename TI>
__global__ __launch_bounds__(256) void k_fd_pad(const TI* __restrict__ r, int Nz, int Ny, int Pz, int Py,
                                                typename Mfma<KIND>::T* __restrict__ out) {
  const size_t n = (size_t)Pz * Py;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % Pz), j = (int)(i / Pz);
    out[i] = (k < Nz && j < Ny) ? to_op<KIND>(r[(size_t)j * Nz + k]) : to_op<KIND>((TI)0);
  }
}

// one GEMM with the configuration for its size: 64 x 64 tiles when the
// grid has one per CU or more, 32 x 32 tiles below
template <int KIND, int VECB, typename TB, typename TC>
void gemm(hipStream_t s, const typename Mfma<KIND>::T* A, int lda, const TB* B, int ldb, int kB, int nB,
          TC* C, int ldc, int mC, int nC, const typename Mfma<KIND>::Acc* S, int lds, int M, int N, int K) {
  GemmArgs<KIND, TB, TC> g{A, lda, B, ldb, kB, nB, C, ldc, mC, nC, S, lds, M, N, K};
  if constexpr (KIND == 2) {
    static_assert(KIND != 2, "bf16 GEMMs go through gemm_bf16");
  } else {
    using Large = std::conditional_t<KIND == 0, CfgF64Large, CfgF32Large>;
    using Small = std::conditional_t<KIND == 0, CfgF64Small, CfgF32Small>;
    if ((M / 64) * (N / 64) >= 200) {
      const unsigned nb = (unsigned)((M / Large::BM) * (N / Large::BN));
      hipLaunchKernelGGL((k_fd_gemm<KIND, Large, VECB, TB, TC>), dim3(nb), dim3(Large::NT), 0, s, g);
    } else {
      const unsigned nb = (unsigned)((M / Small::BM) * (N / Small::BN));
      hipLaunchKernelGGL((k_fd_gemm<KIND, Small, VECB, TB, TC>), dim3(nb), dim3(Small::NT), 0, s, g);
    }
  }
}

template <bool TRC, typename TC>
void gemm_bf16(hipStream_t s, const unsigned short* A, int lda, const unsigned short* B, int ldb, TC* C, int ldc,
               int mC, int nC, const float* S, int lds, int M, int N, int K) {
  GemmArgs<2, unsigned short, TC> g{A, lda, B, ldb, K, N, C, ldc, mC, nC, S, lds, M, N, K};
  hipLaunchKernelGGL((k_fd_gemm_bf16<TRC, TC>), dim3((unsigned)((M / 64) * (N / 64))), dim3(256), 0, s, g);
}

unsigned pad_blocks(const FdDiag& f) {
  const size_t b = ((size_t)f.Pz * f.Py + 255) / 256;
  return (unsigned)(b < 4096 ? b : 4096);
}

// z = W (inv .* (W^T R V)) V^T with R the z-fast vector r (column-major
// Nz x Ny): r padded into t, then four GEMMs, the scale in the second one's
// store.  TI / TO: the vectors' types; the intermediates are in the GEMM's
// operand type.
template <int KIND, typename TI, typename TO>
void fd_transform(hipStream_t s, const FdDiag& f, const typename Mfma<KIND>::T* V,
                  const typename Mfma<KIND>::T* VT, const typename Mfma<KIND>::T* W,
                  const typename Mfma<KIND>::T* WT, const typename Mfma<KIND>::Acc* inv,
                  typename Mfma<KIND>::T* g, typename Mfma<KIND>::T* t, const TI* r, TO* z) {
  using TM = typename Mfma<KIND>::T;
  using Acc = typename Mfma<KIND>::Acc;
  const int Ny = f.Ny, Nz = f.Nz, Py = f.Py, Pz = f.Pz;
  hipLaunchKernelGGL((k_fd_pad<KIND, TI>), dim3(pad_blocks(f)), dim3(256), 0, s, r, Nz, Ny, Pz, Py, t);
  if constexpr (KIND == 2) {
    // k-fast operands: W^T's rows are W's columns, and the first and third
    // products are stored transposed for the GEMM after them
    gemm_bf16<true, TM>(s, W, Pz, t, Pz, g, Py, Pz, Py, (const Acc*)nullptr, 0, Pz, Py, Pz);
    gemm_bf16<false, TM>(s, g, Py, V, Py, t, Pz, Pz, Py, inv, Pz, Pz, Py, Py);
    gemm_bf16<true, TM>(s, WT, Pz, t, Pz, g, Py, Pz, Py, (const Acc*)nullptr, 0, Pz, Py, Pz);
    gemm_bf16<false, TO>(s, g, Py, VT, Py, z, Nz, Nz, Ny, (const Acc*)nullptr, 0, Pz, Py, Py);
  } else {
    gemm<KIND, 1, TM, TM>(s, WT, Pz, t, Pz, Pz, Py, g, Pz, Pz, Py, (const Acc*)nullptr, 0, Pz, Py, Pz);
    gemm<KIND, 1, TM, TM>(s, g, Pz, V, Py, Py, Py, t, Pz, Pz, Py, inv, Pz, Pz, Py, Py);
    gemm<KIND, 1, TM, TM>(s, W, Pz, t, Pz, Pz, Py, g, Pz, Pz, Py, (const Acc*)nullptr, 0, Pz, Py, Pz);
    gemm<KIND, 1, TM, TO>(s, g, Pz, VT, Py, Py, Py, z, Nz, Nz, Ny, (const Acc*)nullptr, 0, Pz, Py, Py);
  }
}

}  // namespace

FdDiag::~FdDiag() {
  void* bufs[] = {V64, W64, VT64, WT64, inv64, V32, W32, VT32, WT32, inv32, Vb, Wb, VTb, WTb,
                  g64, t64, g32, t32, gb, tb};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (blas) rocblas_destroy_handle((rocblas_handle)blas);
}

int fdiag_build(Ctx& c, FdDiag& f) {
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  if (c.d.laplace != 0) return fail(SC_ERR_ARG, "fast-diagonalisation preconditioner needs laplace type yz");
  const int Py = (Ny + kFdTile - 1) / kFdTile * kFdTile, Pz = (Nz + kFdTile - 1) / kFdTile * kFdTile;
  const bool resize = f.Py != Py || f.Pz != Pz;
  f.Ny = Ny;
  f.Nz = Nz;
  // depth profile of mu (column mean over y when mu is not a function of z alone)
  std::vector<double> mz(Nz, 0.0);
  f.muSeparable = true;
  for (int k = 0; k < Nz; ++k) {
    double s = 0;
    for (int j = 0; j < Ny; ++j) {
      const double v = c.mu_h[(size_t)j * Nz + k];
      s += v;
      if (v != c.mu_h[k]) f.muSeparable = false;
    }
    mz[k] = s / Ny;
  }
  // y: mu = 1, c = qy ; prefactor of the y-part Hz*zr*mu_k
  std::vector<double> cy(c.qy), c3y(c.qy);
  std::vector<double> Ly = dense_line(c.ty, cy, c3y, c.d.bcL, c.d.bcR);
  std::vector<double> Dy(Ny), Dz(Nz);
  for (int j = 0; j < Ny; ++j) Dy[j] = c.ty.H[j] * c.yq[j];
  // z: c = mu_k rz_k, c3 the two-point average ; prefactor Hy*yq
  std::vector<double> cz(Nz), c3z(Nz);
  for (int k = 0; k < Nz; ++k) cz[k] = mz[k] * c.rz[k];
  for (int k = 0; k < Nz; ++k) c3z[k] = k + 1 < Nz ? 0.5 * (cz[k] + cz[k + 1]) : cz[k];
  std::vector<double> Lz = dense_line(c.tz, cz, c3z, c.d.bcT, c.d.bcB);
  for (int k = 0; k < Nz; ++k) Dz[k] = c.tz.H[k] * c.zr[k] * mz[k];
  // A_sep = Ly (x) diag(Dz) + diag(Dy) (x) Lz ; A_sep is negative definite, so
  // decompose -L to work with positive eigenvalues and flip the sign back.
  for (auto& x : Ly) x = -x;
  for (auto& x : Lz) x = -x;
  if (!f.blas) {
    rocblas_handle hb;
    if (rocblas_create_handle(&hb) != rocblas_status_success) return fail(SC_ERR_HIP, "rocblas_create_handle failed");
    f.blas = hb;
  }
  std::vector<double> V, W;
  double a1 = 0, a2 = 0;
  if (int e = gen_eig((rocblas_handle)f.blas, c.stream, Ny, std::move(Ly), Dy, f.lam, V, &a1)) return e;
  if (int e = gen_eig((rocblas_handle)f.blas, c.stream, Nz, std::move(Lz), Dz, f.theta, W, &a2)) return e;
  f.maxAsym = std::max(a1, a2);
  // the factors zero-padded to Py x Py / Pz x Pz (column-major), their
  // transposes, and the inverse eigenvalue sums at (k, j) of a Pz x Py array
  auto pad = [](const std::vector<double>& a, int n, int p, bool trans) {
    std::vector<double> o((size_t)p * p, 0.0);
    for (int col = 0; col < n; ++col)
      for (int row = 0; row < n; ++row)
        o[(size_t)col * p + row] = trans ? a[(size_t)row * n + col] : a[(size_t)col * n + row];
    return o;
  };
  const std::vector<double> Vp = pad(V, Ny, Py, false), VTp = pad(V, Ny, Py, true);
  const std::vector<double> Wp = pad(W, Nz, Pz, false), WTp = pad(W, Nz, Pz, true);
  std::vector<double> inv((size_t)Pz * Py, 0.0);
  const double lmax = std::max(std::fabs(f.lam.back()), std::fabs(f.theta.back()));
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k) {
      const double den = -(f.lam[j] + f.theta[k]);
      inv[(size_t)j * Pz + k] = std::fabs(den) > 1e-13 * lmax ? 1.0 / den : 0.0;
    }
  if (resize) {
    void* bufs[] = {f.V64, f.W64, f.VT64, f.WT64, f.inv64, f.V32, f.W32, f.VT32, f.WT32, f.inv32,
                    f.Vb, f.Wb, f.VTb, f.WTb, f.g64, f.t64, f.g32, f.t32, f.gb, f.tb};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    f.V64 = f.W64 = f.VT64 = f.WT64 = f.inv64 = f.g64 = f.t64 = nullptr;
    f.V32 = f.W32 = f.VT32 = f.WT32 = f.inv32 = f.g32 = f.t32 = nullptr;
    f.Vb = f.Wb = f.VTb = f.WTb = f.gb = f.tb = nullptr;
  }
  f.Py = Py;
  f.Pz = Pz;
  auto up64 = [&](double*& p, const std::vector<double>& h) -> int {
    if (!p) SC_HIP(hipMalloc(&p, h.size() * sizeof(double)));
    SC_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  auto up32 = [&](float*& p, const std::vector<double>& h) -> int {
    std::vector<float> t(h.begin(), h.end());
    if (!p) SC_HIP(hipMalloc(&p, t.size() * sizeof(float)));
    SC_HIP(hipMemcpy(p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
  };
  auto upb = [&](void*& p, const std::vector<double>& h) -> int {
    std::vector<__hip_bfloat16> t(h.size());
    for (size_t i = 0; i < h.size(); ++i) t[i] = __float2bfloat16((float)h[i]);
    if (!p) SC_HIP(hipMalloc(&p, t.size() * sizeof(__hip_bfloat16)));
    SC_HIP(hipMemcpy(p, t.data(), t.size() * sizeof(__hip_bfloat16), hipMemcpyHostToDevice));
    return 0;
  };
  int e = up64(f.V64, Vp);
  if (!e) e = up64(f.W64, Wp);
  if (!e) e = up64(f.VT64, VTp);
  if (!e) e = up64(f.WT64, WTp);
  if (!e) e = up64(f.inv64, inv);
  if (!e) e = up32(f.V32, Vp);
  if (!e) e = up32(f.W32, Wp);
  if (!e) e = up32(f.VT32, VTp);
  if (!e) e = up32(f.WT32, WTp);
  if (!e) e = up32(f.inv32, inv);
  if (!e) e = upb(f.Vb, Vp);
  if (!e) e = upb(f.Wb, Wp);
  if (!e) e = upb(f.VTb, VTp);
  if (!e) e = upb(f.WTb, WTp);
  if (e) return e;
  const size_t P = (size_t)Pz * Py;
  if (!f.g64) {
    SC_HIP(hipMalloc(&f.g64, P * sizeof(double)));
    SC_HIP(hipMalloc(&f.t64, P * sizeof(double)));
    SC_HIP(hipMalloc(&f.g32, P * sizeof(float)));
    SC_HIP(hipMalloc(&f.t32, P * sizeof(float)));
    SC_HIP(hipMalloc(&f.gb, P * sizeof(__hip_bfloat16)));
    SC_HIP(hipMalloc(&f.tb, P * sizeof(__hip_bfloat16)));
  }
  f.bcKey = c.d.bcR | c.d.bcT << 1 | c.d.bcL << 2 | c.d.bcB << 3;
  return SC_OK;
}

int fdiag_apply(Ctx& c, FdDiag& f, int prec, const double* r, double* z) {
  hipStream_t s = c.stream;
  using u16 = unsigned short;
  if (prec == FD_F64)
    fd_transform<0, double, double>(s, f, f.V64, f.VT64, f.W64, f.WT64, f.inv64, f.g64, f.t64, r, z);
  else if (prec == FD_F32)
    fd_transform<1, double, double>(s, f, f.V32, f.VT32, f.W32, f.WT32, f.inv32, f.g32, f.t32, r, z);
  else
    fd_transform<2, double, double>(s, f, (const u16*)f.Vb, (const u16*)f.VTb, (const u16*)f.Wb,
                                         (const u16*)f.WTb, f.inv32, (u16*)f.gb, (u16*)f.tb, r, z);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int fdiag_apply_f32(Ctx& c, FdDiag& f, int prec, const float* r, float* z) {
  hipStream_t s = c.stream;
  using u16 = unsigned short;
  if (prec == FD_F32)
    fd_transform<1, float, float>(s, f, f.V32, f.VT32, f.W32, f.WT32, f.inv32, f.g32, f.t32, r, z);
  else if (prec == FD_BF16)
    fd_transform<2, float, float>(s, f, (const u16*)f.Vb, (const u16*)f.VTb, (const u16*)f.Wb,
                                       (const u16*)f.WTb, f.inv32, (u16*)f.gb, (u16*)f.tb, r, z);
  else
    return fail(SC_ERR_ARG, "the FP32 preconditioner application takes FP32 or bf16 GEMMs");
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc
