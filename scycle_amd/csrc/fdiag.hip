// Fast-diagonalisation preconditioner: set-up (1-D operators, generalised
// eigen-decompositions with rocSOLVER) and application (four GEMMs, rocBLAS).
#include <hip/hip_bf16.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>

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

__global__ void k_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ void k_to_f64(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}
template <typename T>
__global__ void k_scale(size_t n, T* __restrict__ a, const T* __restrict__ s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] *= s[i];
}
__global__ void k_f64_to_bf16(size_t n, const double* __restrict__ a, __hip_bfloat16* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = __float2bfloat16((float)a[i]);
}
__global__ void k_f32_to_bf16(size_t n, const float* __restrict__ a, const float* __restrict__ s,
                              __hip_bfloat16* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = __float2bfloat16(s ? a[i] * s[i] : a[i]);
}

unsigned nblk(size_t n) {
  size_t b = (n + 255) / 256;
  return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 4096);
}

}  // namespace

FdDiag::~FdDiag() {
  void* bufs[] = {V64, W64, VT64, WT64, inv64, V32, W32, VT32, WT32, inv32, Vb, Wb, VTb, WTb,
                  g64, t64, r32, g32, t32, rb, gb};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (blas) rocblas_destroy_handle((rocblas_handle)blas);
}

int fdiag_build(Ctx& c, FdDiag& f) {
  const int Ny = c.d.Ny, Nz = c.d.Nz;
  const size_t N = (size_t)Ny * Nz;
  if (c.d.laplace != 0) return fail(SC_ERR_ARG, "fast-diagonalisation preconditioner needs laplace type yz");
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
  std::vector<double> inv(N);
  const double lmax = std::max(std::fabs(f.lam.back()), std::fabs(f.theta.back()));
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k) {
      const double den = -(f.lam[j] + f.theta[k]);
      inv[(size_t)j * Nz + k] = std::fabs(den) > 1e-13 * lmax ? 1.0 / den : 0.0;
    }
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
  // explicit transposes, so that all four GEMMs run in the faster NN form
  auto transpose = [](const std::vector<double>& a, int n) {
    std::vector<double> t(a.size());
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) t[(size_t)j * n + i] = a[(size_t)i * n + j];
    return t;
  };
  const std::vector<double> VT = transpose(V, Ny), WT = transpose(W, Nz);
  int e = up64(f.V64, V);
  if (!e) e = up64(f.W64, W);
  if (!e) e = up64(f.VT64, VT);
  if (!e) e = up64(f.WT64, WT);
  if (!e) e = up64(f.inv64, inv);
  if (!e) e = up32(f.V32, V);
  if (!e) e = up32(f.W32, W);
  if (!e) e = up32(f.VT32, VT);
  if (!e) e = up32(f.WT32, WT);
  if (!e) e = up32(f.inv32, inv);
  if (!e) e = upb(f.Vb, V);
  if (!e) e = upb(f.Wb, W);
  if (!e) e = upb(f.VTb, VT);
  if (!e) e = upb(f.WTb, WT);
  if (e) return e;
  if (!f.g64) {
    SC_HIP(hipMalloc(&f.g64, N * sizeof(double)));
    SC_HIP(hipMalloc(&f.t64, N * sizeof(double)));
    SC_HIP(hipMalloc(&f.r32, N * sizeof(float)));
    SC_HIP(hipMalloc(&f.g32, N * sizeof(float)));
    SC_HIP(hipMalloc(&f.t32, N * sizeof(float)));
    SC_HIP(hipMalloc(&f.rb, N * sizeof(__hip_bfloat16)));
    SC_HIP(hipMalloc(&f.gb, N * sizeof(__hip_bfloat16)));
  }
  rocblas_set_stream((rocblas_handle)f.blas, c.stream);
  f.bcKey = c.d.bcR | c.d.bcT << 1 | c.d.bcL << 2 | c.d.bcB << 3;
  return SC_OK;
}

#define SC_BLAS(call)                                                                  \
  do {                                                                                 \
    rocblas_status _s = (call);                                                        \
    if (_s != rocblas_status_success)                                                  \
      return fail(SC_ERR_HIP, std::string(#call " -> ") + rocblas_status_to_string(_s)); \
  } while (0)

// Column-major view: r as R_c (Nz x Ny, ld Nz).  z = W (inv .* (W^T R_c V)) V^T.
int fdiag_apply(Ctx& c, FdDiag& f, int prec, const double* r, double* z) {
  const int Ny = f.Ny, Nz = f.Nz;
  const size_t N = (size_t)Ny * Nz;
  hipStream_t s = c.stream;
  rocblas_handle hb = (rocblas_handle)f.blas;
  rocblas_set_stream(hb, s);
  const rocblas_operation nN = rocblas_operation_none;
  if (prec == FD_F64) {
    const double one = 1.0, zero = 0.0;
    SC_BLAS(rocblas_dgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.WT64, Nz, r, Nz, &zero, f.g64, Nz));
    SC_BLAS(rocblas_dgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g64, Nz, f.V64, Ny, &zero, f.t64, Nz));
    hipLaunchKernelGGL(k_scale<double>, dim3(nblk(N)), dim3(256), 0, s, N, f.t64, f.inv64);
    SC_BLAS(rocblas_dgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.W64, Nz, f.t64, Nz, &zero, f.g64, Nz));
    SC_BLAS(rocblas_dgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g64, Nz, f.VT64, Ny, &zero, z, Nz));
  } else if (prec == FD_F32) {
    const float one = 1.0f, zero = 0.0f;
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(N)), dim3(256), 0, s, N, r, f.r32);
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.WT32, Nz, f.r32, Nz, &zero, f.g32, Nz));
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g32, Nz, f.V32, Ny, &zero, f.t32, Nz));
    hipLaunchKernelGGL(k_scale<float>, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, f.inv32);
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.W32, Nz, f.t32, Nz, &zero, f.g32, Nz));
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g32, Nz, f.VT32, Ny, &zero, f.t32, Nz));
    hipLaunchKernelGGL(k_to_f64, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, z);
  } else {
    // bf16 inputs, FP32 accumulation and output; intermediates re-rounded to bf16
    const float one = 1.0f, zero = 0.0f;
    const rocblas_datatype bf = rocblas_datatype_bf16_r, f32 = rocblas_datatype_f32_r;
    __hip_bfloat16* rb = (__hip_bfloat16*)f.rb;
    __hip_bfloat16* gb = (__hip_bfloat16*)f.gb;
    hipLaunchKernelGGL(k_f64_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, r, rb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Nz, &one, f.WTb, bf, Nz, rb, bf, Nz, &zero, f.g32, f32,
                            Nz, f.g32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.g32, (const float*)nullptr, gb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Ny, &one, gb, bf, Nz, f.Vb, bf, Ny, &zero, f.t32, f32,
                            Nz, f.t32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, (const float*)f.inv32, rb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Nz, &one, f.Wb, bf, Nz, rb, bf, Nz, &zero, f.g32, f32,
                            Nz, f.g32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.g32, (const float*)nullptr, gb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Ny, &one, gb, bf, Nz, f.VTb, bf, Ny, &zero, f.t32, f32,
                            Nz, f.t32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_to_f64, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, z);
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

int fdiag_apply_f32(Ctx& c, FdDiag& f, int prec, const float* r, float* z) {
  const int Ny = f.Ny, Nz = f.Nz;
  const size_t N = (size_t)Ny * Nz;
  hipStream_t s = c.stream;
  rocblas_handle hb = (rocblas_handle)f.blas;
  rocblas_set_stream(hb, s);
  const rocblas_operation nN = rocblas_operation_none;
  const float one = 1.0f, zero = 0.0f;
  if (prec == FD_F32) {
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.WT32, Nz, r, Nz, &zero, f.g32, Nz));
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g32, Nz, f.V32, Ny, &zero, f.t32, Nz));
    hipLaunchKernelGGL(k_scale<float>, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, f.inv32);
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Nz, &one, f.W32, Nz, f.t32, Nz, &zero, f.g32, Nz));
    SC_BLAS(rocblas_sgemm(hb, nN, nN, Nz, Ny, Ny, &one, f.g32, Nz, f.VT32, Ny, &zero, z, Nz));
  } else if (prec == FD_BF16) {
    const rocblas_datatype bf = rocblas_datatype_bf16_r, f32 = rocblas_datatype_f32_r;
    __hip_bfloat16* rb = (__hip_bfloat16*)f.rb;
    __hip_bfloat16* gb = (__hip_bfloat16*)f.gb;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, r, (const float*)nullptr, rb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Nz, &one, f.WTb, bf, Nz, rb, bf, Nz, &zero, f.g32, f32,
                            Nz, f.g32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.g32, (const float*)nullptr, gb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Ny, &one, gb, bf, Nz, f.Vb, bf, Ny, &zero, f.t32, f32,
                            Nz, f.t32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.t32, (const float*)f.inv32, rb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Nz, &one, f.Wb, bf, Nz, rb, bf, Nz, &zero, f.g32, f32,
                            Nz, f.g32, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(N)), dim3(256), 0, s, N, f.g32, (const float*)nullptr, gb);
    SC_BLAS(rocblas_gemm_ex(hb, nN, nN, Nz, Ny, Ny, &one, gb, bf, Nz, f.VTb, bf, Ny, &zero, z, f32,
                            Nz, z, f32, Nz, f32, rocblas_gemm_algo_standard, 0, 0));
  } else {
    return fail(SC_ERR_ARG, "the FP32 preconditioner application takes FP32 or bf16 GEMMs");
  }
  SC_HIP(hipGetLastError());
  return SC_OK;
}

}  // namespace sc
