// Fast-diagonalisation preconditioner: the separable part of the SBP-SAT
// operator, A_sep = Ay (x) Dz + Dy (x) Az, inverted exactly through the
// generalised eigen-decompositions Ay V = Dy V Lambda, Az W = Dz W Theta:
//   A_sep^-1 = (V (x) W) diag(1/(lambda_j + theta_k)) (V^T (x) W^T).
// For a shear modulus that varies with depth only the two operators differ
// only by the reference's mu-averaging term, so CG converges in a handful of
// iterations; each application is four dense GEMMs on the matrix cores
// (this library's own MFMA kernels, the eigenvalue scaling and the precision
// conversions fused into their loads and stores).
#pragma once
#include <vector>

#include "sc_internal.hpp"

namespace sc {

enum FdPrec { FD_F64 = 0, FD_F32 = 1, FD_BF16 = 2 };

// GEMM tile edge: the factors and the intermediates are zero-padded to a
// multiple of it in both directions.
constexpr int kFdTile = 64;

struct FdDiag {
  int Ny = 0, Nz = 0;
  int Py = 0, Pz = 0;            // Ny, Nz rounded up to kFdTile
  int bcKey = -1;                // BC signature the factors were built for
  bool muSeparable = true;       // mu depends on depth only
  double maxAsym = 0;            // max |L - L^T| / max |L| of the 1-D operators
  std::vector<double> lam, theta;
  // device factors, column-major and zero-padded: V, VT (Py x Py), W, WT
  // (Pz x Pz), inv = 1/(-(lambda_j + theta_k)) at (k, j) (Pz x Py)
  double *V64 = nullptr, *W64 = nullptr, *VT64 = nullptr, *WT64 = nullptr, *inv64 = nullptr;
  float *V32 = nullptr, *W32 = nullptr, *VT32 = nullptr, *WT32 = nullptr, *inv32 = nullptr;
  void *Vb = nullptr, *Wb = nullptr, *VTb = nullptr, *WTb = nullptr;   // bf16
  // intermediates of one application (Pz x Py each)
  double *g64 = nullptr, *t64 = nullptr;
  float *g32 = nullptr, *t32 = nullptr;
  void *gb = nullptr, *tb = nullptr;
  void* blas = nullptr;           // rocblas_handle of the set-up's eigensolver
  ~FdDiag();
};

int fdiag_build(Ctx& c, FdDiag& f);
// z = A_sep^-1 r (FP64 in/out) with the GEMMs in the given precision.
int fdiag_apply(Ctx& c, FdDiag& f, int prec, const double* r, double* z);
// z = A_sep^-1 r with FP32 vectors (the inner solve of iterative
// refinement): prec FD_F32 (FP32 GEMMs) or FD_BF16 (bf16 GEMMs, FP32 accumulation).
int fdiag_apply_f32(Ctx& c, FdDiag& f, int prec, const float* r, float* z);

}  // namespace sc
