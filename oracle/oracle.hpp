// CPU FP64 oracle for the SCycle quasi-dynamic / fully-dynamic step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product path (scycle_amd/, the
// C-ABI library) may include, link or call this code.  It is used by tests/,
// by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline
// leg (timed as the CPU reference, never as the thing measured).
//
// It restates the reference's algorithm as the reference composes it: the
// 2-D SBP-SAT operator is ASSEMBLED as a sparse matrix from Kronecker products
// of 1-D SBP matrices (sbpOps_m_varGrid.cpp / sbpOps_m_constGrid.cpp), the
// opposite construction from the GPU library's matrix-free stencils, so a
// parity test compares two independent derivations.
//
// Layout: index I = j*Nz + k (z fastest), fault is j = 0 (domain.cpp:379-383,
// 449-457).  Units follow the reference (km, GPa, g/cm^3, MPa, m, m/s, s).
//
// How each part is pinned (the reference itself cannot be built here: it
// needs PETSc, MPI and Boost):
//   1-D SBP matrices    SBP identity, exactness per order, remainder structure
//                       (tests/test_oracle_sbp.py)
//   2-D operator        symmetry/definiteness, Neumann null space, closed-form
//                       order-2 interior stencil, var-grid == const-grid on a
//                       uniform grid, MMS convergence rates, the mu average
//                       with global index I+1 (tests/test_oracle_sbp.py)
//   friction, roots     strength limits, steady states of both state laws,
//                       closed-form and bisection roots (test_oracle_physics)
//   RK32 / RK43         stability polynomial / fifth-order local error,
//                       adaptive runs meeting their tolerance (same file)
//   wave step, FD fault constant state at rest, leapfrog energy, a_y sides,
//                       fault equations solved independently (same file)
//   power law           s^n and Arrhenius scaling, cap, SAT term (same file)
//   QD <-> FD switching a_y as the reference's loop writes it, the time step
//                       formula's cases, the phase sequence and switching
//                       thresholds of a nucleating event (same file); the
//                       phases themselves are the QD model and FD pieces
//                       pinned above
//   PETSc Vec / .in I/O the reference's own test fixture (tests/golden/)
// Parity unpinned: MUMPS's and PETSc's floating-point summation order (the
// assembled CSR matches the reference's algebra, not its bits).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- sparse --
struct Csr {
  int nr = 0, nc = 0;
  std::vector<int64_t> rp;  // nr+1
  std::vector<int> ci;
  std::vector<double> v;
};

// Map-backed 1-D matrix with the same set/get semantics as the reference's
// Spmat (spmat.hpp:51-70): operator()(r,c,val) inserts/overwrites.
struct Sp1 {
  int nr, nc;
  std::map<int, std::map<int, double>> m;
  Sp1(int r, int c) : nr(r), nc(c) {}
  void set(int r, int c, double val) { m[r][c] = val; }
  double get(int r, int c) const;
  void eye();
  void scale(double s);
};

Csr to_csr(const Sp1& a);
Csr diag(const std::vector<double>& d);
Csr eye(int n);
Csr kron(const Csr& a, const Csr& b);  // drops zero products like kronConvert
Csr matmul(const Csr& a, const Csr& b);
Csr matmul3(const Csr& a, const Csr& b, const Csr& c);  // (a*b)*c
Csr transpose(const Csr& a);
Csr axpy(const Csr& y, double alpha, const Csr& x);  // y + alpha*x
void scale(Csr& a, double s);
void spmv(const Csr& a, const double* x, double* y);          // y = a*x
void spmv_add(const Csr& a, const double* x, double* y);      // y += a*x
std::vector<double> diagonal(const Csr& a);

// ------------------------------------------------------------ 1-D SBP ----
// Restatement of sbp_Spmat / sbp_Spmat2 / sbp_Spmat4 (spmat.cpp:310-627).
struct Sbp1D {
  int n;
  Sp1 H, Hinv, D1, D1int, BS;  // order 2/4
  Sp1 D2, C2;                  // order 2 remainder
  Sp1 D3, D4, C3, C4;          // order 4 remainder
  explicit Sbp1D(int n_) : n(n_), H(n_, n_), Hinv(n_, n_), D1(n_, n_), D1int(n_, n_),
      BS(n_, n_), D2(n_, n_), C2(n_, n_), D3(n_, n_), D4(n_, n_), C3(n_, n_), C4(n_, n_) {}
};
int sbp_spmat(int order, int N, double scale, Sbp1D& o, const std::string& type);
int sbp_spmat2(int N, double scale, Sbp1D& o);
int sbp_spmat4(int N, double scale, Sbp1D& o);

// ------------------------------------------------------------ 2-D SBP ----
enum Bc { DIRICHLET = 0, NEUMANN = 1 };

struct SbpDesc {
  int order = 4;
  int Ny = 0, Nz = 0;
  double Ly = 1, Lz = 1;
  bool var_grid = true;             // variableGridSpacing vs constantGridSpacing
  std::string compat = "fullyCompatible";
  std::string laplace = "yz";
  int multByH = 1;
  int bcR = DIRICHLET, bcT = NEUMANN, bcL = DIRICHLET, bcB = NEUMANN;
};

// Assembled 2-D operator, mirroring SbpOps_m_varGrid / SbpOps_m_constGrid
// (computeMatrices, sbpOps_m_varGrid.cpp:280-307).
struct Sbp2D {
  SbpDesc d;
  double dy, dz, alphaDy, alphaDz, alphaT, h11y, h11z;
  std::vector<double> mu, y, z;   // N each (y,z empty => unit metric)
  Csr muM, yq, zr, qy, rz, J, Jinv, muqy, murz;
  Csr Dq_Iz, Iy_Dr, Dy_Iz, Iy_Dz;
  Csr H, Hinv, Hy_Iz, Iy_Hz, Hyinv_Iz, Iy_Hzinv;
  Csr E0y_Iz, ENy_Iz, Iy_E0z, Iy_ENz, e0y_Iz, eNy_Iz, Iy_e0z, Iy_eNz;
  Csr BSy_Iz, Iy_BSz, muxBySy_IzT, Iy_muxBzSzT;
  Csr D2, A;
  Csr AL, AR, AT, AB, rhsL, rhsR, rhsT, rhsB;
  int N() const { return d.Ny * d.Nz; }
};

int sbp2d_build(Sbp2D& s, const SbpDesc& d, const double* mu, const double* y, const double* z);
int sbp2d_change_bc(Sbp2D& s, int bcR, int bcT, int bcL, int bcB);
void sbp2d_set_rhs(const Sbp2D& s, double* rhs, const double* bcL, const double* bcR,
                   const double* bcT, const double* bcB);
void sbp2d_muxDy(const Sbp2D& s, const double* in, double* out);
void sbp2d_muxDz(const Sbp2D& s, const double* in, double* out);
void sbp2d_Dy(const Sbp2D& s, const double* in, double* out);
void sbp2d_Dz(const Sbp2D& s, const double* in, double* out);
void sbp2d_Dyxmu(const Sbp2D& s, const double* in, double* out);
void sbp2d_Dzxmu(const Sbp2D& s, const double* in, double* out);
void sbp2d_bnd(const Sbp2D& s, int which, const double* in, double* out);

// ----------------------------------------------------------- solvers -----
struct PcgStats { int iters = 0; double relres = 0; };
// Jacobi-preconditioned CG on A u = b (A symmetric definite), warm start in u.
PcgStats pcg(const Csr& A, const double* b, double* u, double rtol, int maxit);
// Banded Cholesky of -A (A symmetric negative definite): MUMPS stand-in.
struct BandChol {
  int n = 0, bw = 0;
  std::vector<double> L;  // n*(bw+1), row-major band of the lower factor
  int factor(const Csr& A);
  void solve(const double* b, double* x) const;  // solves A x = b
};

}  // namespace orc
