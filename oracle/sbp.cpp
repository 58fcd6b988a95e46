// Oracle restatement of the SBP operators (test infrastructure only; see
// oracle.hpp).  1-D matrices follow sbp_Spmat / sbp_Spmat2 / sbp_Spmat4
// (spmat.cpp:310-627) entry by entry; the 2-D operator is composed from
// Kronecker products exactly as SbpOps_m_varGrid::computeMatrices
// (sbpOps_m_varGrid.cpp:280-307, 333-1504) and SbpOps_m_constGrid
// (sbpOps_m_constGrid.cpp:261-1318) compose it.
#include <cassert>
#include <cmath>

#include "oracle.hpp"

namespace orc {

// spmat.cpp:310-491
int sbp_spmat(int order, int N, double scale, Sbp1D& o, const std::string& type) {
  if (N == 1) {
    o.H.eye();
    return 0;
  }
  auto& H = o.H;
  auto& Hinv = o.Hinv;
  auto& D1int = o.D1int;
  auto& BS = o.BS;
  if (order == 2) {
    H.eye();
    H.set(0, 0, 0.5);
    H.set(N - 1, N - 1, 0.5);
    H.scale(1 / scale);
    for (int i = 0; i < N; ++i) Hinv.set(i, i, 1 / H.get(i, i));
    D1int.set(0, 0, -1.0 * scale);
    D1int.set(0, 1, scale);
    for (int i = 1; i < N - 1; ++i) {
      D1int.set(i, i - 1, -0.5 * scale);
      D1int.set(i, i + 1, 0.5 * scale);
    }
    D1int.set(N - 1, N - 1, scale);
    D1int.set(N - 1, N - 2, -1 * scale);
    o.D1 = D1int;
    if (type == "fullyCompatible") {
      BS.set(0, 0, -D1int.get(0, 0));
      BS.set(0, 1, -D1int.get(0, 1));
      BS.set(N - 1, N - 2, D1int.get(N - 1, N - 2));
      BS.set(N - 1, N - 1, D1int.get(N - 1, N - 1));
    } else if (type == "compatible") {
      BS.set(0, 0, 1.5 * scale);
      BS.set(0, 1, -2.0 * scale);
      BS.set(0, 2, 0.5 * scale);
      BS.set(N - 1, N - 3, 0.5 * scale);
      BS.set(N - 1, N - 2, -2.0 * scale);
      BS.set(N - 1, N - 1, 1.5 * scale);
      o.D1.set(0, 0, -BS.get(0, 0));
      o.D1.set(0, 1, -BS.get(0, 1));
      o.D1.set(0, 2, -BS.get(0, 2));
      o.D1.set(N - 1, N - 3, BS.get(N - 1, N - 3));
      o.D1.set(N - 1, N - 2, BS.get(N - 1, N - 2));
      o.D1.set(N - 1, N - 1, BS.get(N - 1, N - 1));
    } else {
      return 2;
    }
    return 0;
  }
  if (order == 4) {
    if (!(N > 8)) return 3;
    H.eye();
    H.set(0, 0, 17.0 / 48.0);
    H.set(1, 1, 59.0 / 48.0);
    H.set(2, 2, 43.0 / 48.0);
    H.set(3, 3, 49.0 / 48.0);
    H.set(N - 1, N - 1, 17.0 / 48.0);
    H.set(N - 2, N - 2, 59.0 / 48.0);
    H.set(N - 3, N - 3, 43.0 / 48.0);
    H.set(N - 4, N - 4, 49.0 / 48.0);
    H.scale(1 / scale);
    for (int i = 0; i < N; ++i) Hinv.set(i, i, 1 / H.get(i, i));
    auto hi = [&](int i) { return Hinv.get(i, i); };
    for (int i = 4; i < N - 4; ++i) {
      D1int.set(i, i - 2, 1.0 / 12.0 * hi(i));
      D1int.set(i, i - 1, -2.0 / 3.0 * hi(i));
      D1int.set(i, i + 1, 2.0 / 3.0 * hi(i));
      D1int.set(i, i + 2, -1.0 / 12.0 * hi(i));
    }
    D1int.set(0, 0, -1.0 / 2.0 * hi(0));
    D1int.set(0, 1, 59.0 / 96.0 * hi(0));
    D1int.set(0, 2, -1.0 / 12.0 * hi(0));
    D1int.set(0, 3, -1.0 / 32.0 * hi(0));
    D1int.set(1, 0, -59.0 / 96.0 * hi(1));
    D1int.set(1, 2, 59.0 / 96.0 * hi(1));
    D1int.set(2, 0, 1.0 / 12.0 * hi(2));
    D1int.set(2, 1, -59.0 / 96.0 * hi(2));
    D1int.set(2, 3, 59.0 / 96.0 * hi(2));
    D1int.set(2, 4, -1.0 / 12.0 * hi(2));
    D1int.set(3, 0, 1.0 / 32.0 * hi(3));
    D1int.set(3, 2, -59.0 / 96.0 * hi(3));
    D1int.set(3, 4, 2.0 / 3.0 * hi(3));
    D1int.set(3, 5, -1.0 / 12.0 * hi(3));
    D1int.set(N - 1, N - 1, 1.0 / 2.0 * hi(N - 1));
    D1int.set(N - 1, N - 2, -59.0 / 96.0 * hi(N - 1));
    D1int.set(N - 1, N - 3, 1.0 / 12.0 * hi(N - 1));
    D1int.set(N - 1, N - 4, 1.0 / 32.0 * hi(N - 1));
    D1int.set(N - 2, N - 1, 59.0 / 96.0 * hi(N - 2));
    D1int.set(N - 2, N - 3, -59.0 / 96.0 * hi(N - 2));
    D1int.set(N - 3, N - 1, -1.0 / 12.0 * hi(N - 3));
    D1int.set(N - 3, N - 2, 59.0 / 96.0 * hi(N - 3));
    D1int.set(N - 3, N - 4, -59.0 / 96.0 * hi(N - 3));
    D1int.set(N - 3, N - 5, 1.0 / 12.0 * hi(N - 3));
    D1int.set(N - 4, N - 1, -1.0 / 32.0 * hi(N - 4));
    D1int.set(N - 4, N - 3, 59.0 / 96.0 * hi(N - 4));
    D1int.set(N - 4, N - 5, -2.0 / 3.0 * hi(N - 4));
    D1int.set(N - 4, N - 6, 1.0 / 12.0 * hi(N - 4));
    o.D1 = D1int;
    if (type == "fullyCompatible") {
      BS.set(0, 0, 24.0 / 17.0 * scale);
      BS.set(0, 1, -59.0 / 34.0 * scale);
      BS.set(0, 2, 4.0 / 17.0 * scale);
      BS.set(0, 3, 3.0 / 34.0 * scale);
      BS.set(N - 1, N - 1, 24.0 / 17.0 * scale);
      BS.set(N - 1, N - 2, -59.0 / 34.0 * scale);
      BS.set(N - 1, N - 3, 4.0 / 17.0 * scale);
      BS.set(N - 1, N - 4, 3.0 / 34.0 * scale);
    } else if (type == "compatible") {
      // spmat.cpp:461-467, including its unscaled last row and the
      // duplicated D1(N-1,N-4) assignment.
      BS.set(0, 0, 11.0 / 6.0 * scale);
      BS.set(0, 1, -3.0 * scale);
      BS.set(0, 2, 1.5 * scale);
      BS.set(0, 3, -1.0 / 3.0 * scale);
      BS.set(N - 1, N - 1, 11.0 / 6.0);
      BS.set(N - 1, N - 2, -3.0);
      BS.set(N - 1, N - 3, 1.5);
      BS.set(N - 1, N - 4, -1.0 / 3.0);
      o.D1.set(0, 0, -BS.get(0, 0));
      o.D1.set(0, 1, -BS.get(0, 1));
      o.D1.set(0, 2, -BS.get(0, 2));
      o.D1.set(0, 3, -BS.get(0, 3));
      o.D1.set(N - 1, N - 4, BS.get(N - 1, N - 4));
      o.D1.set(N - 1, N - 4, BS.get(N - 1, N - 4));
      o.D1.set(N - 1, N - 2, BS.get(N - 1, N - 2));
      o.D1.set(N - 1, N - 1, BS.get(N - 1, N - 1));
    }
    return 0;
  }
  return 1;
}

// spmat.cpp:494-544
int sbp_spmat2(int N, double scale, Sbp1D& o) {
  if (N == 1) return 0;
  if (!(N > 2)) return 1;
  double s2 = scale * scale;
  o.D2.set(0, 0, s2);
  o.D2.set(0, 1, -2.0 * s2);
  o.D2.set(0, 2, s2);
  for (int i = 1; i < N - 1; ++i) {
    o.D2.set(i, i - 1, s2);
    o.D2.set(i, i, -2.0 * s2);
    o.D2.set(i, i + 1, s2);
  }
  o.D2.set(N - 1, N - 3, s2);
  o.D2.set(N - 1, N - 2, -2.0 * s2);
  o.D2.set(N - 1, N - 1, s2);
  o.C2.eye();
  o.C2.set(0, 0, 0);
  o.C2.set(N - 1, N - 1, 0);
  return 0;
}

// spmat.cpp:547-627 (scale is unused by the reference for D3/D4)
int sbp_spmat4(int N, double /*scale*/, Sbp1D& o) {
  if (N == 1) return 0;
  if (!(N > 8)) return 1;
  auto& D3 = o.D3;
  auto& D4 = o.D4;
  auto& C3 = o.C3;
  auto& C4 = o.C4;
  D3.set(0, 0, -1); D3.set(0, 1, 3); D3.set(0, 2, -3); D3.set(0, 3, 1);
  D3.set(1, 0, -1); D3.set(1, 1, 3); D3.set(1, 2, -3); D3.set(1, 3, 1);
  D3.set(2, 0, -185893.0 / 301051.0);
  D3.set(2, 1, 79000249461.0 / 54642863857.0);
  D3.set(2, 2, -33235054191.0 / 54642863857.0);
  D3.set(2, 3, -36887526683.0 / 54642863857.0);
  D3.set(2, 4, 26183621850.0 / 54642863857.0);
  D3.set(2, 5, -4386.0 / 181507.0);
  for (int i = 3; i < N - 4; ++i) {
    D3.set(i, i - 1, -1.0);
    D3.set(i, i, 3);
    D3.set(i, i + 1, -3);
    D3.set(i, i + 2, 1.0);
  }
  D3.set(N - 3, N - 1, -D3.get(2, 0));
  D3.set(N - 3, N - 2, -D3.get(2, 1));
  D3.set(N - 3, N - 3, -D3.get(2, 2));
  D3.set(N - 3, N - 4, -D3.get(2, 3));
  D3.set(N - 3, N - 5, -D3.get(2, 4));
  D3.set(N - 3, N - 6, -D3.get(2, 5));
  D3.set(N - 2, N - 4, -1); D3.set(N - 2, N - 3, 3); D3.set(N - 2, N - 2, -3); D3.set(N - 2, N - 1, 1);
  D3.set(N - 1, N - 4, -1); D3.set(N - 1, N - 3, 3); D3.set(N - 1, N - 2, -3); D3.set(N - 1, N - 1, 1);

  D4.set(0, 0, 1); D4.set(0, 1, -4); D4.set(0, 2, 6); D4.set(0, 3, -4); D4.set(0, 4, 1);
  D4.set(1, 0, 1); D4.set(1, 1, -4); D4.set(1, 2, 6); D4.set(1, 3, -4); D4.set(1, 4, 1);
  for (int i = 2; i < N - 2; ++i) {
    D4.set(i, i - 2, 1);
    D4.set(i, i - 1, -4);
    D4.set(i, i, 6);
    D4.set(i, i + 1, -4);
    D4.set(i, i + 2, 1);
  }
  D4.set(N - 2, N - 5, 1); D4.set(N - 2, N - 4, -4); D4.set(N - 2, N - 3, 6); D4.set(N - 2, N - 2, -4); D4.set(N - 2, N - 1, 1);
  D4.set(N - 1, N - 5, 1); D4.set(N - 1, N - 4, -4); D4.set(N - 1, N - 3, 6); D4.set(N - 1, N - 2, -4); D4.set(N - 1, N - 1, 1);

  C3.eye();
  C3.set(0, 0, 0);
  C3.set(1, 1, 0);
  C3.set(2, 2, 163928591571.0 / 53268010936.0);
  C3.set(3, 3, 189284.0 / 185893.0);
  C3.set(N - 5, N - 5, C3.get(3, 3));
  C3.set(N - 4, N - 4, 0);
  C3.set(N - 3, N - 3, C3.get(2, 2));
  C3.set(N - 2, N - 2, 0);
  C3.set(N - 1, N - 1, 0);

  C4.eye();
  C4.set(0, 0, 0);
  C4.set(1, 1, 0);
  C4.set(2, 2, 1644330.0 / 301051.0);
  C4.set(3, 3, 156114.0 / 181507.0);
  C4.set(N - 4, N - 4, C4.get(3, 3));
  C4.set(N - 3, N - 3, C4.get(2, 2));
  C4.set(N - 2, N - 2, 0);
  C4.set(N - 1, N - 1, 0);
  return 0;
}

// ---------------------------------------------------------------------------
namespace {

Csr unit_vec_kron_left(int n, int idx, int nOther) {  // kron(e_idx (n x 1), I_nOther)
  Sp1 e(n, 1);
  if (n > 1) e.set(idx, 0, 1.0);
  return kron(to_csr(e), eye(nOther));
}
Csr unit_vec_kron_right(int nOther, int n, int idx) {  // kron(I_nOther, e_idx)
  Sp1 e(n, 1);
  if (n > 1) e.set(idx, 0, 1.0);
  return kron(eye(nOther), to_csr(e));
}
Csr unit_mat(int n, int idx) {
  Sp1 e(n, n);
  if (n > 1) e.set(idx, idx, 1.0);
  return to_csr(e);
}

// mu3 diagonal of the 4th-order remainder (sbpOps_m_varGrid.cpp:1285-1311 for
// z, :1428-1448 for y; identical in sbpOps_m_constGrid.cpp).  The neighbour is
// global index I+1 in both directions (serial ownership range [0, N)).
std::vector<double> mu3_y(const std::vector<double>& m) {
  const int64_t N = (int64_t)m.size();
  std::vector<double> o(m);
  o[N - 1] += m[N - 2];
  for (int64_t i = 1; i < N; ++i) o[i - 1] += m[i];
  for (auto& x : o) x *= 0.5;
  return o;
}
std::vector<double> mu3_z(const std::vector<double>& m) {
  const int64_t N = (int64_t)m.size();
  std::vector<double> o(m);
  o[0] += m[1];
  o[N - 1] += m[N - 2];
  for (int64_t i = 1; i < N - 1; ++i) o[i - 1] += m[i];
  for (auto& x : o) x *= 0.5;
  return o;
}

std::vector<double> diag_vals(const Csr& d) { return diagonal(d); }

// R_y (sbpOps_m_varGrid.cpp:1376-1497) / R_z (:1217-1360).  `cvec` is muqy
// (y) or murz (z) in the var grid, mu in the const grid; `ydir` selects
// kron(X, Iz) vs kron(Iy, X).
Csr build_R(const Sbp2D& s, const Csr& cmat, const std::vector<double>& cvec, bool ydir,
            const Csr& Iy, const Csr& Iz) {
  const int n = ydir ? s.d.Ny : s.d.Nz;
  const double h = ydir ? s.dy : s.dz;
  auto kr = [&](const Sp1& m) { return ydir ? kron(to_csr(m), Iz) : kron(Iy, to_csr(m)); };
  if (s.d.order == 2) {
    Sbp1D o(n);
    sbp_spmat2(n, 1.0 / h, o);
    Csr D2 = kr(o.D2), C2 = kr(o.C2);
    Csr temp = matmul(transpose(D2), C2);
    Csr R = matmul3(temp, cmat, D2);
    scale(R, 0.25 * std::pow(h, 3));
    return R;
  }
  Sbp1D o(n);
  sbp_spmat4(n, 1 / h, o);
  Csr mu3 = diag(ydir ? mu3_y(cvec) : mu3_z(cvec));
  Csr D3 = kr(o.D3), C3 = kr(o.C3);
  Csr temp1 = matmul(transpose(D3), C3);
  Csr temp2 = matmul3(temp1, mu3, D3);
  scale(temp2, 1.0 / h / 18.0);
  Csr D4 = kr(o.D4), C4 = kr(o.C4);
  temp1 = matmul(transpose(D4), C4);
  Csr R = matmul3(temp1, cmat, D4);
  scale(R, 1.0 / h / 144.0);
  return axpy(R, 1.0, temp2);  // MatAYPX(R, 1.0, temp2)
}

void build_bc(Sbp2D& s) {
  const bool vg = s.d.var_grid;
  // Dirichlet (sbpOps_m_varGrid.cpp:675-703; sbpOps_m_constGrid.cpp:504-534)
  auto dir = [&](double alphaD, const Csr& L, const Csr& mu, const Csr& Hinv, const Csr& BD1T,
                 const Csr& E) {
    Csr LxHxHinv = vg ? matmul3(s.H, L, Hinv) : matmul(s.H, Hinv);
    Csr HinvxmuxE = matmul3(LxHxHinv, mu, E);
    Csr out = matmul3(LxHxHinv, BD1T, E);
    return axpy(out, alphaD, HinvxmuxE);
  };
  // Neumann, operator part (varGrid :606-637; constGrid :438-470)
  auto neuA = [&](const Csr& L, const Csr& Hinv, double Bfact, const Csr& E, const Csr& mu,
                  const Csr& D1) {
    Csr out;
    if (vg) {
      Csr LxHinv = matmul(L, Hinv);
      Csr Exmu = matmul(E, mu);
      Csr LxHinvxExmu = matmul(LxHinv, Exmu);
      out = matmul3(s.H, LxHinvxExmu, D1);
    } else {
      Csr HinvxExmu = matmul3(Hinv, E, mu);
      out = matmul3(s.H, HinvxExmu, D1);
    }
    scale(out, Bfact * s.alphaT);
    return out;
  };
  // Neumann, rhs part (varGrid :643-668; constGrid :474-500)
  auto neuR = [&](const Csr& L, const Csr& Hinv, double Bfact, const Csr& e) {
    Csr out = vg ? matmul3(matmul(s.H, L), Hinv, e) : matmul3(s.H, Hinv, e);
    scale(out, Bfact * s.alphaT);
    return out;
  };
  const Csr& Lz = s.zr;  // metric factor for y-sides (unused in const grid)
  const Csr& Ly = s.yq;
  const Csr& cy = vg ? s.muqy : s.muM;
  const Csr& cz = vg ? s.murz : s.muM;
  if (s.d.bcR == DIRICHLET) {
    s.AR = dir(s.alphaDy, Lz, cy, s.Hyinv_Iz, s.muxBySy_IzT, s.ENy_Iz);
    s.rhsR = dir(s.alphaDy, Lz, cy, s.Hyinv_Iz, s.muxBySy_IzT, s.eNy_Iz);
  } else {
    s.AR = neuA(Lz, s.Hyinv_Iz, 1., s.ENy_Iz, s.muM, s.Dy_Iz);
    s.rhsR = neuR(Lz, s.Hyinv_Iz, 1., s.eNy_Iz);
  }
  if (s.d.bcT == DIRICHLET) {
    s.AT = dir(s.alphaDz, Ly, cz, s.Iy_Hzinv, s.Iy_muxBzSzT, s.Iy_E0z);
    s.rhsT = dir(s.alphaDz, Ly, cz, s.Iy_Hzinv, s.Iy_muxBzSzT, s.Iy_e0z);
  } else {
    s.AT = neuA(Ly, s.Iy_Hzinv, -1., s.Iy_E0z, s.muM, s.Iy_Dz);
    s.rhsT = neuR(Ly, s.Iy_Hzinv, -1., s.Iy_e0z);
  }
  if (s.d.bcL == DIRICHLET) {
    s.AL = dir(s.alphaDy, Lz, cy, s.Hyinv_Iz, s.muxBySy_IzT, s.E0y_Iz);
    s.rhsL = dir(s.alphaDy, Lz, cy, s.Hyinv_Iz, s.muxBySy_IzT, s.e0y_Iz);
  } else {
    s.AL = neuA(Lz, s.Hyinv_Iz, -1., s.E0y_Iz, s.muM, s.Dy_Iz);
    s.rhsL = neuR(Lz, s.Hyinv_Iz, -1., s.e0y_Iz);
  }
  if (s.d.bcB == DIRICHLET) {
    s.AB = dir(s.alphaDz, Ly, cz, s.Iy_Hzinv, s.Iy_muxBzSzT, s.Iy_ENz);
    s.rhsB = dir(s.alphaDz, Ly, cz, s.Iy_Hzinv, s.Iy_muxBzSzT, s.Iy_eNz);
  } else {
    s.AB = neuA(Ly, s.Iy_Hzinv, 1., s.Iy_ENz, s.muM, s.Iy_Dz);
    s.rhsB = neuR(Ly, s.Iy_Hzinv, 1., s.Iy_eNz);
  }
}

void assemble_A(Sbp2D& s) {
  const std::string& t = s.d.laplace;
  s.A = s.D2;
  if (t == "yz") {
    s.A = axpy(s.A, 1.0, s.AL);
    s.A = axpy(s.A, 1.0, s.AR);
    s.A = axpy(s.A, 1.0, s.AT);
    s.A = axpy(s.A, 1.0, s.AB);
  } else if (t == "y") {
    s.A = axpy(s.A, 1.0, s.AL);
    s.A = axpy(s.A, 1.0, s.AR);
  } else {
    s.A = axpy(s.A, 1.0, s.AT);
    s.A = axpy(s.A, 1.0, s.AB);
  }
}

}  // namespace

int sbp2d_build(Sbp2D& s, const SbpDesc& d, const double* mu, const double* y, const double* z) {
  s.d = d;
  const int Ny = d.Ny, Nz = d.Nz, N = Ny * Nz;
  if (d.order != 2 && d.order != 4) return 1;
  if (d.var_grid) {
    s.dy = 1. / (Ny - 1.);
    s.dz = 1. / (Nz - 1.);
    if (Ny == 1) s.dy = 1.;
    if (Nz == 1) s.dz = 1.;
  } else {
    s.dy = d.Ly / (Ny - 1.);
    s.dz = d.Lz / (Nz - 1.);
    if (Ny == 1) s.dy = d.Ly;
    if (Nz == 1) s.dz = d.Lz;
  }
  s.alphaT = -1.0;
  if (d.order == 2) {
    s.alphaDy = -4.0 / s.dy;
    s.alphaDz = -4.0 / s.dz;
    s.h11y = 0.5 * s.dy;
    s.h11z = 0.5 * s.dz;
  } else {
    s.alphaDy = 2.0 * -48.0 / 17.0 / s.dy;
    s.alphaDz = 2.0 * -48.0 / 17.0 / s.dz;
    s.h11y = 17.0 / 48.0 * s.dy;
    s.h11z = 17.0 / 48.0 * s.dz;
  }
  s.mu.assign(mu, mu + N);
  if (y) s.y.assign(y, y + N); else s.y.clear();
  if (z) s.z.assign(z, z + N); else s.z.clear();

  Sbp1D ty(Ny), tz(Nz);
  if (sbp_spmat(d.order, Ny, 1. / s.dy, ty, d.compat)) return 2;
  if (sbp_spmat(d.order, Nz, 1. / s.dz, tz, d.compat)) return 2;
  Csr Iy = eye(Ny), Iz = eye(Nz);
  Csr D1y = to_csr(ty.D1), D1z = to_csr(tz.D1);

  s.muM = diag(s.mu);
  // Jacobian (sbpOps_m_varGrid.cpp:333-385)
  if (d.var_grid) {
    s.Dq_Iz = kron(D1y, Iz);
    s.Iy_Dr = kron(Iy, D1z);
    std::vector<double> t(N, 1.0), inv(N);
    if (!s.y.empty()) spmv(s.Dq_Iz, s.y.data(), t.data());
    for (int i = 0; i < N; ++i) inv[i] = 1.0 / t[i];
    s.yq = diag(t);
    s.qy = diag(inv);
    std::fill(t.begin(), t.end(), 1.0);
    if (!s.z.empty()) spmv(s.Iy_Dr, s.z.data(), t.data());
    for (int i = 0; i < N; ++i) inv[i] = 1.0 / t[i];
    s.zr = diag(t);
    s.rz = diag(inv);
    s.J = matmul(s.yq, s.zr);
    s.Jinv = matmul(s.qy, s.rz);
    s.muqy = matmul(s.muM, s.qy);
    s.murz = matmul(s.muM, s.rz);
  } else {
    s.yq = s.zr = s.qy = s.rz = s.J = s.Jinv = eye(N);
    s.muqy = s.murz = s.muM;
  }
  // Es / es (sbpOps_m_varGrid.cpp:440-497)
  s.E0y_Iz = kron(unit_mat(Ny, 0), Iz);
  s.ENy_Iz = kron(unit_mat(Ny, Ny - 1), Iz);
  s.Iy_E0z = kron(Iy, unit_mat(Nz, 0));
  s.Iy_ENz = kron(Iy, unit_mat(Nz, Nz - 1));
  s.e0y_Iz = unit_vec_kron_left(Ny, 0, Nz);
  s.eNy_Iz = unit_vec_kron_left(Ny, Ny - 1, Nz);
  s.Iy_e0z = unit_vec_kron_right(Ny, Nz, 0);
  s.Iy_eNz = unit_vec_kron_right(Ny, Nz, Nz - 1);
  // Hs (sbpOps_m_varGrid.cpp:545-571)
  s.Hy_Iz = kron(to_csr(ty.H), Iz);
  s.Iy_Hz = kron(Iy, to_csr(tz.H));
  s.H = matmul(s.Hy_Iz, s.Iy_Hz);
  s.Hyinv_Iz = kron(to_csr(ty.Hinv), Iz);
  s.Iy_Hzinv = kron(Iy, to_csr(tz.Hinv));
  s.Hinv = matmul(s.Hyinv_Iz, s.Iy_Hzinv);
  // Bs (sbpOps_m_varGrid.cpp:499-528)
  s.BSy_Iz = kron(to_csr(ty.BS), Iz);
  s.Iy_BSz = kron(Iy, to_csr(tz.BS));
  s.muxBySy_IzT = matmul(transpose(s.BSy_Iz), d.var_grid ? s.muqy : s.muM);
  s.Iy_muxBzSzT = matmul(transpose(s.Iy_BSz), d.var_grid ? s.murz : s.muM);
  // 1st derivatives (varGrid :415-437; constGrid :1319-1343)
  if (d.var_grid) {
    s.Dy_Iz = matmul(s.qy, s.Dq_Iz);
    s.Iy_Dz = matmul(s.rz, s.Iy_Dr);
  } else {
    s.Dy_Iz = kron(D1y, Iz);
    s.Iy_Dz = kron(Iy, D1z);
  }
  build_bc(s);
  // D2 (varGrid :1133-1207 Dyymu/Dzzmu, :777-817 constructD2)
  Csr Dyymu, Dzzmu;
  const bool wantY = d.laplace == "yz" || d.laplace == "y";
  const bool wantZ = d.laplace == "yz" || d.laplace == "z";
  if (wantY) {
    std::vector<double> cy = d.var_grid ? diag_vals(s.muqy) : s.mu;
    Csr R = build_R(s, d.var_grid ? s.muqy : s.muM, cy, true, Iy, Iz);
    Csr HinvR = matmul(s.Hyinv_Iz, R);
    Dyymu = d.var_grid ? matmul3(s.Dq_Iz, s.muqy, s.Dq_Iz) : matmul3(s.Dy_Iz, s.muM, s.Dy_Iz);
    Dyymu = axpy(Dyymu, -1., HinvR);
    if (d.var_grid) Dyymu = d.multByH ? matmul3(s.H, s.zr, Dyymu) : matmul(s.zr, Dyymu);
    else if (d.multByH) Dyymu = matmul(s.H, Dyymu);
  }
  if (wantZ) {
    std::vector<double> cz = d.var_grid ? diag_vals(s.murz) : s.mu;
    Csr R = build_R(s, d.var_grid ? s.murz : s.muM, cz, false, Iy, Iz);
    Csr HinvR = matmul(s.Iy_Hzinv, R);
    Dzzmu = d.var_grid ? matmul3(s.Iy_Dr, s.murz, s.Iy_Dr) : matmul3(s.Iy_Dz, s.muM, s.Iy_Dz);
    Dzzmu = axpy(Dzzmu, -1., HinvR);
    if (d.var_grid) Dzzmu = d.multByH ? matmul3(s.H, s.yq, Dzzmu) : matmul(s.yq, Dzzmu);
    else if (d.multByH) Dzzmu = matmul(s.H, Dzzmu);
  }
  if (wantY && wantZ) s.D2 = axpy(Dyymu, 1.0, Dzzmu);
  else if (wantY) s.D2 = Dyymu;
  else s.D2 = Dzzmu;
  assemble_A(s);
  return 0;
}

int sbp2d_change_bc(Sbp2D& s, int bcR, int bcT, int bcL, int bcB) {
  s.d.bcR = bcR;
  s.d.bcT = bcT;
  s.d.bcL = bcL;
  s.d.bcB = bcB;
  build_bc(s);
  assemble_A(s);
  return 0;
}

// setRhs (sbpOps_m_varGrid.cpp:1046-1083)
void sbp2d_set_rhs(const Sbp2D& s, double* rhs, const double* bcL, const double* bcR,
                   const double* bcT, const double* bcB) {
  const int N = s.N();
  for (int i = 0; i < N; ++i) rhs[i] = 0.0;
  const std::string& t = s.d.laplace;
  if (t == "yz" || t == "y") {
    spmv_add(s.rhsL, bcL, rhs);
    spmv_add(s.rhsR, bcR, rhs);
  }
  if (t == "yz" || t == "z") {
    spmv_add(s.rhsT, bcT, rhs);
    spmv_add(s.rhsB, bcB, rhs);
  }
}

// muxDy / muxDz / Dy / Dz (sbpOps_m_varGrid.cpp:1602-1690)
void sbp2d_Dy(const Sbp2D& s, const double* in, double* out) { spmv(s.Dy_Iz, in, out); }
void sbp2d_Dz(const Sbp2D& s, const double* in, double* out) { spmv(s.Iy_Dz, in, out); }
void sbp2d_muxDy(const Sbp2D& s, const double* in, double* out) {
  std::vector<double> t(s.N());
  spmv(s.Dy_Iz, in, t.data());
  spmv(s.muM, t.data(), out);
}
void sbp2d_muxDz(const Sbp2D& s, const double* in, double* out) {
  std::vector<double> t(s.N());
  spmv(s.Iy_Dz, in, t.data());
  spmv(s.muM, t.data(), out);
}
// Dyxmu, Dzxmu: mu first, then the derivative (sbpOps_m_varGrid.cpp:1641-1660,
// 1704-1723)
void sbp2d_Dyxmu(const Sbp2D& s, const double* in, double* out) {
  std::vector<double> t(s.N());
  spmv(s.muM, in, t.data());
  spmv(s.Dy_Iz, t.data(), out);
}
void sbp2d_Dzxmu(const Sbp2D& s, const double* in, double* out) {
  std::vector<double> t(s.N());
  spmv(s.muM, in, t.data());
  spmv(s.Iy_Dz, t.data(), out);
}
// Hyinvxe0y ... HzinvxENz: the boundary restriction, then the Kronecker
// inverse 1-D norm (sbpOps_m_varGrid.cpp:1760-1889).  which 0, 1 take Nz
// values (e0y, eNy); 2..5 a grid field (E0y, ENy, E0z, ENz).
void sbp2d_bnd(const Sbp2D& s, int which, const double* in, double* out) {
  const Csr* e[6] = {&s.e0y_Iz, &s.eNy_Iz, &s.E0y_Iz, &s.ENy_Iz, &s.Iy_E0z, &s.Iy_ENz};
  const Csr& Hi = which < 4 ? s.Hyinv_Iz : s.Iy_Hzinv;
  std::vector<double> t(s.N());
  spmv(*e[which], in, t.data());
  spmv(Hi, t.data(), out);
}

}  // namespace orc
