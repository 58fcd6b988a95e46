// Reference input-deck support: the `.in` key/value format, PETSc-binary Vec
// files, and the mediator set-up that turns a deck into a QD driver
// (Domain::loadSettings domain.cpp:133-196, LinearElastic / Fault /
// StrikeSlip_LinearElastic_qd ::loadSettings, genFuncs.cpp:380-440 and
// 574-617), plus the steady-state initial conditions solveSS
// (strikeSlip_linearElastic_qd.cpp:895-1006).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <memory>
#include <vector>

#include "input.hpp"
#include "qd_internal.hpp"

namespace sc {

// ------------------------------------------------------------ .in format --
// "var = value # comment": the separator is exactly " = ", the value is cut
// at its first space; rhsFull keeps everything after the separator.
std::vector<InLine> read_in_file(const std::string& path, bool* ok) {
  std::vector<InLine> out;
  std::ifstream f(path);
  *ok = f.good();
  std::string line;
  const std::string delim = " = ";
  while (std::getline(f, line)) {
    size_t pos = line.find(delim);
    InLine l;
    l.var = line.substr(0, pos);
    if (line.length() > pos + delim.length()) l.rhs = line.substr(pos + delim.length());
    l.rhsFull = l.rhs;
    l.rhs = l.rhs.substr(0, l.rhs.find(' '));
    out.push_back(l);
  }
  return out;
}

// "[v1 v2 ... vn]" with single spaces; a one-element list parses as empty
// (genFuncs.cpp:413-440).  Values are appended to vec.
void parse_vector(const std::string& str, std::vector<double>& vec) {
  size_t pos = str.find(']');
  if (str.empty()) return;
  std::string rem = str.substr(1, pos - 1);
  pos = rem.find(' ');
  while (pos != std::string::npos) {
    pos = rem.find(' ');
    vec.push_back(std::atof(rem.substr(0, pos).c_str()));
    rem = rem.substr(pos + 1);
  }
}
void parse_vector(const std::string& str, std::vector<std::string>& vec) {
  size_t pos = str.find(']');
  if (str.empty()) return;
  std::string rem = str.substr(1, pos - 1);
  pos = rem.find(' ');
  while (pos != std::string::npos) {
    pos = rem.find(' ');
    vec.push_back(rem.substr(0, pos));
    rem = rem.substr(pos + 1);
  }
}

// setVec: piecewise-linear through (vals, depths), last segment extended.
void set_vec(std::vector<double>& out, const std::vector<double>& coord,
             const std::vector<double>& vals, const std::vector<double>& depths) {
  out.assign(coord.size(), vals.empty() ? 0.0 : vals[0]);
  if (coord.size() == 1 || depths.size() < 2) return;
  const size_t n = depths.size();
  for (size_t i = 0; i < coord.size(); ++i) {
    const double z = coord[i];
    for (size_t ind = 0; ind + 1 < n; ++ind) {
      const double z0 = depths[ind], z1 = depths[ind + 1], v0 = vals[ind], v1 = vals[ind + 1];
      if ((z >= z0 && z <= z1) || (z > z1 && ind == n - 2)) out[i] = (v1 - v0) / (z1 - z0) * (z - z0) + v0;
    }
  }
}

// --------------------------------------------------------- PETSc binary --
namespace {
uint32_t be32(const unsigned char* p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}
double be64(const unsigned char* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v = v << 8 | p[i];
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}
void put_be32(std::vector<unsigned char>& o, uint32_t v) {
  for (int i = 3; i >= 0; --i) o.push_back((unsigned char)(v >> (8 * i)));
}
void put_be64(std::vector<unsigned char>& o, double d) {
  uint64_t v;
  std::memcpy(&v, &d, 8);
  for (int i = 7; i >= 0; --i) o.push_back((unsigned char)(v >> (8 * i)));
}
}  // namespace

constexpr uint32_t kVecClassId = 1211214;

// Reads every Vec record of a PETSc binary file (an appended time series has
// several).  Returns the number of records, or -1 on a format error.
int read_petsc_vecs(const std::string& path, std::vector<std::vector<double>>& recs) {
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) return -1;
  std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  size_t p = 0;
  recs.clear();
  while (p + 8 <= buf.size()) {
    if (be32(&buf[p]) != kVecClassId) return -1;
    const uint32_t n = be32(&buf[p + 4]);
    p += 8;
    if (p + (size_t)n * 8 > buf.size()) return -1;
    std::vector<double> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = be64(&buf[p + 8 * (size_t)i]);
    p += (size_t)n * 8;
    recs.push_back(std::move(v));
  }
  return p == buf.size() ? (int)recs.size() : -1;
}

int write_petsc_vec(const std::string& path, const double* v, int n, bool append) {
  std::vector<unsigned char> o;
  put_be32(o, kVecClassId);
  put_be32(o, (uint32_t)n);
  for (int i = 0; i < n; ++i) put_be64(o, v[i]);
  std::FILE* f = std::fopen(path.c_str(), append ? "ab" : "wb");
  if (!f) return -1;
  size_t w = std::fwrite(o.data(), 1, o.size(), f);
  std::fclose(f);
  return w == o.size() ? 0 : -1;
}

// loadVecFromInputFile: overrides `out` when inputDir+name exists.
static bool load_override(const std::string& dir, const std::string& name, std::vector<double>& out) {
  std::vector<std::vector<double>> recs;
  if (read_petsc_vecs(dir + name, recs) < 1) return false;
  if (recs[0].size() != out.size()) return false;
  out = recs[0];
  return true;
}

// ------------------------------------------------------------------ deck --
int load_deck(const std::string& path, Deck& d) {
  bool ok;
  auto lines = read_in_file(path, &ok);
  if (!ok) return fail(SC_ERR_IO, "cannot read input file " + path);
  for (const auto& l : lines) {
    const std::string& v = l.var;
    const std::string& r = l.rhs;
    // Domain (domain.cpp:167-188)
    if (v == "order") d.order = std::atoi(r.c_str());
    else if (v == "Ny" && d.Ny < 0) d.Ny = std::atoi(r.c_str());
    else if (v == "Nz" && d.Nz < 0) d.Nz = std::atoi(r.c_str());
    else if (v == "Ly") d.Ly = std::atof(r.c_str());
    else if (v == "Lz") d.Lz = std::atof(r.c_str());
    else if (v == "inputDir") d.inputDir = r;
    else if (v == "outputDir") { d.outputDir = r; d.haveOutputDir = true; }
    else if (v == "enableCheckpointing") d.ckpt = std::atoi(r.c_str());
    else if (v == "interval") d.interval = (int)std::atof(r.c_str());
    else if (v == "stride1D") d.stride1D = (int)std::atof(r.c_str());
    else if (v == "stride2D") d.stride2D = (int)std::atof(r.c_str());
    else if (v == "sbpCompatibilityType") d.compat = r;
    else if (v == "gridSpacingType") d.gridSpacingType = r;
    else if (v == "bulkDeformationType") d.bulk = r;
    else if (v == "momentumBalanceType") d.momBal = r;
    else if (v == "bCoordTrans") d.bCoordTrans = std::atof(r.c_str());
    else if (v == "vL") d.vL = std::atof(r.c_str());
    // LinearElastic (linearElastic.cpp:129-139) and Fault (fault.cpp:63-101)
    else if (v == "muVals") parse_vector(l.rhsFull, d.muVals);
    else if (v == "muDepths") parse_vector(l.rhsFull, d.muDepths);
    else if (v == "rhoVals") parse_vector(l.rhsFull, d.rhoVals);
    else if (v == "rhoDepths") parse_vector(l.rhsFull, d.rhoDepths);
    else if (v == "DcVals") parse_vector(l.rhsFull, d.DcVals);
    else if (v == "DcDepths") parse_vector(l.rhsFull, d.DcDepths);
    else if (v == "sNVals") parse_vector(l.rhsFull, d.sNVals);
    else if (v == "sNDepths") parse_vector(l.rhsFull, d.sNDepths);
    else if (v == "sN_cap") d.sNcap = std::atof(r.c_str());
    else if (v == "sN_floor") d.sNfloor = std::atof(r.c_str());
    else if (v == "aVals") parse_vector(l.rhsFull, d.aVals);
    else if (v == "aDepths") parse_vector(l.rhsFull, d.aDepths);
    else if (v == "bVals") parse_vector(l.rhsFull, d.bVals);
    else if (v == "bDepths") parse_vector(l.rhsFull, d.bDepths);
    else if (v == "stateVals") parse_vector(l.rhsFull, d.stateVals);
    else if (v == "stateDepths") parse_vector(l.rhsFull, d.stateDepths);
    else if (v == "lockedVals") parse_vector(l.rhsFull, d.lockedVals);
    else if (v == "lockedDepths") parse_vector(l.rhsFull, d.lockedDepths);
    else if (v == "stateLaw") d.stateLaw = r;
    else if (v == "rootTol") d.rootTol = std::atof(r.c_str());
    else if (v == "f0") d.f0 = std::atof(r.c_str());
    else if (v == "v0") d.v0 = std::atof(r.c_str());
    // QD mediator (strikeSlip_linearElastic_qd.cpp:158-186)
    else if (v == "guessSteadyStateICs") d.guessSS = std::atoi(r.c_str());
    else if (v == "timeIntegrator") d.timeIntegrator = r;
    else if (v == "timeControlType") d.timeControlType = r;
    else if (v == "maxStepCount") d.maxStepCount = (int)std::atof(r.c_str());
    else if (v == "initTime") d.initTime = std::atof(r.c_str());
    else if (v == "maxTime") d.maxTime = std::atof(r.c_str());
    else if (v == "minDeltaT") d.minDeltaT = std::atof(r.c_str());
    else if (v == "maxDeltaT") d.maxDeltaT = std::atof(r.c_str());
    else if (v == "initDeltaT") d.initDeltaT = std::atof(r.c_str());
    else if (v == "timeStepTol") d.timeStepTol = std::atof(r.c_str());
    else if (v == "timeIntInds") parse_vector(l.rhsFull, d.timeIntInds);
    else if (v == "normType") d.normType = r;
    else if (v == "momBal_bcR_qd") d.bcR = r;
    else if (v == "momBal_bcT_qd") d.bcT = r;
    else if (v == "momBal_bcL_qd") d.bcL = r;
    else if (v == "momBal_bcB_qd") d.bcB = r;
    else if (v == "linSolver") d.linSolver = r;
    else if (v == "kspTol") d.kspTol = std::atof(r.c_str());
  }
  return SC_OK;
}

static int side_type(const std::string& s, int* mat) {
  if (s == "symmFault" || s == "rigidFault" || s == "remoteLoading") *mat = SC_DIRICHLET;
  else if (s == "freeSurface" || s == "outGoingCharacteristics") *mat = SC_NEUMANN;
  else return 1;
  return 0;
}

// Builds all host arrays of a QD run from a deck.
int build_from_deck(const Deck& d, QdInputs& in) {
  if (d.Ny < 3 || d.Nz < 3) return fail(SC_ERR_ARG, "deck: Ny and Nz must be >= 3");
  if (d.bulk != "linearElastic" || d.momBal != "quasidynamic")
    return fail(SC_ERR_ARG, "deck: only bulkDeformationType = linearElastic with momentumBalanceType = quasidynamic is served by this driver");
  const int Ny = d.Ny, Nz = d.Nz, N = Ny * Nz;
  const double dq = 1.0 / (Ny - 1.0), dr = 1.0 / (Nz - 1.0);
  in.y.resize(N);
  in.z.resize(N);
  for (int I = 0; I < N; ++I) {
    const double q = dq * (I / Nz), r = dr * (I - Nz * (I / Nz));
    in.y[I] = d.bCoordTrans <= 0 ? q * d.Ly : d.Ly * std::sinh(d.bCoordTrans * q) / std::sinh(d.bCoordTrans);
    in.z[I] = r * d.Lz;
  }
  load_override(d.inputDir, "y", in.y);
  load_override(d.inputDir, "z", in.z);
  // bulk material: mu interpolated along y, rho along z (linearElastic.cpp:353-354)
  std::vector<double> rho, cs;
  set_vec(in.mu, in.y, d.muVals, d.muDepths);
  set_vec(rho, in.z, d.rhoVals, d.rhoDepths);
  cs.resize(N);
  for (int i = 0; i < N; ++i) cs[i] = std::sqrt(std::fabs(in.mu[i] / rho[i]));
  std::vector<double> bcL(Nz, 0.0);
  in.bcRShift.assign(Nz, 0.0);
  load_override(d.inputDir, "momBal_bcL", bcL);
  load_override(d.inputDir, "momBal_bcR", in.bcRShift);
  load_override(d.inputDir, "momBal_mu", in.mu);
  load_override(d.inputDir, "momBal_rho", rho);
  load_override(d.inputDir, "momBal_cs", cs);
  // fault fields along the fault's z (fault.cpp:227-312)
  std::vector<double> zf(in.z.begin(), in.z.begin() + Nz);
  set_vec(in.a, zf, d.aVals, d.aDepths);
  set_vec(in.b, zf, d.bVals, d.bDepths);
  set_vec(in.sN, zf, d.sNVals, d.sNDepths);
  set_vec(in.Dc, zf, d.DcVals, d.DcDepths);
  if (!d.lockedVals.empty()) set_vec(in.locked, zf, d.lockedVals, d.lockedDepths);
  else in.locked.assign(Nz, 0.0);
  std::vector<double> muf, rhof;
  set_vec(muf, zf, d.muVals, d.muDepths);
  set_vec(rhof, zf, d.rhoVals, d.rhoDepths);
  if (!d.stateVals.empty()) set_vec(in.psi, zf, d.stateVals, d.stateDepths);
  else in.psi.assign(Nz, d.f0);
  for (auto& s : in.sN) s = std::max(std::min(s, d.sNcap), d.sNfloor);
  load_override(d.inputDir, "sNEff", in.sN);
  load_override(d.inputDir, "psi", in.psi);
  std::vector<double> tauQS(Nz, 0.0);
  in.prestress.assign(Nz, 0.0);
  load_override(d.inputDir, "tauQS", tauQS);
  load_override(d.inputDir, "prestress", in.prestress);
  load_override(d.inputDir, "a", in.a);
  load_override(d.inputDir, "b", in.b);
  // mediator
  int mR, mT, mL, mB;
  if (side_type(d.bcR, &mR) || side_type(d.bcT, &mT) || side_type(d.bcL, &mL) || side_type(d.bcB, &mB))
    return fail(SC_ERR_ARG, "deck: unknown momBal_bc*_qd type");
  const double fts = d.bcL == "rigidFault" ? 1.0 : 2.0;
  in.eta.resize(Nz);
  for (int k = 0; k < Nz; ++k) in.eta[k] = std::sqrt(std::fabs(muf[k] * rhof[k])) * (1.0 / fts);
  in.slip.resize(Nz);
  for (int k = 0; k < Nz; ++k) in.slip[k] = bcL[k] * fts;
  sc_qd_desc& q = in.desc;
  std::memset(&q, 0, sizeof(q));
  q.grid.Ny = Ny;
  q.grid.Nz = Nz;
  q.grid.Ly = d.Ly;
  q.grid.Lz = d.Lz;
  q.grid.order = d.order;
  q.grid.var_grid = d.gridSpacingType == "variableGridSpacing";
  if (!q.grid.var_grid && d.gridSpacingType != "constantGridSpacing")
    return fail(SC_ERR_ARG, "deck: unknown gridSpacingType");
  q.grid.compat = d.compat == "compatible";
  q.grid.laplace = 0;
  q.grid.bcR = mR;
  q.grid.bcT = mT;
  q.grid.bcL = mL;
  q.grid.bcB = mB;
  q.remoteLoadingR = d.bcR == "remoteLoading";
  q.faultTypeScale = fts;
  q.vL = d.vL;
  q.f0 = d.f0;
  q.v0 = d.v0;
  q.rootTol = d.rootTol;
  q.maxNumIts = 10000;
  if (d.stateLaw == "agingLaw") q.stateLaw = SC_AGING_LAW;
  else if (d.stateLaw == "slipLaw") q.stateLaw = SC_SLIP_LAW;
  else return fail(SC_ERR_ARG, "deck: stateLaw must be agingLaw or slipLaw for this driver");
  q.jac_mode = 0;
  q.solve.mode = SC_SOLVE_CG64;
  q.solve.rtol = 1e-12;
  q.solve.inner_rtol = 1e-5;
  q.solve.max_iters = 1000000;
  q.solve.max_outer = 50;
  q.solve.check_every = 64;
  q.solve.precond = SC_PREC_JACOBI;
  q.solve.reuse = 0;
  // run settings (odeSolver set-up in StrikeSlip_LinearElastic_qd::integrate)
  sc_qd_run& r = in.run;
  std::memset(&r, 0, sizeof(r));
  if (d.timeIntegrator == "RK32") r.integrator = 0;
  else if (d.timeIntegrator == "RK43") r.integrator = 1;
  else return fail(SC_ERR_ARG, "deck: timeIntegrator must be RK32 or RK43");
  r.control = d.timeControlType == "P" ? 0 : 1;
  r.normType = d.normType == "L2_relative" ? 1 : 0;
  r.maxSteps = d.maxStepCount;
  r.initTime = d.initTime;
  r.maxTime = d.maxTime;
  r.maxDt = d.maxDeltaT;
  r.tol = d.timeStepTol;
  // computeMinTimeStep (strikeSlip_linearElastic_qd.cpp:262-322): the grid
  // spacing over the shear-wave speed, minimum over the body
  double minTs = INFINITY;
  {
    std::vector<double> yq(Ny, d.Ly / (Ny - 1.0)), zr(Nz, d.Lz / (Nz - 1.0));
    if (q.grid.var_grid) {
      Dir1D ty, tz;
      int e = build_dir(d.order, Ny, 1. / (Ny - 1.), q.grid.compat, ty);
      if (!e) e = build_dir(d.order, Nz, 1. / (Nz - 1.), q.grid.compat, tz);
      if (e) return e;
      for (int j = 0; j < Ny; ++j) {
        double s = 0;
        for (int o = 0; o < kBand; ++o) {
          int c = j + o - 4;
          if (c >= 0 && c < Ny) s += ty.D1[(size_t)j * kBand + o] * in.y[(size_t)c * Nz];
        }
        yq[j] = s / (Ny - 1);
      }
      for (int k = 0; k < Nz; ++k) {
        double s = 0;
        for (int o = 0; o < kBand; ++o) {
          int c = k + o - 4;
          if (c >= 0 && c < Nz) s += tz.D1[(size_t)k * kBand + o] * in.z[c];
        }
        zr[k] = s / (Nz - 1);
      }
    }
    for (int j = 0; j < Ny; ++j)
      for (int k = 0; k < Nz; ++k) {
        const double c = cs[(size_t)j * Nz + k];
        minTs = std::min(minTs, std::min(yq[j] / c, zr[k] / c));
      }
  }
  r.minDt = d.minDeltaT == -1 ? minTs : d.minDeltaT;
  r.initDt = d.initDeltaT;
  if (r.initDt < d.minDeltaT || r.initDt < 1e-14) r.initDt = d.minDeltaT;
  in.guessSS = d.guessSS;
  in.haveStateVals = !d.stateVals.empty();
  return SC_OK;
}

// Steady-state initial conditions (strikeSlip_linearElastic_qd.cpp:895-1006,
// Fault::guessSS fault.cpp:633-709): traction from the steady-state friction at
// vL imposed on a Neumann fault, solve, shift u >= 0, take the remote-boundary
// displacement as bcRShift and the fault displacement as the initial slip.
int solve_ss(QdDriver& q, std::vector<double>& psi, bool haveStateVals, std::vector<double>& slip,
             std::vector<double>& V) {
  const int Ny = q.Ny, Nz = q.Nz, N = q.N;
  sc::Ctx& c = q.ctx->c;
  std::vector<double> a(Nz), b(Nz), sN(Nz);
  SC_HIP(hipMemcpy(a.data(), q.a_d, Nz * 8, hipMemcpyDeviceToHost));
  SC_HIP(hipMemcpy(b.data(), q.b_d, Nz * 8, hipMemcpyDeviceToHost));
  SC_HIP(hipMemcpy(sN.data(), q.sN_d, Nz * 8, hipMemcpyDeviceToHost));
  const double vL = q.desc.vL, v0 = q.desc.v0, f0 = q.desc.f0;
  V.assign(Nz, vL);
  if (!haveStateVals)
    for (int k = 0; k < Nz; ++k) psi[k] = f0 - b[k] * std::log(std::fabs(vL) / v0);
  std::vector<double> tau(Nz);
  for (int k = 0; k < Nz; ++k) tau[k] = sN[k] * a[k] * std::asinh((double)0.5 * vL * std::exp(psi[k] / a[k]) / v0);
  const sc_grid_desc g = q.desc.grid;
  int e = sc_change_bc_types(q.ctx, g.bcR, g.bcT, SC_NEUMANN, g.bcB);
  if (e) return e;
  // the traction is the fault's datum; bcR is used as it stands (0 at set-up,
  // linearElastic.cpp:311-312, 383; bcRShift after an earlier solveSS)
  SC_HIP(hipMemcpy(q.bcL_d, tau.data(), Nz * 8, hipMemcpyHostToDevice));
  e = launch_set_rhs(c, q.rhs_d, q.bcL_d, q.bcR_d, q.bcT_d, q.bcB_d, c.stream);
  if (e) return e;
  sc_solve_stats st;
  e = solve(c, q.rhs_d, q.u_d, q.desc.solve, st);
  if (e) return e;
  std::vector<double> u(N);
  SC_HIP(hipMemcpy(u.data(), q.u_d, (size_t)N * 8, hipMemcpyDeviceToHost));
  double mn = *std::min_element(u.begin(), u.end());
  if (mn < 0) {
    mn = std::fabs(mn);
    for (auto& x : u) x += mn;
  }
  SC_HIP(hipMemcpy(q.u_d, u.data(), (size_t)N * 8, hipMemcpyHostToDevice));
  // solveSSb (:974-998): bcRShift and bcR from the remote boundary's
  // displacement, slip and bcL from the fault's
  std::vector<double> shift(u.begin() + (size_t)(Ny - 1) * Nz, u.end());
  SC_HIP(hipMemcpy(q.shift_d, shift.data(), Nz * 8, hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(q.bcR_d, shift.data(), Nz * 8, hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(q.bcL_d, u.data(), Nz * 8, hipMemcpyHostToDevice));
  slip.resize(Nz);
  for (int k = 0; k < Nz; ++k) slip[k] = u[k] * q.desc.faultTypeScale;
  return sc_change_bc_types(q.ctx, g.bcR, g.bcT, g.bcL, g.bcB);
}

}  // namespace sc

// ---------------------------------------------------------------- C ABI ---
using sc::fail;

extern "C" int sc_qd_create_from_file(const char* in_path, sc_qd** out) {
  if (!in_path || !out) return fail(SC_ERR_ARG, "sc_qd_create_from_file: null argument");
  sc::Deck d;
  int e = sc::load_deck(in_path, d);
  if (e) return e;
  sc::QdInputs in;
  e = sc::build_from_deck(d, in);
  if (e) return e;
  e = sc_qd_create(&in.desc, in.mu.data(), in.y.data(), in.z.data(), in.a.data(), in.b.data(),
                   in.Dc.data(), in.sN.data(), in.eta.data(), in.locked.data(),
                   in.prestress.data(), in.bcRShift.data(), out);
  if (e) return e;
  sc::QdDriver& q = *(*out)->q;
  q.file_run = in.run;
  q.have_file_run = true;
  if (d.haveOutputDir) {
    // the deck's output series and checkpoints; a checkpoint in the output
    // directory replaces the initial conditions (and solveSS) with its state
    const sc_qd_output o{d.outputDir.c_str(), d.stride1D, d.stride2D, d.ckpt, d.interval};
    int n = 0;
    e = q.set_output(&o);
    if (!e) e = q.load_checkpoint(&q.file_run, &n);
    if (e || n > 0) {
      if (e) {
        sc_qd_destroy(*out);
        *out = nullptr;
      }
      return e;
    }
  }
  std::vector<double> V(d.Nz, 0.0);
  if (in.guessSS) {
    e = sc::solve_ss(q, in.psi, in.haveStateVals, in.slip, V);
    if (e) {
      sc_qd_destroy(*out);
      *out = nullptr;
      return e;
    }
  }
  return sc_qd_set_state(*out, in.psi.data(), in.slip.data(), V.data());
}

extern "C" int sc_qd_file_run(sc_qd* h, sc_qd_run* run_out) {
  if (!h || !run_out) return fail(SC_ERR_ARG, "null argument");
  if (!h->q->have_file_run) return fail(SC_ERR_ARG, "driver was not created from an input deck");
  *run_out = h->q->file_run;
  return SC_OK;
}

extern "C" int sc_qd_solve_ss(sc_qd* h, double* psi_h, int have_state_vals, double* slip_h) {
  if (!h || !psi_h) return fail(SC_ERR_ARG, "null argument");
  sc::QdDriver& q = *h->q;
  std::vector<double> psi(psi_h, psi_h + q.Nz), slip, V;
  int e = sc::solve_ss(q, psi, have_state_vals != 0, slip, V);
  if (e) return e;
  std::memcpy(psi_h, psi.data(), sizeof(double) * q.Nz);
  if (slip_h) std::memcpy(slip_h, slip.data(), sizeof(double) * q.Nz);
  return sc_qd_set_state(h, psi.data(), slip.data(), V.data());
}

// Host-only utilities (no device work): deck parsing and Vec files.
extern "C" int sc_in_get(const char* in_path, const char* key, char* buf, int buflen) {
  if (!in_path || !key || !buf || buflen <= 0) return fail(SC_ERR_ARG, "null argument");
  bool ok;
  auto lines = sc::read_in_file(in_path, &ok);
  if (!ok) return fail(SC_ERR_IO, std::string("cannot read ") + in_path);
  int found = 0;
  for (const auto& l : lines)
    if (l.var == key) {
      std::snprintf(buf, buflen, "%s", l.rhs.c_str());
      found = 1;
    }
  return found ? SC_OK : fail(SC_ERR_ARG, std::string("key not found: ") + key);
}

extern "C" int sc_in_get_vector(const char* in_path, const char* key, double* vals, int cap,
                                int* n) {
  if (!in_path || !key || !n) return fail(SC_ERR_ARG, "null argument");
  bool ok;
  auto lines = sc::read_in_file(in_path, &ok);
  if (!ok) return fail(SC_ERR_IO, std::string("cannot read ") + in_path);
  std::vector<double> v;
  for (const auto& l : lines)
    if (l.var == key) sc::parse_vector(l.rhsFull, v);
  *n = (int)v.size();
  for (int i = 0; i < (int)v.size() && i < cap; ++i) vals[i] = v[i];
  return SC_OK;
}

extern "C" int sc_vec_read(const char* path, int rec, double* vals, int cap, int* n) {
  if (!path || !n) return fail(SC_ERR_ARG, "null argument");
  std::vector<std::vector<double>> recs;
  int nr = sc::read_petsc_vecs(path, recs);
  if (nr < 0) return fail(SC_ERR_IO, std::string("not a PETSc binary Vec file: ") + path);
  if (rec < 0 || rec >= nr) return fail(SC_ERR_ARG, "record index out of range");
  *n = (int)recs[rec].size();
  for (int i = 0; i < *n && i < cap; ++i) vals[i] = recs[rec][i];
  return SC_OK;
}

extern "C" int sc_vec_count(const char* path, int* nrec) {
  if (!path || !nrec) return fail(SC_ERR_ARG, "null argument");
  std::vector<std::vector<double>> recs;
  int nr = sc::read_petsc_vecs(path, recs);
  if (nr < 0) return fail(SC_ERR_IO, std::string("not a PETSc binary Vec file: ") + path);
  *nrec = nr;
  return SC_OK;
}

extern "C" int sc_vec_write(const char* path, const double* vals, int n, int append) {
  if (!path || (!vals && n > 0) || n < 0) return fail(SC_ERR_ARG, "bad argument");
  if (sc::write_petsc_vec(path, vals, n, append != 0)) return fail(SC_ERR_IO, std::string("cannot write ") + path);
  return SC_OK;
}
