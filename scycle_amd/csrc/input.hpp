// Input-deck structures (reference `.in` format).
#pragma once
#include <string>
#include <vector>

#include "sc_internal.hpp"

namespace sc {

struct InLine {
  std::string var, rhs, rhsFull;
};

std::vector<InLine> read_in_file(const std::string& path, bool* ok);
void parse_vector(const std::string& str, std::vector<double>& vec);
void parse_vector(const std::string& str, std::vector<std::string>& vec);
void set_vec(std::vector<double>& out, const std::vector<double>& coord,
             const std::vector<double>& vals, const std::vector<double>& depths);
int read_petsc_vecs(const std::string& path, std::vector<std::vector<double>>& recs);
int write_petsc_vec(const std::string& path, const double* v, int n, bool append);

// Keys read by Domain, LinearElastic, Fault and StrikeSlip_LinearElastic_qd,
// with the reference's defaults.
struct Deck {
  int order = 4, Ny = -1, Nz = -1;
  double Ly = -1, Lz = -1, bCoordTrans = -1, vL = 1e-9;
  std::string inputDir = "unspecified_", outputDir = "data/";
  std::string compat = "fullyCompatible", gridSpacingType = "variableGridSpacing";
  std::string bulk = "linearElastic", momBal = "quasidynamic";
  std::vector<double> muVals, muDepths, rhoVals, rhoDepths;
  std::vector<double> DcVals, DcDepths, sNVals, sNDepths, aVals, aDepths, bVals, bDepths;
  std::vector<double> stateVals, stateDepths, lockedVals, lockedDepths;
  double sNcap = 1e14, sNfloor = 0., rootTol = 1e-12, f0 = 0.6, v0 = 1e-6;
  std::string stateLaw = "agingLaw";
  int guessSS = 0;
  std::string timeIntegrator = "RK43", timeControlType = "PID", normType = "L2_absolute";
  int maxStepCount = 100000000;
  double initTime = 0, maxTime = 1e15, minDeltaT = -1, maxDeltaT = 1e10, initDeltaT = 1e-3,
         timeStepTol = 1e-8;
  std::vector<std::string> timeIntInds;
  std::string bcR = "remoteLoading", bcT = "freeSurface", bcL = "symmFault", bcB = "freeSurface";
  std::string linSolver = "MUMPSCHOLESKY";
  double kspTol = 1e-10;
  // output and checkpoints (strikeSlip_linearElastic_qd.cpp:14,167-168;
  // domain.cpp:17,187-188)
  bool haveOutputDir = false;
  int stride1D = 1, stride2D = 1, ckpt = 0, interval = 10000;
};

struct QdInputs {
  sc_qd_desc desc{};
  sc_qd_run run{};
  std::vector<double> mu, y, z, a, b, Dc, sN, eta, locked, prestress, bcRShift, psi, slip;
  int guessSS = 0;
  bool haveStateVals = false;
};

int load_deck(const std::string& path, Deck& d);
int build_from_deck(const Deck& d, QdInputs& in);

}  // namespace sc
