// Internal declarations of the quasi-dynamic driver.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sc_internal.hpp"

namespace sc {

// Device vectors of one integrated variable (odeSolver.cpp keeps the same set
// per map key: var, dvar, stage states k_i, stage rates f_i, y_lo/y_hi).
struct VecSet {
  int n = 0;  // length: Nz for psi/slip, N for the viscous strains
  double* var = nullptr;
  double* dvar = nullptr;
  double* k[7] = {};
  double* f[7] = {};
  double* lo = nullptr;
  double* hi = nullptr;
};

struct RkState {
  int ord = 3, control = 1, normType = 0, stepCount = 0;
  double tol = 1e-8, kappa = 0.9, minDt = 0, maxDt = 1e30, deltaT = 0, currT = 0;
  double errA[2] = {0, 0};
};

// Power-law viscoelastic terms (StrikeSlip_PowerLaw_qd with dislocation creep)
struct PowerLawState {
  double *T = nullptr, *A = nullptr, *B = nullptr, *n = nullptr;  // N each
  double cap = 1e30;
  double *effVisc = nullptr, *sxy = nullptr;                       // last d_dt
};

struct QdDriver {
  sc_qd_desc desc{};
  sc_ctx* ctx = nullptr;
  int Ny = 0, Nz = 0, N = 0;
  std::vector<void*> owned;
  double *a_d = nullptr, *b_d = nullptr, *Dc_d = nullptr, *sN_d = nullptr, *eta_d = nullptr;
  double *locked_d = nullptr, *prestress_d = nullptr, *shift_d = nullptr;
  double *bcL_d = nullptr, *bcR_d = nullptr, *bcT_d = nullptr, *bcB_d = nullptr;
  double *rhs_d = nullptr, *u_d = nullptr;
  double *tauQS_d = nullptr, *tau_d = nullptr, *strength_d = nullptr, *V_d = nullptr;
  double* err_d = nullptr;              // 3 error sums per selected field
  double* err_part_d = nullptr;         // block partials of the error sums
  static constexpr int kErrBlocks = 512;
  int* status_d = nullptr;
  void* host_scratch = nullptr;
  std::map<std::string, VecSet> vars;
  int errMask = 3;                      // RK error fields: 1 psi, 2 slip, 4 gVxy, 8 gVxz
  std::unique_ptr<PowerLawState> pl;    // set for the power-law mediator
  sc_qd_counters counters{};
  sc_qd_run file_run{};
  bool have_file_run = false;

  ~QdDriver();
  double* dalloc(size_t n, const double* init_h);
  int init(const sc_qd_desc& d, const double* mu, const double* y, const double* z,
           const double* a, const double* b, const double* Dc, const double* sN,
           const double* eta, const double* locked, const double* prestress,
           const double* bcRShift);
  int add_var(const char* key, int n);
  int init_powerlaw(const double* T, const double* A, const double* B, const double* n,
                    double cap, int errMask);
  // d_dt of the mediator; gxy.. are the viscous strains and their rates (power law only)
  int d_dt(double t, const double* psi, const double* slip, double* dpsi, double* dslip,
           const double* gxy = nullptr, const double* gxz = nullptr, double* dgxy = nullptr,
           double* dgxz = nullptr);
  int d_dt_vars(double t);              // on var -> dvar of every integrated variable
  int check_status();
  double step_size(const RkState& r, double totErr) const;
  int error(const RkState& r, double* totErr);
  int lincomb_all(const char* out_sel, int stage, int nt, const int* fidx, const double* c);
  int stage(int st, double t);
  int integrate(const sc_qd_run& run, double* tEnd, int* steps, double* dts, int dtsCap);
};

}  // namespace sc

struct sc_qd {
  std::unique_ptr<sc::QdDriver> q;
};
