// Oracle restatement of the fault, integrator, wave-step and power-law parts
// of the SCycle step (test infrastructure only; see oracle.hpp).
#pragma once
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "oracle.hpp"

namespace orc {

// ----------------------------------------------------------- rate-state --
double strength_psi(double sN, double psi, double V, double a, double v0);      // fault.cpp:2246-2250
double agingLaw_psi(double psi, double V, double b, double f0, double v0, double Dc);  // fault.cpp:1984-1995
double slipLaw_psi(double psi, double V, double a, double b, double f0, double v0, double Dc);  // fault.cpp:2072-2088

// BracketedNewton::findRoot (rootFinder.cpp:174-257).  resid(x, &f, &J).
// Returns 0 on convergence, 1 otherwise; *its receives the iteration count.
int bracketed_newton(const std::function<void(double, double*, double*)>& resid, double left,
                     double right, double x0, double atol, int maxIts, double* out, int* its);

// Quasi-dynamic slip-velocity solve over all fault nodes
// (ComputeVel_qd::computeVel fault.cpp:954-1015, getResid :1020-1059).
// jac_mode 0 reproduces the reference's derivative expression (A*V instead of
// A*B, fault.cpp:1051); 1 uses the exact derivative.  V is in/out (warm start).
int computeVel_qd(int N, const double* eta, const double* tauQS, const double* sN,
                  const double* psi, const double* a, const double* locked, double v0, double vL,
                  double rootTol, int maxIts, int jac_mode, double* V, long* rootIts);

// ---------------------------------------------------------------- RK -----
using VarMap = std::map<std::string, std::vector<double>>;
struct OdeCtx {
  virtual ~OdeCtx() {}
  virtual int d_dt(double t, const VarMap& var, VarMap& dvar) = 0;
  virtual int timeMonitor(double t, double dt, int stepCount, int& stop) { (void)t; (void)dt; (void)stepCount; (void)stop; return 0; }
};

// RK32 / RK43 with P/PID control (odeSolver.cpp:280-627, 780-1175).
struct RkSolver {
  int ord;                       // 3 (RK32) or 4 (RK43)
  std::string controlType = "PID";
  std::string normType = "L2_absolute";
  double initT = 0, finalT = 0, currT = 0, deltaT = 0, newDeltaT = 0;
  double totTol = 1e-9, kappa = 0.9, minDeltaT = 0, maxDeltaT = 0;
  int maxNumSteps = 0, stepCount = 0, numRejected = 0;
  int pidStep0 = 0;               // PID after four steps past this one
  double errA[2] = {0, 0};       // circular buffer: [0] latest, [1] previous
  std::vector<std::string> errInds;
  std::vector<double> scaleV;
  VarMap var, dvar;
  std::vector<double> dtHistory;  // accepted step sizes
  std::vector<double> errHistory, newDtHistory;  // their error estimates and the next step chosen
  int force_dt = 0;               // replay mode: use replayDt[stepCount-1], no control
  std::vector<double> replayDt;
  RkSolver(int order, int maxSteps, double finalT_, double dt0, const std::string& ctl)
      : ord(order), controlType(ctl), initT(0), finalT(finalT_), deltaT(dt0), maxDeltaT(finalT_),
        maxNumSteps(maxSteps) {}
  double computeStepSize(double totErr);
  double computeError(const VarMap& lo, const VarMap& hi);
  int integrate(OdeCtx* obj);
};

// ---------------------------------------------------------- QD model -----
struct QdConfig {
  SbpDesc sbp;
  double faultTypeScale = 2.0, vL = 1e-9, f0 = 0.6, v0 = 1e-6, rootTol = 1e-12;
  int maxNumIts = 10000;
  int stateLaw = 0;   // 0 aging, 1 slip
  int jac_mode = 0;
  int remoteLoadingR = 1;  // bcR == remoteLoading
  int solver = 0;          // 0 banded Cholesky, 1 Jacobi-PCG
  double cg_rtol = 1e-12;
  int cg_maxit = 200000;
};

// StrikeSlip_LinearElastic_qd + Fault_qd + LinearElastic
// (strikeSlip_linearElastic_qd.cpp:756-891; fault.cpp:717-918;
// linearElastic.cpp:465-509,711-738).
struct QdModel : OdeCtx {
  QdConfig c;
  Sbp2D sbp;
  BandChol chol;
  int Ny, Nz, N;
  std::vector<double> a, b, Dc, sN, eta, locked, prestress;
  std::vector<double> bcL, bcR, bcT, bcB, bcRShift, rhs, u, sxy;
  std::vector<double> tauQS, tau, V, strength, psi, slip;
  long rootIts = 0;
  long nSolves = 0, cgIters = 0;
  int init(const QdConfig& cfg, const double* mu, const double* y, const double* z);
  int solveMomentumBalance(double t);
  int d_dt(double t, const VarMap& var, VarMap& dvar) override;
  // solveSS + solveSSb (strikeSlip_linearElastic_qd.cpp:895-1006) with
  // Fault::guessSS (fault.cpp:633-709): psi (recomputed at vL unless
  // haveStateVals), V = vL, slip, bcL, bcR, bcRShift and u set as the
  // reference leaves them
  int solveSS(bool haveStateVals);
};

// ------------------------------------------------------------- FD -----
// propagateWaves (strikeSlip_linearElastic_fd.cpp:445-516) for the body.
void fd_propagate(const Sbp2D& s, double dt, const double* u, const double* uPrev,
                  const double* rho, const double* ay, double* uNext, double* d2u);
// Absorbing coefficient a_y (strikeSlip_linearElastic_fd.cpp:608-669).
void fd_ay(const Sbp2D& s, const double* cs, int outR, int outT, int outB, int faultL, double* ay);
// Fault_fd::d_dt minus the body scatter (fault.cpp:1436-1575, 1588-1879).
struct FdFault {
  int N;
  double v0, f0, vL, rootTol, faultTypeScale;
  int maxNumIts = 10000, stateLaw = 1;
  std::vector<double> a, b, Dc, sNEff, rho, alphay, tau0, locked, slip0;
  std::vector<double> psi, slip, V, u, uPrev, d2u, an, Phi, fricPen, strength;
  long rootIts = 0;
  int step(double dt, const double* psiCur, const double* psiPrev, const double* slipCur,
           const double* slipPrev, double* psiNext, double* slipNext);
};

// ----------------------------------------------------- QD <-> FD -----
// strikeSlip_linearElastic_qd_fd (strikeSlip_linearElastic_qd_fd.cpp): QD
// phases with a QdModel (QD material SAT types), FD phases with fd_propagate
// and FdFault on an operator with the FD material SAT types.
struct QdFdConfig {
  int fdBcR = NEUMANN, fdBcT = NEUMANN, fdBcL = NEUMANN, fdBcB = NEUMANN;
  int outR = 1, outT = 0, outB = 1;
  int ayCompat = 1;                       // computePenaltyVectors as it runs (:486-500)
  double deltaT_fd = -1, CFL = -1;
  double trigger_qd2fd = 1e-3, trigger_fd2qd = 1e-3, limit_qd = 1e-8, limit_fd = 1e-1;
  int maxNumCycles = 1000;
  int fdStop = 0;                          // 1: FD phases end on the switching test
};
struct QdFdRunCfg {
  int integrator = 0, control = 1, normType = 0, maxStepCount = 0;
  double initTime = 0, maxTime = 0, minDt = 1e-3, maxDt = 1e10, tol = 1e-8;
};
struct QdFdPhase {
  int regime, stepStart, stepEnd;
  double tStart, tEnd, R;
};
struct QdFdModel {
  QdModel* qd = nullptr;                   // not owned
  QdFdConfig c;
  Sbp2D fs;                                // the FD phase's operator
  FdFault ff;
  int Ny = 0, Nz = 0, N = 0;
  std::vector<double> rho, cs, ay, u0, du, duPrev, duNext, d2u;
  std::vector<double> psiF[3], slipF[3];   // FD fault levels: prev, cur, next
  VarMap qvar;                             // QD state (psi, slip)
  double dtFd = 0, currT = 0, lastR = 0;
  int stepCount = 0, cycleCount = 0;
  bool inDynamic = false, allowed = false, started = false;
  std::vector<QdFdPhase> log;
  int init(QdModel* q, const QdFdConfig& cfg, const double* mu, const double* y, const double* z,
           const double* rho, const double* cs);
  bool checkSwitch(double R);
  double qdRatio() const;
  double fdRatio() const;
  int qdPhase(const QdFdRunCfg& r, int regime, int maxSteps, double finalT, double minDt, double maxDt);
  int prepareFd2Qd();
  int prepareQd2Fd(const QdFdRunCfg& r);
  int fdPhase(const QdFdRunCfg& r);
  int fdStep();
  int integrate(const QdFdRunCfg& r);
};

// ------------------------------------------------------ power-law QD -----
// StrikeSlip_PowerLaw_qd::d_dt (strikeSlip_powerLaw_qd.cpp:633-686, 787-823)
// with PowerLaw (powerLaw.cpp) in its dislocation-creep form: integrated
// variables psi, slip (Nz) and the viscous strains gVxy, gVxz (N).
struct PowerLawModel : QdModel {
  Csr Bm, Cm;                          // PowerLaw::initializeMomBalMats (powerLaw.cpp:1239-1318)
  std::vector<double> T, Ad, Bd, nd;   // temperature, A, Q/R, n (N each)
  double effViscCap = 1e30;
  std::vector<double> visc, gTxy, gTxz, sxz, effVisc, sat;
  int initPL(const double* T, const double* A, const double* B, const double* n, double cap);
  int d_dt(double t, const VarMap& var, VarMap& dvar) override;
};
// B gxy + C gxz  (PowerLaw::computeViscStrainSourceTerms, powerLaw.cpp:1651-1672)
void powerlaw_source(const PowerLawModel& m, const double* gxy, const double* gxz, double* out);
// PowerLaw::computeViscousStrainRateSAT (powerLaw.cpp:1745-1800)
void powerlaw_sat(const Sbp2D& s, int bcLDirichlet, int bcRDirichlet, const double* u,
                  const double* gL, const double* gR, double* out);

// -------------------------------------------------------- power law -----
// DislocationCreep::computeInvEffVisc + PowerLaw::computeViscosity +
// computeViscStrainRates (powerLaw.cpp:438-474, 1674-1799).
void powerlaw_rates(int N, const double* sxy, const double* sxz, const double* T,
                    const double* A, const double* B, const double* n, const double* mu,
                    const double* sat, double effViscCap, double* effVisc, double* gxy_t,
                    double* gxz_t);

}  // namespace orc
