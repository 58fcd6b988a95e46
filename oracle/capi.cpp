// ctypes entry points of the CPU oracle (test infrastructure only; see
// oracle.hpp).  Loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg; never by the product library.
#include <chrono>
#include <cstring>

#include "oracle.hpp"
#include "physics.hpp"

using namespace orc;

namespace {
struct OrcSbp {
  Sbp2D s;
  BandChol chol;
  bool chol_ok = false;
};
const Csr* pick(const OrcSbp* o, int which) {
  switch (which) {
    case 0: return &o->s.A;
    case 1: return &o->s.H;
    case 2: return &o->s.Hinv;
    case 3: return &o->s.Dy_Iz;
    case 4: return &o->s.Iy_Dz;
    case 5: return &o->s.D2;
    case 6: return &o->s.rhsL;
    case 7: return &o->s.rhsR;
    case 8: return &o->s.rhsT;
    case 9: return &o->s.rhsB;
    case 10: return &o->s.Jinv;
    default: return nullptr;
  }
}
void dense(const Sp1& m, int n, double* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(double) * n * n);
  for (auto& r : m.m)
    for (auto& c : r.second) out[(size_t)r.first * n + c.first] = c.second;
}
}  // namespace

extern "C" {

void* orc_sbp_new(int order, int Ny, int Nz, double Ly, double Lz, int var_grid, int compat,
                  int laplace, int bcR, int bcT, int bcL, int bcB, const double* mu,
                  const double* y, const double* z) {
  SbpDesc d;
  d.order = order;
  d.Ny = Ny;
  d.Nz = Nz;
  d.Ly = Ly;
  d.Lz = Lz;
  d.var_grid = var_grid != 0;
  d.compat = compat ? "compatible" : "fullyCompatible";
  d.laplace = laplace == 0 ? "yz" : (laplace == 1 ? "y" : "z");
  d.bcR = bcR;
  d.bcT = bcT;
  d.bcL = bcL;
  d.bcB = bcB;
  auto* o = new OrcSbp;
  if (sbp2d_build(o->s, d, mu, y, z)) {
    delete o;
    return nullptr;
  }
  return o;
}
void orc_sbp_free(void* h) { delete (OrcSbp*)h; }
int orc_sbp_change_bc(void* h, int bcR, int bcT, int bcL, int bcB) {
  auto* o = (OrcSbp*)h;
  o->chol_ok = false;
  return sbp2d_change_bc(o->s, bcR, bcT, bcL, bcB);
}
long orc_sbp_nnz(void* h, int which) {
  const Csr* m = pick((OrcSbp*)h, which);
  return m ? (long)m->rp[m->nr] : -1;
}
int orc_sbp_shape(void* h, int which, int* nr, int* nc) {
  const Csr* m = pick((OrcSbp*)h, which);
  if (!m) return 1;
  *nr = m->nr;
  *nc = m->nc;
  return 0;
}
void orc_sbp_csr(void* h, int which, long* rp, int* ci, double* v) {
  const Csr* m = pick((OrcSbp*)h, which);
  for (int i = 0; i <= m->nr; ++i) rp[i] = (long)m->rp[i];
  std::memcpy(ci, m->ci.data(), sizeof(int) * m->ci.size());
  std::memcpy(v, m->v.data(), sizeof(double) * m->v.size());
}
void orc_sbp_apply(void* h, const double* in, double* out) { spmv(((OrcSbp*)h)->s.A, in, out); }
void orc_sbp_set_rhs(void* h, double* rhs, const double* bcL, const double* bcR, const double* bcT,
                     const double* bcB) {
  sbp2d_set_rhs(((OrcSbp*)h)->s, rhs, bcL, bcR, bcT, bcB);
}
void orc_sbp_muxdy(void* h, const double* in, double* out) { sbp2d_muxDy(((OrcSbp*)h)->s, in, out); }
void orc_sbp_muxdz(void* h, const double* in, double* out) { sbp2d_muxDz(((OrcSbp*)h)->s, in, out); }
void orc_sbp_dy(void* h, const double* in, double* out) { sbp2d_Dy(((OrcSbp*)h)->s, in, out); }
void orc_sbp_dz(void* h, const double* in, double* out) { sbp2d_Dz(((OrcSbp*)h)->s, in, out); }
void orc_sbp_dyxmu(void* h, const double* in, double* out) { sbp2d_Dyxmu(((OrcSbp*)h)->s, in, out); }
void orc_sbp_dzxmu(void* h, const double* in, double* out) { sbp2d_Dzxmu(((OrcSbp*)h)->s, in, out); }
void orc_sbp_bnd(void* h, int which, const double* in, double* out) {
  sbp2d_bnd(((OrcSbp*)h)->s, which, in, out);
}
void orc_sbp_h(void* h, const double* in, double* out) { spmv(((OrcSbp*)h)->s.H, in, out); }
void orc_sbp_hinv(void* h, const double* in, double* out) { spmv(((OrcSbp*)h)->s.Hinv, in, out); }
void orc_sbp_diag(void* h, double* out) {
  auto d = diagonal(((OrcSbp*)h)->s.A);
  std::memcpy(out, d.data(), sizeof(double) * d.size());
}
void orc_sbp_metric(void* h, double* yq, double* zr) {
  auto* o = (OrcSbp*)h;
  auto a = diagonal(o->s.yq), b = diagonal(o->s.zr);
  std::memcpy(yq, a.data(), sizeof(double) * a.size());
  std::memcpy(zr, b.data(), sizeof(double) * b.size());
}
void orc_sbp_scalars(void* h, double* out) {
  auto& s = ((OrcSbp*)h)->s;
  out[0] = s.dy; out[1] = s.dz; out[2] = s.alphaDy; out[3] = s.alphaDz;
  out[4] = s.h11y; out[5] = s.h11z; out[6] = s.alphaT;
}
int orc_sbp1d(int order, int N, double scale, int compat, double* H, double* Hinv, double* D1,
              double* BS, double* D2, double* C2, double* D3, double* D4, double* C3, double* C4) {
  Sbp1D o(N);
  int e = sbp_spmat(order, N, scale, o, compat ? "compatible" : "fullyCompatible");
  if (e) return e;
  if (order == 2) e = sbp_spmat2(N, scale, o);
  else e = sbp_spmat4(N, scale, o);
  if (e) return e;
  dense(o.H, N, H); dense(o.Hinv, N, Hinv); dense(o.D1, N, D1); dense(o.BS, N, BS);
  dense(o.D2, N, D2); dense(o.C2, N, C2); dense(o.D3, N, D3); dense(o.D4, N, D4);
  dense(o.C3, N, C3); dense(o.C4, N, C4);
  return 0;
}
int orc_pcg(void* h, const double* b, double* u, double rtol, int maxit, double* relres) {
  PcgStats st = pcg(((OrcSbp*)h)->s.A, b, u, rtol, maxit);
  if (relres) *relres = st.relres;
  return st.iters;
}
int orc_chol_solve(void* h, const double* b, double* u) {
  auto* o = (OrcSbp*)h;
  if (!o->chol_ok) {
    if (o->chol.factor(o->s.A)) return 1;
    o->chol_ok = true;
  }
  o->chol.solve(b, u);
  return 0;
}

double orc_strength(double sN, double psi, double V, double a, double v0) { return strength_psi(sN, psi, V, a, v0); }
double orc_aging(double psi, double V, double b, double f0, double v0, double Dc) { return agingLaw_psi(psi, V, b, f0, v0, Dc); }
double orc_sliplaw(double psi, double V, double a, double b, double f0, double v0, double Dc) { return slipLaw_psi(psi, V, a, b, f0, v0, Dc); }
int orc_vel_qd(int N, const double* eta, const double* tauQS, const double* sN, const double* psi,
               const double* a, const double* locked, double v0, double vL, double tol, int maxIts,
               int jac_mode, double* V, long* its) {
  return computeVel_qd(N, eta, tauQS, sN, psi, a, locked, v0, vL, tol, maxIts, jac_mode, V, its);
}

// ------------------------------------------------------------- QD run ----
struct OrcQdCfg {
  int order, Ny, Nz, var_grid;
  double Ly, Lz;
  int bcR, bcT, bcL, bcB;          // material SAT types
  int remoteLoadingR;
  double faultTypeScale, vL, f0, v0, rootTol;
  int stateLaw, jac_mode, solver;
  double cg_rtol;
};
struct OrcQdRun {
  int integrator;  // 0 RK32, 1 RK43
  int control;     // 0 P, 1 PID
  int normType;    // 0 L2_absolute, 1 L2_relative
  int maxSteps;
  double initTime, maxTime, initDt, minDt, maxDt, tol;
  int nReplay;
  const double* replayDt;
  double currErr, prevErr;  // PID error history at the start (odeSolver errA[0], errA[1])
};

}  // extern "C"

namespace {
QdModel* setup_qd(QdModel* m, const OrcQdCfg* c, const double* mu, const double* y,
                  const double* z, const double* a, const double* b, const double* Dc,
                  const double* sN, const double* eta, const double* locked,
                  const double* prestress, const double* bcRShift) {
  QdConfig q;
  q.sbp.order = c->order;
  q.sbp.Ny = c->Ny;
  q.sbp.Nz = c->Nz;
  q.sbp.Ly = c->Ly;
  q.sbp.Lz = c->Lz;
  q.sbp.var_grid = c->var_grid != 0;
  q.sbp.bcR = c->bcR;
  q.sbp.bcT = c->bcT;
  q.sbp.bcL = c->bcL;
  q.sbp.bcB = c->bcB;
  q.remoteLoadingR = c->remoteLoadingR;
  q.faultTypeScale = c->faultTypeScale;
  q.vL = c->vL;
  q.f0 = c->f0;
  q.v0 = c->v0;
  q.rootTol = c->rootTol;
  q.stateLaw = c->stateLaw;
  q.jac_mode = c->jac_mode;
  q.solver = c->solver;
  q.cg_rtol = c->cg_rtol;
  int Nz = c->Nz;
  m->a.assign(a, a + Nz);
  m->b.assign(b, b + Nz);
  m->Dc.assign(Dc, Dc + Nz);
  m->sN.assign(sN, sN + Nz);
  m->eta.assign(eta, eta + Nz);
  if (locked) m->locked.assign(locked, locked + Nz);
  if (prestress) m->prestress.assign(prestress, prestress + Nz);
  if (m->init(q, mu, y, z)) {
    delete m;
    return nullptr;
  }
  if (bcRShift) m->bcRShift.assign(bcRShift, bcRShift + Nz);
  return m;
}
}  // namespace

extern "C" {

void* orc_qd_new(const OrcQdCfg* c, const double* mu, const double* y, const double* z,
                 const double* a, const double* b, const double* Dc, const double* sN,
                 const double* eta, const double* locked, const double* prestress,
                 const double* bcRShift) {
  return setup_qd(new QdModel, c, mu, y, z, a, b, Dc, sN, eta, locked, prestress, bcRShift);
}
void orc_qd_free(void* h) { delete (QdModel*)h; }

// ------------------------------------------------------ power-law QD -----
void* orc_pl_new(const OrcQdCfg* c, const double* mu, const double* y, const double* z,
                 const double* a, const double* b, const double* Dc, const double* sN,
                 const double* eta, const double* locked, const double* prestress,
                 const double* bcRShift, const double* T, const double* A, const double* B,
                 const double* n, double cap) {
  auto* m = (PowerLawModel*)setup_qd(new PowerLawModel, c, mu, y, z, a, b, Dc, sN, eta, locked,
                                     prestress, bcRShift);
  if (m && m->initPL(T, A, B, n, cap)) {
    delete m;
    return nullptr;
  }
  return m;
}
void orc_pl_free(void* h) { delete (PowerLawModel*)h; }
void orc_pl_source(void* h, const double* gxy, const double* gxz, double* out) {
  powerlaw_source(*(PowerLawModel*)h, gxy, gxz, out);
}
void orc_pl_sat(void* h, const double* u, const double* gL, const double* gR, double* out) {
  auto* m = (PowerLawModel*)h;
  powerlaw_sat(m->sbp, m->sbp.d.bcL == 0, m->sbp.d.bcR == 0, u, gL, gR, out);
}
// One d_dt; outputs (any may be null): rates, fault tau, u, sxy, effVisc.
int orc_pl_ddt(void* h, double t, const double* psi, const double* slip, const double* gxy,
               const double* gxz, double* V_inout, double* dpsi, double* dslip, double* dgxy,
               double* dgxz, double* u, double* sxy, double* effVisc) {
  auto* m = (PowerLawModel*)h;
  const int Nz = m->Nz, N = m->N;
  if (V_inout) m->V.assign(V_inout, V_inout + Nz);
  VarMap var, dvar;
  var["psi"].assign(psi, psi + Nz);
  var["slip"].assign(slip, slip + Nz);
  var["gVxy"].assign(gxy, gxy + N);
  var["gVxz"].assign(gxz, gxz + N);
  int e = m->d_dt(t, var, dvar);
  if (e) return e;
  if (V_inout) std::memcpy(V_inout, m->V.data(), sizeof(double) * Nz);
  if (dpsi) std::memcpy(dpsi, dvar["psi"].data(), sizeof(double) * Nz);
  if (dslip) std::memcpy(dslip, dvar["slip"].data(), sizeof(double) * Nz);
  if (dgxy) std::memcpy(dgxy, dvar["gVxy"].data(), sizeof(double) * N);
  if (dgxz) std::memcpy(dgxz, dvar["gVxz"].data(), sizeof(double) * N);
  if (u) std::memcpy(u, m->u.data(), sizeof(double) * N);
  if (sxy) std::memcpy(sxy, m->sxy.data(), sizeof(double) * N);
  if (effVisc) std::memcpy(effVisc, m->effVisc.data(), sizeof(double) * N);
  return 0;
}
// RK integration of (psi, slip, gVxy, gVxz) with the given error indices
// (errMask bit 0 psi, 1 slip, 2 gVxy, 3 gVxz); state arrays in/out.
int orc_pl_integrate(void* h, const OrcQdRun* r, int errMask, double* psi, double* slip,
                     double* gxy, double* gxz, double* V, double* tEnd, double* dts, int dtsCap) {
  auto* m = (PowerLawModel*)h;
  const int Nz = m->Nz, N = m->N;
  RkSolver rk(r->integrator == 0 ? 3 : 4, r->maxSteps, r->maxTime, r->initDt,
              r->control == 0 ? "P" : "PID");
  rk.initT = rk.currT = r->initTime;
  rk.errA[0] = r->currErr;
  rk.errA[1] = r->prevErr;
  rk.totTol = r->tol;
  rk.minDeltaT = r->minDt;
  rk.maxDeltaT = r->maxDt;
  rk.normType = r->normType == 0 ? "L2_absolute" : "L2_relative";
  const char* names[4] = {"psi", "slip", "gVxy", "gVxz"};
  for (int i = 0; i < 4; ++i)
    if (errMask & (1 << i)) rk.errInds.push_back(names[i]);
  rk.var["psi"].assign(psi, psi + Nz);
  rk.var["slip"].assign(slip, slip + Nz);
  rk.var["gVxy"].assign(gxy, gxy + N);
  rk.var["gVxz"].assign(gxz, gxz + N);
  if (r->nReplay > 0) {
    rk.force_dt = 1;
    rk.replayDt.assign(r->replayDt, r->replayDt + r->nReplay);
    if (rk.maxNumSteps > r->nReplay) rk.maxNumSteps = r->nReplay;
  }
  if (V) m->V.assign(V, V + Nz);
  int e = rk.integrate(m);
  if (e) return -e;
  std::memcpy(psi, rk.var["psi"].data(), sizeof(double) * Nz);
  std::memcpy(slip, rk.var["slip"].data(), sizeof(double) * Nz);
  std::memcpy(gxy, rk.var["gVxy"].data(), sizeof(double) * N);
  std::memcpy(gxz, rk.var["gVxz"].data(), sizeof(double) * N);
  if (V) std::memcpy(V, m->V.data(), sizeof(double) * Nz);
  if (tEnd) *tEnd = rk.currT;
  for (int i = 0; i < (int)rk.dtHistory.size() && i < dtsCap; ++i) dts[i] = rk.dtHistory[i];
  return rk.stepCount;
}

// One d_dt evaluation (the reference's StrikeSlip_LinearElastic_qd::d_dt).
// V is the root-solve warm start on entry (persisting fault state).
int orc_qd_ddt(void* h, double t, const double* psi, const double* slip, double* V_inout,
               double* dpsi, double* dslip, double* tau, double* u) {
  auto* m = (QdModel*)h;
  int Nz = m->Nz;
  if (V_inout) m->V.assign(V_inout, V_inout + Nz);
  VarMap var, dvar;
  var["psi"].assign(psi, psi + Nz);
  var["slip"].assign(slip, slip + Nz);
  int e = m->d_dt(t, var, dvar);
  if (e) return e;
  if (V_inout) std::memcpy(V_inout, m->V.data(), sizeof(double) * Nz);
  if (dpsi) std::memcpy(dpsi, dvar["psi"].data(), sizeof(double) * Nz);
  if (dslip) std::memcpy(dslip, dvar["slip"].data(), sizeof(double) * Nz);
  if (tau) std::memcpy(tau, m->tau.data(), sizeof(double) * Nz);
  if (u) std::memcpy(u, m->u.data(), sizeof(double) * m->N);
  return 0;
}

// Integrate from (psi, slip); on return psi/slip hold the final state, and
// dts[0..steps) the accepted step sizes.  Returns the number of steps or <0.
int orc_qd_integrate(void* h, const OrcQdRun* r, double* psi, double* slip, double* V,
                     double* tau, double* u, double* tEnd, double* dts, int dtsCap,
                     double* seconds) {
  auto* m = (QdModel*)h;
  int Nz = m->Nz;
  RkSolver rk(r->integrator == 0 ? 3 : 4, r->maxSteps, r->maxTime, r->initDt,
              r->control == 0 ? "P" : "PID");
  rk.initT = rk.currT = r->initTime;
  rk.errA[0] = r->currErr;
  rk.errA[1] = r->prevErr;
  rk.totTol = r->tol;
  rk.minDeltaT = r->minDt;
  rk.maxDeltaT = r->maxDt;
  rk.normType = r->normType == 0 ? "L2_absolute" : "L2_relative";
  rk.errInds = {"psi", "slip"};
  rk.var["psi"].assign(psi, psi + Nz);
  rk.var["slip"].assign(slip, slip + Nz);
  if (r->nReplay > 0) {
    rk.force_dt = 1;
    rk.replayDt.assign(r->replayDt, r->replayDt + r->nReplay);
    if (rk.maxNumSteps > r->nReplay) rk.maxNumSteps = r->nReplay;
  }
  if (V) m->V.assign(V, V + Nz);
  auto t0 = std::chrono::steady_clock::now();
  int e = rk.integrate(m);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  if (e) return -e;
  std::memcpy(psi, rk.var["psi"].data(), sizeof(double) * Nz);
  std::memcpy(slip, rk.var["slip"].data(), sizeof(double) * Nz);
  if (V) std::memcpy(V, m->V.data(), sizeof(double) * Nz);
  if (tau) std::memcpy(tau, m->tau.data(), sizeof(double) * Nz);
  if (u) std::memcpy(u, m->u.data(), sizeof(double) * m->N);
  if (tEnd) *tEnd = rk.currT;
  for (int i = 0; i < (int)rk.dtHistory.size() && i < dtsCap; ++i) dts[i] = rk.dtHistory[i];
  return rk.stepCount;
}
// solveSS on the model: psi in/out (in: the deck's psi, used when
// haveStateVals), outputs slip, V, bcRShift (Nz) and the shifted u (N); any
// output may be null.
int orc_qd_solve_ss(void* h, double* psi, int haveStateVals, double* slip, double* V,
                    double* bcRShift, double* u) {
  auto* m = (QdModel*)h;
  const int Nz = m->Nz;
  m->psi.assign(psi, psi + Nz);
  int e = m->solveSS(haveStateVals != 0);
  if (e) return e;
  std::memcpy(psi, m->psi.data(), sizeof(double) * Nz);
  if (slip) std::memcpy(slip, m->slip.data(), sizeof(double) * Nz);
  if (V) std::memcpy(V, m->V.data(), sizeof(double) * Nz);
  if (bcRShift) std::memcpy(bcRShift, m->bcRShift.data(), sizeof(double) * Nz);
  if (u) std::memcpy(u, m->u.data(), sizeof(double) * m->N);
  return 0;
}
long orc_qd_counter(void* h, int which) {
  auto* m = (QdModel*)h;
  return which == 0 ? m->nSolves : (which == 1 ? m->cgIters : m->rootIts);
}

// ------------------------------------------------------------------ FD ---
int orc_fd_propagate(void* h, double dt, const double* u, const double* uPrev, const double* rho,
                     const double* ay, double* uNext, double* d2u) {
  fd_propagate(((OrcSbp*)h)->s, dt, u, uPrev, rho, ay, uNext, d2u);
  return 0;
}
void orc_fd_ay(void* h, const double* cs, int outR, int outT, int outB, int faultL, double* ay) {
  fd_ay(((OrcSbp*)h)->s, cs, outR, outT, outB, faultL, ay);
}
void* orc_fdfault_new(int N, double v0, double f0, double vL, double rootTol, double fts,
                      int stateLaw, const double* a, const double* b, const double* Dc,
                      const double* sNEff, const double* rho, const double* alphay,
                      const double* tau0, const double* slip0) {
  auto* f = new FdFault;
  f->N = N; f->v0 = v0; f->f0 = f0; f->vL = vL; f->rootTol = rootTol; f->faultTypeScale = fts;
  f->stateLaw = stateLaw;
  f->a.assign(a, a + N); f->b.assign(b, b + N); f->Dc.assign(Dc, Dc + N);
  f->sNEff.assign(sNEff, sNEff + N); f->rho.assign(rho, rho + N); f->alphay.assign(alphay, alphay + N);
  f->tau0.assign(tau0, tau0 + N); f->slip0.assign(slip0, slip0 + N);
  f->psi.assign(N, 0); f->slip.assign(N, 0); f->V.assign(N, 0); f->u.assign(N, 0);
  f->uPrev.assign(N, 0); f->d2u.assign(N, 0); f->an.assign(N, 0); f->Phi.assign(N, 0);
  f->fricPen.assign(N, 0); f->strength.assign(N, 0);
  return f;
}
void orc_fdfault_free(void* h) { delete (FdFault*)h; }
int orc_fdfault_step(void* h, double dt, const double* d2u, double* V_inout, const double* psi,
                     const double* psiPrev, const double* slip, const double* slipPrev,
                     double* psiNext, double* slipNext, double* uFault) {
  auto* f = (FdFault*)h;
  std::memcpy(f->d2u.data(), d2u, sizeof(double) * f->N);
  std::memcpy(f->V.data(), V_inout, sizeof(double) * f->N);
  int e = f->step(dt, psi, psiPrev, slip, slipPrev, psiNext, slipNext);
  std::memcpy(V_inout, f->V.data(), sizeof(double) * f->N);
  if (uFault) std::memcpy(uFault, f->u.data(), sizeof(double) * f->N);
  return e;
}

// ------------------------------------------------------- QD <-> FD -----
struct OrcQdFdCfg {
  int fdBcR, fdBcT, fdBcL, fdBcB, outR, outT, outB, ayCompat;
  double deltaT_fd, CFL, trigger_qd2fd, trigger_fd2qd, limit_qd, limit_fd;
  int maxNumCycles, fdStop;
};
struct OrcQdFdRun {
  int integrator, control, normType, maxStepCount;
  double initTime, maxTime, minDt, maxDt, tol;
};
// qd: a model from orc_qd_new (not owned; freed by the caller after this one)
void* orc_qdfd_new(void* qd, const OrcQdFdCfg* c, const double* mu, const double* y, const double* z,
                   const double* rho, const double* cs) {
  auto* m = new QdFdModel;
  QdFdConfig cfg;
  cfg.fdBcR = c->fdBcR; cfg.fdBcT = c->fdBcT; cfg.fdBcL = c->fdBcL; cfg.fdBcB = c->fdBcB;
  cfg.outR = c->outR; cfg.outT = c->outT; cfg.outB = c->outB; cfg.ayCompat = c->ayCompat;
  cfg.deltaT_fd = c->deltaT_fd; cfg.CFL = c->CFL;
  cfg.trigger_qd2fd = c->trigger_qd2fd; cfg.trigger_fd2qd = c->trigger_fd2qd;
  cfg.limit_qd = c->limit_qd; cfg.limit_fd = c->limit_fd;
  cfg.maxNumCycles = c->maxNumCycles; cfg.fdStop = c->fdStop;
  if (m->init((QdModel*)qd, cfg, mu, y, z, rho, cs)) {
    delete m;
    return nullptr;
  }
  return m;
}
void orc_qdfd_free(void* h) { delete (QdFdModel*)h; }
double orc_qdfd_dt(void* h) { return ((QdFdModel*)h)->dtFd; }
void orc_qdfd_ay(void* h, double* ay) {
  auto* m = (QdFdModel*)h;
  std::memcpy(ay, m->ay.data(), sizeof(double) * m->N);
}
// psi, slip, V: initial QD state in, final state out; u (N): total
// displacement out; phases to log[0..cap) as (regime, stepStart, stepEnd,
// tStart, tEnd, R).  Returns the number of phases or < 0.
int orc_qdfd_integrate(void* h, const OrcQdFdRun* r, double* psi, double* slip, double* V, double* u,
                       double* tEnd, int* steps, int* cycles, int* regime, QdFdPhase* log, int cap) {
  auto* m = (QdFdModel*)h;
  const int Nz = m->Nz, N = m->N;
  m->qvar["psi"].assign(psi, psi + Nz);
  m->qvar["slip"].assign(slip, slip + Nz);
  m->qd->V.assign(V, V + Nz);
  QdFdRunCfg rc;
  rc.integrator = r->integrator; rc.control = r->control; rc.normType = r->normType;
  rc.maxStepCount = r->maxStepCount; rc.initTime = r->initTime; rc.maxTime = r->maxTime;
  rc.minDt = r->minDt; rc.maxDt = r->maxDt; rc.tol = r->tol;
  const int e = m->integrate(rc);
  if (e) return -e;
  if (m->inDynamic) {
    std::memcpy(psi, m->psiF[1].data(), sizeof(double) * Nz);
    std::memcpy(slip, m->slipF[1].data(), sizeof(double) * Nz);
    std::memcpy(V, m->ff.V.data(), sizeof(double) * Nz);
    for (int i = 0; i < N; ++i) u[i] = m->du[i] + 1.0 * m->u0[i];
  } else {
    std::memcpy(psi, m->qvar["psi"].data(), sizeof(double) * Nz);
    std::memcpy(slip, m->qvar["slip"].data(), sizeof(double) * Nz);
    std::memcpy(V, m->qd->V.data(), sizeof(double) * Nz);
    std::memcpy(u, m->qd->u.data(), sizeof(double) * N);
  }
  *tEnd = m->currT;
  *steps = m->stepCount;
  *cycles = m->cycleCount;
  *regime = m->inDynamic ? 1 : 0;
  const int n = (int)m->log.size();
  for (int i = 0; i < n && i < cap; ++i) log[i] = m->log[i];
  return n;
}

// Integrates y' = lam*y (two components: lam and lam2) with the RK32 / RK43
// solver over n fixed steps of size dt (replay mode: no step-size control) or,
// with n < 0, adaptively to finalT with tolerance tol.  Test hook for the
// stability polynomials and the controller.  Returns the step count.
namespace {
struct LinOde : OdeCtx {
  double lam, lam2;
  long calls = 0;
  int d_dt(double, const VarMap& var, VarMap& dvar) override {
    ++calls;
    const auto& y = var.at("y");
    auto& d = dvar["y"];
    d.assign(y.size(), 0.0);
    d[0] = lam * y[0];
    if (y.size() > 1) d[1] = lam2 * y[1];
    return 0;
  }
};
}  // namespace

int orc_rk_linear(int ord, double lam, double lam2, double dt, int n, double finalT, double tol,
                  double y0, double* y_out, double* t_out, long* calls) {
  LinOde ode;
  ode.lam = lam;
  ode.lam2 = lam2;
  RkSolver rk(ord, n >= 0 ? n : 1000000, n >= 0 ? dt * n * 2.0 : finalT, dt, "PID");
  rk.totTol = tol;
  rk.var["y"] = {y0, y0};
  rk.errInds = {"y"};
  if (n >= 0) {
    rk.force_dt = 1;
    rk.replayDt.assign((size_t)n, dt);
  }
  int e = rk.integrate(&ode);
  y_out[0] = rk.var["y"][0];
  y_out[1] = rk.var["y"][1];
  *t_out = rk.currT;
  *calls = ode.calls;
  return e ? -1 : rk.stepCount;
}

// Adaptive RK32/RK43 run of y' = lam*y (both components alike) from y0 with
// step dt0 to finalT, control 0 P / 1 PID: for each accepted step its size,
// its error estimate and the step size chosen after it.  Test hook for the
// controller's use of the error history.  Returns the accepted step count.
int orc_rk_linear_trace(int ord, double lam, double dt0, double finalT, double tol, double y0,
                        int control, double* dts, double* errs, double* newDts, int cap) {
  LinOde ode;
  ode.lam = ode.lam2 = lam;
  RkSolver rk(ord, 1000000, finalT, dt0, control == 0 ? "P" : "PID");
  rk.totTol = tol;
  rk.var["y"] = {y0, y0};
  rk.errInds = {"y"};
  if (rk.integrate(&ode)) return -1;
  const int n = (int)rk.dtHistory.size();
  for (int i = 0; i < n && i < cap; ++i) {
    dts[i] = rk.dtHistory[i];
    errs[i] = rk.errHistory[i];
    newDts[i] = rk.newDtHistory[i];
  }
  return n;
}

void orc_powerlaw(int N, const double* sxy, const double* sxz, const double* T, const double* A,
                  const double* B, const double* n, const double* mu, const double* sat,
                  double cap, double* effVisc, double* gxy_t, double* gxz_t) {
  powerlaw_rates(N, sxy, sxz, T, A, B, n, mu, sat, cap, effVisc, gxy_t, gxz_t);
}

}  // extern "C"
