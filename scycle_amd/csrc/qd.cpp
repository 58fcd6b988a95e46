// Quasi-dynamic mediator + explicit adaptive Runge-Kutta driver on the GPU.
//
// QdDriver plays StrikeSlip_LinearElastic_qd (strikeSlip_linearElastic_qd.cpp)
// with its LinearElastic and Fault_qd members; integrate() is the host loop of
// OdeSolver RK32/RK43 (odeSolver.cpp:505-627, 973-1175) with every vector on
// the device.  Per d_dt: boundary data -> setRhs -> solve (warm start from the
// previous displacement, as KSPSetInitialGuessNonzero, or from the projection
// onto earlier solutions) -> sigma_xy on the fault line -> rate-and-state
// kernel.  Inside integrate() a solve whose projected guess is accepted on
// the device needs no host round trip: the host reads back one block per RK
// attempt (error sums, fault status, missed-guess flag) and repeats the
// attempt with checked solves in the rare case that a guess missed.
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "qd_internal.hpp"

namespace sc {

int launch_err(hipStream_t s, int n, const double* lo, const double* hi, double* out3,
               double* partials);

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

QdDriver::~QdDriver() {
  if (ctx) {
    (void)hipStreamSynchronize(ctx->c.stream);
    for (void* p : owned) (void)hipFree(p);
    if (host_scratch) (void)hipHostFree(host_scratch);
    if (ctl_d) (void)hipFree(ctl_d);
    if (ctl_h) (void)hipHostFree(ctl_h);
    if (ctlDts_d) (void)hipFree(ctlDts_d);
    drop_stage_graph();
    stage_bufs_free(stageBufs);
    if (adjB) (void)hipFree(adjB);
    if (adjCtx) sc_ctx_destroy(adjCtx);
    sc_ctx_destroy(ctx);
  }
}

void QdDriver::drop_stage_graph() {
  if (stageExec) (void)hipGraphExecDestroy(stageExec);
  if (stageGraph) (void)hipGraphDestroy(stageGraph);
  if (tailExec) (void)hipGraphExecDestroy(tailExec);
  if (tailGraph) (void)hipGraphDestroy(tailGraph);
  stageExec = nullptr;
  stageGraph = nullptr;
  tailExec = nullptr;
  tailGraph = nullptr;
  stageKey.clear();
}

double* QdDriver::dalloc(size_t n, const double* init_h) {
  double* p = nullptr;
  if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
  owned.push_back(p);
  if (init_h) (void)hipMemcpy(p, init_h, n * sizeof(double), hipMemcpyHostToDevice);
  else (void)hipMemset(p, 0, n * sizeof(double));
  return p;
}

int QdDriver::init(const sc_qd_desc& d, const double* mu, const double* y, const double* z,
                   const double* a, const double* b, const double* Dc, const double* sN,
                   const double* eta, const double* locked, const double* prestress,
                   const double* bcRShift) {
  desc = d;
  int e = sc_ctx_create(&d.grid, y, z, mu, &ctx);
  if (e) return e;
  Ny = d.grid.Ny;
  Nz = d.grid.Nz;
  N = Ny * Nz;
  mu_host.assign(mu, mu + N);
  if (y) y_host.assign(y, y + N);
  if (z) z_host.assign(z, z + N);
  a_d = dalloc(Nz, a);
  b_d = dalloc(Nz, b);
  Dc_d = dalloc(Nz, Dc);
  sN_d = dalloc(Nz, sN);
  eta_d = dalloc(Nz, eta);
  locked_d = dalloc(Nz, locked);
  prestress_d = dalloc(Nz, prestress);
  shift_d = dalloc(Nz, bcRShift);
  bcL_d = dalloc(Nz, nullptr);
  bcR_d = dalloc(Nz, nullptr);
  bcT_d = dalloc(Ny, nullptr);
  bcB_d = dalloc(Ny, nullptr);
  rhs_d = dalloc(N, nullptr);
  u_d = dalloc(N, nullptr);
  tauQS_d = dalloc(Nz, nullptr);
  tau_d = dalloc(Nz, nullptr);
  strength_d = dalloc(Nz, nullptr);
  V_d = dalloc(Nz, nullptr);
  // one read-back block: the error sums, then the fault status, the
  // missed-guess flag and the attempt's-end status pair
  err_d = dalloc(kChkDoubles, nullptr);
  if (!err_d) return fail(SC_ERR_HIP, "device allocation failed");
  status_d = (int*)(err_d + kErrSlots);
  miss_d = status_d + 1;
  endStatus_d = status_d + 2;
  for (const char* key : {"psi", "slip"})
    if ((e = add_var(key, Nz))) return e;
  err_part_d = dalloc((size_t)3 * kErrBlocks, nullptr);
  if (hipHostMalloc(&host_scratch, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(SC_ERR_HIP, "hipHostMalloc");
  for (void* p : owned)
    if (!p) return fail(SC_ERR_HIP, "device allocation failed");
  return SC_OK;
}

int QdDriver::add_var(const char* key, int n) {
  VecSet& v = vars[key];
  v.n = n;
  v.var = dalloc(n, nullptr);
  v.dvar = dalloc(n, nullptr);
  for (int i = 0; i < 7; ++i) {
    v.k[i] = dalloc(n, nullptr);
    v.f[i] = dalloc(n, nullptr);
  }
  v.lo = dalloc(n, nullptr);
  v.hi = dalloc(n, nullptr);
  for (void* p : owned)
    if (!p) return fail(SC_ERR_HIP, "device allocation failed");
  return SC_OK;
}

int QdDriver::init_powerlaw(const double* T, const double* A, const double* B, const double* n,
                            double cap, int mask) {
  pl.reset(new PowerLawState);
  pl->T = dalloc(N, T);
  pl->A = dalloc(N, A);
  pl->B = dalloc(N, B);
  pl->n = dalloc(N, n);
  pl->K = dalloc(N, nullptr);
  pl->effVisc = dalloc(N, nullptr);
  pl->sxy = dalloc(N, nullptr);
  pl->cap = cap;
  errMask = mask;
  if (!pl->T || !pl->A || !pl->B || !pl->n || !pl->K) return fail(SC_ERR_HIP, "device allocation failed");
  int e = launch_pl_factor(ctx->c, pl->A, pl->B, pl->T, pl->K, ctx->c.stream);
  if (!e) e = add_var("gVxy", N);
  if (!e) e = add_var("gVxz", N);
  SC_HIP(hipStreamSynchronize(ctx->c.stream));
  return e;
}

int QdDriver::d_dt_vars(double t) {
  VecSet& P = vars["psi"];
  VecSet& S = vars["slip"];
  if (!pl) return d_dt(t, P.var, S.var, P.dvar, S.dvar);
  VecSet& X = vars["gVxy"];
  VecSet& Z = vars["gVxz"];
  return d_dt(t, P.var, S.var, P.dvar, S.dvar, X.var, Z.var, X.dvar, Z.dvar);
}

int QdDriver::read_flags(int* status, int* miss) {
  hipStream_t s = ctx->c.stream;
  int e = flush_finish();
  if (!e) e = materialize_u();
  if (e) return e;
  SC_HIP(hipMemcpyAsync(host_scratch, status_d, kFlagInts * sizeof(int), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  *status = fault_status((const int*)host_scratch);
  *miss = ((const int*)host_scratch)[1];
  return SC_OK;
}

// StrikeSlip_LinearElastic_qd::d_dt (strikeSlip_linearElastic_qd.cpp:756-798);
// with the power law, StrikeSlip_PowerLaw_qd::d_dt (strikeSlip_powerLaw_qd.cpp
// :633-686) and solveMomentumBalance (:787-823)
sc_fault_qd_args QdDriver::fault_args(const double* psi, double* dpsi) const {
  sc_fault_qd_args fa{};
  fa.N = Nz;
  fa.v0 = desc.v0;
  fa.vL = desc.vL;
  fa.f0 = desc.f0;
  fa.rootTol = desc.rootTol;
  fa.maxNumIts = desc.maxNumIts;
  fa.stateLaw = desc.stateLaw;
  fa.jac_mode = desc.jac_mode;
  fa.eta_d = eta_d;
  fa.sN_d = sN_d;
  fa.a_d = a_d;
  fa.b_d = b_d;
  fa.Dc_d = Dc_d;
  fa.locked_d = locked_d;
  fa.prestress_d = prestress_d;
  fa.psi_d = psi;
  fa.tauQS_d = tauQS_d;
  fa.V_d = V_d;
  fa.dpsi_d = dpsi;
  fa.tau_d = tau_d;
  fa.strength_d = strength_d;
  fa.status_d = status_d;
  return fa;
}

QdStageIn QdDriver::stage_input(const VarsLinComb* L, int st, StepGate g, double tfrac, PendingFinish f) {
  VecSet& P = vars["psi"];
  VecSet& Q = vars["slip"];
  int psiIdx = 0, slipIdx = 0, i = 0;
  for (auto& kv : vars) {
    if (kv.first == "psi") psiIdx = i;
    if (kv.first == "slip") slipIdx = i;
    ++i;
  }
  QdStageIn in;
  in.tol2 = desc.solve.rtol * desc.solve.rtol;
  const double* psi = L ? P.k[st] : P.var;
  double* dpsi = L ? P.f[st] : P.dvar;
  in.fuse = QdFuse{fault_args(psi, dpsi), L ? Q.f[st] : Q.dvar};
  if (L) in.L = *L;
  in.psiIdx = psiIdx;
  in.slipIdx = slipIdx;
  in.slipVar = Q.var;
  in.fts = desc.faultTypeScale;
  in.vL = desc.vL;
  in.remote = desc.remoteLoadingR;
  in.rfix = desc.remoteLoadingR ? shift_d : bcR_d;
  in.bcL = bcL_d;
  in.bcR = bcR_d;
  in.gate = g;
  in.tfrac = tfrac;
  in.fin = f;
  return in;
}

int QdDriver::guess_residual(double t, double* out4) {
  if (pl) return fail(SC_ERR_ARG, "the stage kernel is the linear-elastic mediator's");
  Ctx& c = ctx->c;
  if (!stage_fits(c)) return fail(SC_ERR_ARG, "the stage kernel needs a solution basis (integrate with reuse > 0) and Nz <= 2048");
  hipStream_t s = c.stream;
  int e = flush_finish();
  if (!e) e = stage_bufs_ensure(stageBufs, Nz, band_rows());
  if (!e) e = launch_stage_plan(c, stageBufs, desc.remoteLoadingR ? shift_d : bcR_d, s);
  if (e) return e;
  QdStageIn in = stage_input(nullptr, 0, StepGate{}, 0.0, PendingFinish{});
  in.t0 = t;
  if ((e = launch_qd_stage(c, stageBufs, 2, in, s))) return e;
  uLazy = true;  // u is the launch's guess at the current state
  // the residual blocks' partials, summed as the test of the next launch sums them
  const PendingFinish pf = stage_finish(stageBufs, 2, nullptr);
  std::vector<double> gp(2 * (size_t)pf.nb);
  SC_HIP(hipMemcpyAsync(gp.data(), pf.partials, gp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  // explicit: the guess, the rhs of the stage's boundary data, A u, r = b - Au
  if ((e = materialize_u())) return e;
  double* Au = nullptr;
  double* parts = nullptr;
  constexpr int nb = 512;
  if (hipMalloc(&Au, (size_t)N * sizeof(double)) != hipSuccess || hipMalloc(&parts, 2 * nb * sizeof(double)) != hipSuccess) {
    if (Au) (void)hipFree(Au);
    return fail(SC_ERR_HIP, "device allocation failed");
  }
  e = launch_set_rhs(c, rhs_d, bcL_d, bcR_d, bcT_d, bcB_d, s);
  if (!e) e = launch_apply64(c, u_d, Au, s);
  if (!e) e = launch_resid_norms(rhs_d, Au, (size_t)N, parts, nb, s);
  std::vector<double> h(2 * nb);
  if (!e && hipMemcpyAsync(h.data(), parts, h.size() * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
    e = fail(SC_ERR_HIP, "copy of the residual partials failed");
  if (!e && hipStreamSynchronize(s) != hipSuccess) e = fail(SC_ERR_HIP, "stream synchronisation failed");
  (void)hipFree(Au);
  (void)hipFree(parts);
  if (e) return e;
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < nb; ++i) {
    rr += h[2 * i];
    bb += h[2 * i + 1];
  }
  double g0 = 0.0, g1 = 0.0;
  for (int i = 0; i < pf.nb; ++i) {
    g0 += gp[2 * i];
    g1 += gp[2 * i + 1];
  }
  out4[0] = g0;
  out4[1] = g1;
  out4[2] = rr;
  out4[3] = bb;
  return SC_OK;
}

int QdDriver::bench_stage(double t, int reps, double* ms, double* words) {
  if (pl) return fail(SC_ERR_ARG, "the stage kernel is the linear-elastic mediator's");
  if (reps < 1) return fail(SC_ERR_ARG, "reps must be >= 1");
  Ctx& c = ctx->c;
  if (!stage_fits(c)) return fail(SC_ERR_ARG, "the stage kernel needs a solution basis (integrate with reuse > 0) and Nz <= 2048");
  hipStream_t s = c.stream;
  int e = flush_finish();
  if (!e) e = stage_bufs_ensure(stageBufs, Nz, band_rows());
  if (!e) e = launch_stage_plan(c, stageBufs, desc.remoteLoadingR ? shift_d : bcR_d, s);
  if (e) return e;
  QdStageIn in = stage_input(nullptr, 0, StepGate{}, 0.0, PendingFinish{});
  in.t0 = t;
  if ((e = launch_qd_stage(c, stageBufs, 2, in, s))) return e;  // warm-up
  hipEvent_t e0, e1;
  SC_HIP(hipEventCreate(&e0));
  SC_HIP(hipEventCreate(&e1));
  SC_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps && !e; ++r) e = launch_qd_stage(c, stageBufs, 2, in, s);
  SC_HIP(hipEventRecord(e1, s));
  SC_HIP(hipEventSynchronize(e1));
  float msTot = 0.f;
  SC_HIP(hipEventElapsedTime(&msTot, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e) return e;
  uLazy = true;  // u is the last launch's guess at the current state
  *ms = (double)msTot / reps;
  const double M = c.projCount;
  // per fault node the plan and tractions (2 M), slip and right datum, nine
  // friction and state inputs and eight outputs (19); per band node the M
  // operator products and the two unit-datum rhs values; the off-band Gram
  // matrix (2 M^2)
  *words = (2.0 * M + 19.0) * Nz + (M + 2.0) * band_rows() * Nz + 2.0 * M * M;
  return SC_OK;
}

int QdDriver::materialize_u() {
  if (!uLazy) return SC_OK;
  uLazy = false;
  return proj_materialize(ctx->c, u_d);
}

int QdDriver::flush_finish() {
  if (!finPending) return SC_OK;
  finPending = false;
  return launch_pending_finish(ctx->c.stream, fin);
}

int QdDriver::d_dt(double t, const double* psi, const double* slip, double* dpsi, double* dslip,
                   const double* gxy, const double* gxz, double* dgxy, double* dgxz,
                   const VarsLinComb* pre, int preSlip, StepGate gate, double tfrac) {
  hipStream_t s = ctx->c.stream;
  const double t0 = now();
  const bool count = gate.ctl == nullptr;  // the device loop counts what ran
  if (count) ++counters.d_dt_calls;
  int e;
  int mnb = 0;
  if (!pl) {
    // the rhs is zero off the boundary frame (rhs_d starts zeroed and only
    // this kernel writes it); with a solution basis the kernel also forms
    // the projections that start the solve
    // (a pending residual test and the stage combination go into the same
    // launch when there is a plan, before it otherwise)
    ProjFramePlan plan;
    e = proj_frame_plan(ctx->c, desc.solve, &plan);
    if (!e)
      e = launch_qd_rhs_frame(ctx->c, rhs_d, slip, shift_d, t, desc.vL, desc.faultTypeScale,
                              desc.remoteLoadingR, bcL_d, bcR_d, bcT_d, bcB_d, s, &plan, &mnb, pre, preSlip,
                              finPending ? &fin : nullptr, gate, tfrac);
    finPending = false;
  } else {
    e = flush_finish();
    if (!e && pre) e = launch_vars_lincomb(s, *pre);
    // the boundary data, setRhs and the viscous source in one launch
    const PlRhsArgs rb{slip, shift_d, t, desc.vL, desc.faultTypeScale, desc.remoteLoadingR, bcR_d,
                       bcT_d, bcB_d, bcL_d};
    if (!e) e = launch_pl_source(ctx->c, gxy, gxz, rhs_d, rb, s);
  }
  if (e) return e;
  sc_fault_qd_args fa = fault_args(psi, dpsi);
  // a checked solve may start CG from u: form the lazy one first
  if (uLazy && !defer && (e = materialize_u())) return e;
  const double ts = now();
  sc_solve_stats st;
  bool deferred = false, fused = false, lazy = false;
  // a deferred guess of the linear-elastic step also computes the fault rates
  const QdFuse fuse{fa, dslip};
  e = solve(ctx->c, rhs_d, u_d, desc.solve, st, defer ? miss_d : nullptr, &deferred, !pl, mnb,
            pl ? nullptr : &fuse, &fused, &fin, gate, pl ? nullptr : &lazy);
  finPending = fin.sc != nullptr;
  uLazy = lazy;  // a solve that wrote u ends the laziness of an earlier one
  if (count) {
    counters.deferred_checks += deferred ? 1 : 0;
    if (deferred) pendingDeferred++;
    else if (!e) hits = (st.iters == 0 && desc.solve.reuse > 0) ? hits + 1 : 0;
    counters.solve_seconds += now() - ts;
    counters.solves += 1;
    counters.cg_iters += st.iters;
    counters.outer_sweeps += st.outer;
  }
  if (e) return e;
  if (pl) {
    // the fault's rates fused in at the fault nodes, and the deferred guess
    // check as one more block
    e = launch_pl_strain(ctx->c, u_d, gxy, gxz, pl->K, pl->n, bcL_d, bcR_d, pl->cap,
                         dgxy, dgxz, tauQS_d, pl->effVisc, pl->sxy, &fa, dslip, finPending ? &fin : nullptr, s);
    if (e) return e;
    finPending = false;
    fused = true;
  }
  // sigma_xy on the fault from u (the power law's strain kernel wrote it already)
  if (!fused) {
    e = launch_qd_fault(ctx->c, pl ? nullptr : u_d, fa, dslip, s);
    if (e) return e;
  }
  if (count) counters.total_seconds += now() - t0;
  return SC_OK;
}

int QdDriver::device_steps(RkState& r, double finalT, int maxSteps, double& newDt, double* dts, int dtsCap,
                           int& nOut, bool* finished, bool* halted) {
  *finished = *halted = false;
  hipStream_t s = ctx->c.stream;
  if (!ctl_d) {
    SC_HIP(hipMalloc(&ctl_d, 2 * sizeof(StepCtl)));  // two for the stage loop's ping-pong
    SC_HIP(hipHostMalloc(&ctl_h, sizeof(StepCtl), hipHostMallocDefault));
  }
  if (dts && dtsCap > ctlDtsCap) {
    if (ctlDts_d) SC_HIP(hipFree(ctlDts_d));
    SC_HIP(hipMalloc(&ctlDts_d, (size_t)dtsCap * sizeof(double)));
    ctlDtsCap = dtsCap;
  }
  // the lo/hi pair and the error fields (all fields are short here)
  const char* names[4] = {"psi", "slip", "gVxy", "gVxz"};
  StepCtl h{};
  h.t = r.currT;
  h.dt = r.deltaT;
  if (h.t + h.dt > finalT) h.dt = finalT - h.t;
  h.newDt = newDt;
  h.tol = r.tol;
  h.minDt = r.minDt;
  h.maxDt = r.maxDt;
  h.finalT = finalT;
  h.kappa = r.kappa;
  h.errA0 = r.errA[0];
  h.errA1 = r.errA[1];
  h.stepCount = r.stepCount;
  h.maxSteps = maxSteps;
  h.pidBase = r.pidBase;
  h.control = r.control;
  h.normType = r.normType;
  h.ord = r.ord;
  h.nOut = nOut;
  h.dtsCap = dts ? dtsCap : 0;
  h.dts = ctlDts_d;
  VarsLoHi S;
  S.ntlo = 2;
  S.nthi = 3;
  S.clo[0] = 0.5; S.clo[1] = 0.5;                        // 0.5 h, 0.5 h
  S.chi[0] = 1; S.chi[1] = 2; S.chi[2] = 1;              // h / 6.0, 2 h / 3.0, h / 6.0
  S.cdhi[0] = 6; S.cdhi[1] = 3; S.cdhi[2] = 6;
  VarsCopy cp;
  for (int fi = 0; fi < 4; ++fi) {
    if (!vars.count(names[fi])) continue;
    VecSet& v = vars[names[fi]];
    const int q = S.nv++;
    S.n[q] = v.n;
    S.base[q] = v.var;
    S.lo[q] = v.lo;
    S.hi[q] = v.hi;
    S.flo[q][0] = v.dvar; S.flo[q][1] = v.f[2];
    S.fhi[q][0] = v.dvar; S.fhi[q][1] = v.f[1]; S.fhi[q][2] = v.f[2];
    const bool sel = (errMask & (1 << fi)) != 0;
    S.out[q] = sel ? err_d + 3 * h.nsel : err_d + 12;
    if (sel) h.selN[h.nsel++] = v.n;
    cp.n[cp.nv] = v.n;
    cp.dst[cp.nv] = v.var;
    cp.src[cp.nv] = v.hi;
    cp.nv++;
  }
  *ctl_h = h;
  SC_HIP(hipMemcpyAsync(ctl_d, ctl_h, sizeof(StepCtl), hipMemcpyHostToDevice, s));
  // the stage combinations as templates of h: k1 = var + (0.5 h) f0,
  // k2 = var + (-h) f0 + (2 h) f1
  auto comb = [&](int st, int nt, const int* fidx, const double* c) {
    VarsLinComb L;
    L.nt = nt;
    for (int i = 0; i < nt; ++i) L.c[i] = c[i];
    for (auto& kv : vars) {
      VecSet& v = kv.second;
      L.n[L.nv] = v.n;
      L.out[L.nv] = v.k[st];
      L.base[L.nv] = v.var;
      for (int i = 0; i < nt; ++i) L.f[L.nv][i] = fidx[i] == 0 ? v.dvar : v.f[fidx[i]];
      ++L.nv;
    }
    return L;
  };
  int slipIdx = 0;
  for (auto& kv : vars) {
    if (kv.first == "slip") break;
    ++slipIdx;
  }
  const int f1[1] = {0}, f2[2] = {0, 1};
  const double c1[1] = {0.5}, c2[2] = {-1.0, 2.0};
  const VarsLinComb L1 = comb(1, 1, f1, c1), L2 = comb(2, 2, f2, c2);
  VecSet& P = vars["psi"];
  VecSet& Q = vars["slip"];
  const StepGate g1{ctl_d, 1}, g2{ctl_d, 2};
  const int nOut0 = nOut;
  int prevAcc = 0, prevRej = 0;
  defer = true;
  // One attempt: both stages, the attempt's end (the last stage's residual
  // test, the lo/hi pair, the step control), the rates after an accepted
  // step.  With a basis and Nz within its tables each d_dt is one stage
  // kernel (no rhs launch) and the attempt's end runs in the last one: the
  // projection plan of the current basis first (the basis does not change
  // inside the loop; a missed guess returns to the host).  The stage kernels
  // of one attempt take positions 0 (stage 2), 1 (stage 3) and 2 (the end
  // and the rates at the accepted state); positions 0 and 1 test the
  // previous position's residual, position 2 tests position 1's.  Otherwise
  // each d_dt is the rhs and guess kernels and k_step_end ends the attempt.
  const bool oneKernel = stage_fits(ctx->c);
  if (oneKernel) {
    int e0 = stage_bufs_ensure(stageBufs, Nz, band_rows());
    if (!e0) e0 = launch_stage_plan(ctx->c, stageBufs, desc.remoteLoadingR ? shift_d : bcR_d, s);
    if (e0) return e0;
  }
  auto stage_in = [&](const VarsLinComb* L, int st, StepGate g, double tfrac, PendingFinish f) {
    return stage_input(L, st, g, tfrac, f);
  };
  // The stage loop's attempts alternate between two states and two control
  // blocks (parity p): an attempt's kernels read state p and control block
  // p, and its last kernel (the attempt's end and the rates at the accepted
  // state, every block alike) writes state and control block 1 - p, so that
  // no launch writes what its other blocks read.  The direct attempt has
  // p = 0 and a graph holds an even number of attempts starting at p = 1:
  // after every batch the latest state and control block are those of p = 1.
  static_assert(kAttemptsPerGraph % 2 == 0 && kAttemptsPerTail % 2 == 0,
                "the stage loop's ping-pong needs an even number of attempts per graph");
  static_assert(kAttemptsPerGraph % kAttemptsPerTail == 0, "the full graph repeats the short one's launches");
  double* varB[2][2] = {{P.var, Q.var}, {nullptr, nullptr}};
  double* dvarB[2][2] = {{P.dvar, Q.dvar}, {nullptr, nullptr}};
  if (oneKernel) {
    double* alt = stageBufs.alt;
    varB[1][0] = alt;
    varB[1][1] = alt + Nz;
    dvarB[1][0] = alt + 2 * (size_t)Nz;
    dvarB[1][1] = alt + 3 * (size_t)Nz;
  }
  int psiIdx = 0, slipIdxL = 0, vi = 0;
  for (auto& kv : vars) {
    if (kv.first == "psi") psiIdx = vi;
    if (kv.first == "slip") slipIdxL = vi;
    ++vi;
  }
  auto at_parity = [&](VarsLinComb L, int p) {  // base and f0 of state p
    L.base[psiIdx] = varB[p][0];
    L.base[slipIdxL] = varB[p][1];
    L.f[psiIdx][0] = dvarB[p][0];
    L.f[slipIdxL][0] = dvarB[p][1];
    return L;
  };
  int selSlot[2], nsSel = 0;
  selSlot[0] = (errMask & 1) ? nsSel++ : -1;
  selSlot[1] = (errMask & 2) ? nsSel++ : -1;
  int parity = 0;
  std::vector<unsigned char>* rec = nullptr;  // set: slot() records its launches' arguments
  // per-launch timing (sc_qd_integrate_timed): events around the three stage
  // launches of each attempt, durations per attempt in launch order
  const bool timed = timing != nullptr && oneKernel;
  hipEvent_t ev[6] = {};
  std::vector<float> attemptMs;
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard() {
      for (int i = 0; i < 6; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } evGuard{ev};
  if (timed)
    for (int i = 0; i < 6; ++i) SC_HIP(hipEventCreate(&ev[i]));
  // the events take the kernels' own start and end (no marker packets between launches)
  auto evt = [&](int i) -> hipEvent_t { return timed && !rec ? ev[i] : nullptr; };
  auto collect = [&]() -> int {
    SC_HIP(hipEventSynchronize(ev[5]));
    for (int k = 0; k < 3; ++k) {
      float ms = 0.f;
      SC_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
      attemptMs.push_back(ms);
    }
    return SC_OK;
  };
  auto slot = [&]() -> int {
    if (oneKernel) {
      const int p = parity;
      parity = 1 - p;
      const StepGate gp{ctl_d + p, 1};
      const PendingFinish prev = finPending ? fin : PendingFinish{};
      const VarsLinComb A1 = at_parity(L1, p), A2 = at_parity(L2, p);
      int e = launch_qd_stage(ctx->c, stageBufs, 0, stage_in(&A1, 1, gp, 0.5, prev), s, rec, evt(0), evt(1));
      if (!e)
        e = launch_qd_stage(ctx->c, stageBufs, 1, stage_in(&A2, 2, gp, 1.0, stage_finish(stageBufs, 0, miss_d)), s,
                            rec, evt(2), evt(3));
      if (!e) {
        QdStageIn in = stage_in(nullptr, 0, gp, 0.0, PendingFinish{});
        in.fuse = QdFuse{fault_args(varB[1 - p][0], dvarB[1 - p][0]), dvarB[1 - p][1]};
        // its root solves report where the next attempt's end reads (blocks
        // of one launch do not see one another's writes in order)
        in.fuse.a.status_d = endStatus_d + (1 - p);
        in.end = true;
        in.ctlOut = ctl_d + (1 - p);
        in.flags = status_d;
        in.flagEnd = endStatus_d + p;
        in.fin3 = stage_finish(stageBufs, 1, miss_d);
        for (int f = 0; f < 2; ++f) {
          in.pv[f] = varB[p][f];
          in.pvOut[f] = varB[1 - p][f];
          in.pf[f][0] = dvarB[p][f];
          in.pf[f][1] = f == 0 ? P.f[1] : Q.f[1];
          in.pf[f][2] = f == 0 ? P.f[2] : Q.f[2];
          in.selSlot[f] = selSlot[f];
        }
        // the RK32 pair: lo = y + h/2 (f0 + f2), hi = y + h/6 f0 + 2h/3 f1 + h/6 f2
        in.clo[0] = 0.5;
        in.clo[1] = 0.5;
        in.chi[0] = 1;
        in.chi[1] = 2;
        in.chi[2] = 1;
        in.cdhi[0] = 6;
        in.cdhi[1] = 3;
        in.cdhi[2] = 6;
        e = launch_qd_stage(ctx->c, stageBufs, 2, in, s, rec, evt(4), evt(5));
      }
      // the rates' residual test is left for the next launch; u for proj_materialize
      fin = stage_finish(stageBufs, 2, miss_d);
      finPending = true;
      uLazy = true;
      return e;
    }
    int e = d_dt(0.0, P.k[1], Q.k[1], P.f[1], Q.f[1], nullptr, nullptr, nullptr, nullptr, &L1, slipIdx, g1, 0.5);
    if (!e) e = d_dt(0.0, P.k[2], Q.k[2], P.f[2], Q.f[2], nullptr, nullptr, nullptr, nullptr, &L2, slipIdx, g1, 1.0);
    if (!e) e = launch_step_end(s, StepEnd{finPending ? fin : PendingFinish{}, S, ctl_d, err_d, status_d, cp}, g1);
    finPending = false;
    if (!e) e = d_dt(0.0, P.var, Q.var, P.dvar, Q.dvar, nullptr, nullptr, nullptr, nullptr, nullptr, -1, g2, 0.0);
    return e;
  };
  // The first attempt goes in directly (it also takes any residual test left
  // from the host's last solve); every later one repeats the same launches
  // with the same arguments (all that changes lives in the control block),
  // so a captured run of attempts is replayed as a graph (attempts past the
  // run's end find the gate closed).  The capture holds
  // no memset nodes.  The one-kernel loop's graph is kept across calls and
  // captured again only when the arguments of its launches change.
  int e = slot();
  if (!e && timed) e = collect();
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto capture = [&](hipGraph_t* gr, hipGraphExec_t* ex, int nAtt) -> int {
    SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    int ec = 0;
    for (int a = 0; a < nAtt && !ec; ++a) ec = slot();
    hipError_t ce = hipStreamEndCapture(s, gr);
    if (!ec && ce != hipSuccess) ec = fail(SC_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(ce));
    if (!ec && hipGraphInstantiate(ex, *gr, nullptr, nullptr, 0) != hipSuccess)
      ec = fail(SC_ERR_HIP, "graph instantiation failed");
    return ec;
  };
  hipGraphExec_t run = nullptr, tail = nullptr;
  if (timed) {
    // no graph: the attempts go in one by one
  } else if (!e && oneKernel) {
    const int parity0 = parity;
    const PendingFinish fin0 = fin;
    const bool finPending0 = finPending, uLazy0 = uLazy;
    std::vector<unsigned char> key;
    rec = &key;
    for (int a = 0; a < kAttemptsPerTail && !e; ++a) e = slot();
    rec = nullptr;
    auto restore = [&]() {
      parity = parity0;
      fin = fin0;
      finPending = finPending0;
      uLazy = uLazy0;
    };
    restore();
    if (!e && (!stageExec || !tailExec || key != stageKey)) {
      drop_stage_graph();
      e = capture(&stageGraph, &stageExec, kAttemptsPerGraph);
      restore();
      if (!e) e = capture(&tailGraph, &tailExec, kAttemptsPerTail);
      if (e) drop_stage_graph();
      else {
        stageKey.swap(key);
        ++graphBuilds;
      }
    }
    run = stageExec;
    tail = tailExec;
  } else if (!e) {
    e = capture(&graph, &exec, kAttemptsPerTail);
    run = tail = exec;
  }
  struct GraphGuard {
    hipGraph_t& g;
    hipGraphExec_t& x;
    ~GraphGuard() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } gguard{graph, exec};
  bool first = true;
  for (;;) {
    if (e) break;
    const int left = maxSteps - ctl_h->stepCount + 1;
    const int slots = std::max(1, std::min(kSlotsPerBatch, left + 1)) - (first ? 1 : 0);
    if (timed) {
      // two attempts between read-backs keep the parity of the graph batches
      for (int k = 0; k < (first ? 0 : 2) && !e; ++k) {
        e = slot();
        if (!e) e = collect();
      }
    } else {
      // whole batches as the full graph, the rest as the short one
      const bool full = run != tail;
      const int nFull = full ? slots / kAttemptsPerGraph : 0;
      const int nTail = (slots - nFull * kAttemptsPerGraph + kAttemptsPerTail - 1) / kAttemptsPerTail;
      for (int k = 0; k < nFull + nTail; ++k) {
        if (hipGraphLaunch(k < nFull ? run : tail, s) != hipSuccess) {
          e = fail(SC_ERR_HIP, "graph launch failed");
          break;
        }
      }
    }
    first = false;
    finPending = true;  // the last replayed attempt's rates
    if (e) break;
    SC_HIP(hipMemcpyAsync(ctl_h, ctl_d + (oneKernel ? 1 : 0), sizeof(StepCtl), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    // what ran since the last read-back: two stage solves per attempt and
    // one more per accepted step
    const int acc = ctl_h->nOut - nOut0, rej = ctl_h->rejected;
    const int dAcc = acc - prevAcc, dRej = rej - prevRej;
    const long long ran = 2LL * (dAcc + dRej + (ctl_h->halt ? 1 : 0)) + dAcc;
    prevAcc = acc;
    prevRej = rej;
    counters.d_dt_calls += ran;
    counters.solves += ran;
    counters.deferred_checks += ran;
    counters.rejected_steps += dRej;
    counters.device_steps += dAcc;
    if (ctl_h->halt) break;
    hits += (int)ran;
    if (ctl_h->done) break;
  }
  defer = false;
  if (e) return e;
  if (timed) {  // the attempts that ran come first; later ones found the gate closed
    const long long ran = (long long)(ctl_h->nOut - nOut0) + ctl_h->rejected + (ctl_h->halt ? 1 : 0);
    const long long nAtt = std::min<long long>(ran, (long long)attemptMs.size() / 3);
    for (long long a = 0; a < nAtt; ++a)
      for (int k = 0; k < 3; ++k) timing->ms[k] += attemptMs[(size_t)(3 * a + k)];
    timing->attempts += nAtt;
  }
  if (oneKernel) {  // the latest state is state 1: back into var and dvar
    for (int f = 0; f < 2; ++f) {
      SC_HIP(hipMemcpyAsync(varB[0][f], varB[1][f], (size_t)Nz * sizeof(double), hipMemcpyDeviceToDevice, s));
      SC_HIP(hipMemcpyAsync(dvarB[0][f], dvarB[1][f], (size_t)Nz * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
  }
  const StepCtl& c = *ctl_h;
  r.currT = c.t;
  r.deltaT = c.dt;
  r.errA[0] = c.errA0;
  r.errA[1] = c.errA1;
  r.stepCount = c.stepCount;
  newDt = c.newDt;
  if (dts && c.nOut > nOut0) {
    const int hi = std::min(c.nOut, dtsCap);
    if (hi > nOut0)
      SC_HIP(hipMemcpy(dts + nOut0, ctlDts_d + nOut0, (size_t)(hi - nOut0) * sizeof(double), hipMemcpyDeviceToHost));
  }
  nOut = c.nOut;
  if (c.halt) {
    // the attempt in progress read a set flag: a root-solve failure ends the
    // run, a missed guess sends the step back to the host
    finPending = false;  // the flags are cleared below; a repeated test would set them again
    int st = 0, miss = 0;
    if ((e = read_flags(&st, &miss))) return e;
    note_checked(true);
    if (st) {
      (void)hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s);
      return fail(st & SC_ERR_NONFINITE ? SC_ERR_NONFINITE : SC_ERR_NOCONV,
                  "fault root solve failed (status " + std::to_string(st) + ")");
    }
    counters.missed_guesses++;
    SC_HIP(hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s));
    *halted = true;
    return SC_OK;
  }
  *finished = true;
  return SC_OK;
}

int QdDriver::check_status() {
  hipStream_t s = ctx->c.stream;
  SC_HIP(hipMemcpyAsync(host_scratch, status_d, kFlagInts * sizeof(int), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  const int st = fault_status((const int*)host_scratch);
  if (st) {
    (void)hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s);
    return fail(st & SC_ERR_NONFINITE ? SC_ERR_NONFINITE : SC_ERR_NOCONV,
                "fault root solve failed (status " + std::to_string(st) + ")");
  }
  return SC_OK;
}

// computeStepSize (odeSolver.cpp:395-445, 842-898)
double QdDriver::step_size(const RkState& r, double totErr) const {
  double ratio;
  if (r.control == 0 || r.stepCount - r.pidBase < 4) {
    ratio = r.kappa * std::pow(r.tol / totErr, 1. / (1. + r.ord));
  } else {
    const double al = 0.49 / r.ord, be = 0.34 / r.ord, ga = 0.1 / r.ord;
    ratio = r.kappa * std::pow(r.tol / totErr, al) * std::pow(r.errA[0] / r.tol, be) *
            std::pow(r.tol / r.errA[1], ga);
  }
  double dt = ratio * r.deltaT;
  dt = std::min(r.deltaT * 5.0, dt);
  dt = std::min(r.maxDt, dt);
  dt = std::max(r.minDt, dt);
  return dt;
}

// y_lo, y_hi of the embedded pair and computeError (odeSolver.cpp:449-501,
// 902-969) over errInds (timeIntInds): psi and slip by default, the viscous
// strains when selected.  Short fields take one fused kernel, long ones
// (the viscous strains) a combination and a block-partial error pass.  The
// same read-back brings the fault status and the missed-guess flag: *missed
// says that a deferred solve since the last check missed its tolerance (both
// flags are then cleared and the error is meaningless).
int QdDriver::lohi_error(const RkState& r, int ntlo, const int* filo, const double* clo, int nthi,
                         const int* fihi, const double* chi, double* totErr, bool* missed) {
  hipStream_t s = ctx->c.stream;
  const char* names[4] = {"psi", "slip", "gVxy", "gVxz"};
  int slot[4], ns = 0;  // error slot of each field, -1 when not selected
  for (int i = 0; i < 4; ++i) slot[i] = ((errMask & (1 << i)) && vars.count(names[i])) ? ns++ : -1;
  VarsLoHi S;
  S.ntlo = ntlo;
  S.nthi = nthi;
  for (int i = 0; i < ntlo; ++i) S.clo[i] = clo[i];
  for (int i = 0; i < nthi; ++i) S.chi[i] = chi[i];
  VarsLinComb Llo, Lhi;
  Llo.nt = ntlo;
  Lhi.nt = nthi;
  for (int i = 0; i < ntlo; ++i) Llo.c[i] = clo[i];
  for (int i = 0; i < nthi; ++i) Lhi.c[i] = chi[i];
  std::vector<int> longSel;
  for (int fi = 0; fi < 4; ++fi) {
    if (!vars.count(names[fi])) continue;
    VecSet& v = vars[names[fi]];
    auto fsel = [&](int id) { return id == 0 ? (const double*)v.dvar : (const double*)v.f[id]; };
    if (v.n <= kShortField) {
      const int q = S.nv++;
      S.n[q] = v.n;
      S.base[q] = v.var;
      S.lo[q] = v.lo;
      S.hi[q] = v.hi;
      for (int i = 0; i < ntlo; ++i) S.flo[q][i] = fsel(filo[i]);
      for (int i = 0; i < nthi; ++i) S.fhi[q][i] = fsel(fihi[i]);
      S.out[q] = slot[fi] >= 0 ? err_d + 3 * slot[fi] : err_d + 12;
    } else {
      // the low-order solution of a long field enters only its error sums:
      // formed only when the norm reads the field
      for (VarsLinComb* L : {&Llo, &Lhi}) {
        if (L == &Llo && slot[fi] < 0) continue;
        const int q = L->nv++;
        L->n[q] = v.n;
        L->out[q] = L == &Llo ? v.lo : v.hi;
        L->base[q] = v.var;
        const int nt = L == &Llo ? ntlo : nthi;
        const int* fidx = L == &Llo ? filo : fihi;
        for (int i = 0; i < nt; ++i) L->f[q][i] = fsel(fidx[i]);
      }
      if (slot[fi] >= 0) longSel.push_back(fi);
    }
  }
  int e = launch_vars_lohi_err(s, S, finPending ? &fin : nullptr);
  finPending = false;
  if (!e && Llo.nv) e = launch_vars_lincomb(s, Llo);
  if (!e && Lhi.nv) e = launch_vars_lincomb(s, Lhi);
  for (int fi : longSel) {
    if (e) break;
    VecSet& v = vars[names[fi]];
    e = launch_err(s, v.n, v.lo, v.hi, err_d + 3 * slot[fi], err_part_d);
  }
  if (e) return e;
  SC_HIP(hipMemcpyAsync(host_scratch, err_d, kChkDoubles * sizeof(double), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  const double* h = (const double*)host_scratch;
  const int status = fault_status((const int*)(h + kErrSlots));
  const int miss = ((const int*)(h + kErrSlots))[1];
  *missed = miss != 0;
  note_checked(*missed);
  if (miss) {
    counters.missed_guesses++;
    SC_HIP(hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s));
    *totErr = 0;
    return SC_OK;
  }
  if (status) {
    (void)hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s);
    return fail(status & SC_ERR_NONFINITE ? SC_ERR_NONFINITE : SC_ERR_NOCONV,
                "fault root solve failed (status " + std::to_string(status) + ")");
  }
  double tot = 0;
  for (int fi = 0; fi < 4; ++fi) {
    if (slot[fi] < 0) continue;
    const double* v = h + 3 * slot[fi];
    const double n = (double)vars[names[fi]].n;
    if (r.normType == 0) tot += std::sqrt(v[0]) / (std::sqrt(n) * 1.0);
    else if (r.ord == 3) tot += std::sqrt(v[0]) / (std::sqrt(v[1]) * 1.0);
    else tot += v[2] / 1.0;
  }
  *totErr = tot;
  return SC_OK;
}

// out = var + sum_i c[i] f_i over every integrated field, one launch
int QdDriver::lincomb_all(const char* out_sel, int stage, int nt, const int* fidx, const double* c) {
  VarsLinComb L;
  L.nt = nt;
  for (int i = 0; i < nt; ++i) L.c[i] = c[i];
  for (auto& kv : vars) {
    VecSet& v = kv.second;
    double* out = nullptr;
    if (!std::strcmp(out_sel, "k")) out = v.k[stage];
    else if (!std::strcmp(out_sel, "lo")) out = v.lo;
    else out = v.hi;
    L.n[L.nv] = v.n;
    L.out[L.nv] = out;
    L.base[L.nv] = v.var;
    for (int i = 0; i < nt; ++i) L.f[L.nv][i] = fidx[i] == 0 ? v.dvar : v.f[fidx[i]];
    ++L.nv;
  }
  return launch_vars_lincomb(ctx->c.stream, L);
}

int QdDriver::stage_comb(int st, double t, int nt, const int* fidx, const double* c) {
  if (pl) {
    const int e = lincomb_all("k", st, nt, fidx, c);
    return e ? e : stage(st, t);
  }
  VarsLinComb L;
  L.nt = nt;
  for (int i = 0; i < nt; ++i) L.c[i] = c[i];
  int slipIdx = -1;
  for (auto& kv : vars) {
    VecSet& v = kv.second;
    if (kv.first == "slip") slipIdx = L.nv;
    L.n[L.nv] = v.n;
    L.out[L.nv] = v.k[st];
    L.base[L.nv] = v.var;
    for (int i = 0; i < nt; ++i) L.f[L.nv][i] = fidx[i] == 0 ? v.dvar : v.f[fidx[i]];
    ++L.nv;
  }
  VecSet& p = vars["psi"];
  VecSet& q = vars["slip"];
  return d_dt(t, p.k[st], q.k[st], p.f[st], q.f[st], nullptr, nullptr, nullptr, nullptr, &L, slipIdx);
}

int QdDriver::stage(int st, double t) {
  VecSet& p = vars["psi"];
  VecSet& q = vars["slip"];
  if (!pl) return d_dt(t, p.k[st], q.k[st], p.f[st], q.f[st]);
  VecSet& x = vars["gVxy"];
  VecSet& z = vars["gVxz"];
  return d_dt(t, p.k[st], q.k[st], p.f[st], q.f[st], x.k[st], z.k[st], x.f[st], z.f[st]);
}

int QdDriver::integrate(const sc_qd_run& run, double* tEnd, int* steps, double* dts, int dtsCap) {
  RkState r;
  r.ord = run.integrator == 0 ? 3 : 4;
  r.control = run.control;
  r.normType = run.normType;
  r.tol = run.tol;
  r.minDt = run.minDt;
  r.maxDt = run.maxDt;
  r.deltaT = run.initDt;
  r.currT = run.initTime;
  r.errA[0] = run.currErr;
  r.errA[1] = run.prevErr;
  const double finalT = run.maxTime;
  int maxSteps = run.maxSteps;
  const bool replay = run.nReplay > 0 && run.replayDt_h;
  if (replay && maxSteps > run.nReplay) maxSteps = run.nReplay;
  r.stepCount = stepBase;
  r.pidBase = pidFresh ? stepBase : 0;
  maxSteps += stepBase;
  hipStream_t s = ctx->c.stream;
  int e;
  if (out) {  // timeMonitor(initTime, initDeltaT, 0) before the solver starts (:729-732)
    const double zero[2] = {0.0, 0.0};
    int stop = 0;  // ignored, as there
    if ((e = monitor(r.currT, r.deltaT, 0, zero, r.deltaT, run.maxSteps, finalT, true, &stop))) return e;
  }
  if (finalT == r.currT || run.maxSteps == 0) {
    *tEnd = r.currT;
    *steps = r.stepCount;
    return SC_OK;
  }
  if (r.deltaT == 0) r.deltaT = (finalT - r.currT) / run.maxSteps;
  // a run that ended in an error may have left the device flags set (and a
  // residual test pending)
  finPending = false;
  SC_HIP(hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s));
  // solves with a projected guess are checked on the device; the flag is read
  // with the error estimate of each attempt
  const bool may_defer = desc.solve.reuse > 0 && run.deferChecks != 1;
  const bool always = run.deferChecks == 2;
  struct DeferGuard {
    bool& d;
    ~DeferGuard() { d = false; }
  } guard{defer};
  defer = false;
  e = d_dt_vars(r.currT);
  if (e) return e;
  double totErr = 0, newDt = r.deltaT;
  int nOut = 0;
  // RK32 of the linear-elastic mediator with deferred checks runs its step
  // control on the device (no per-step read-back) unless the caller asks
  // otherwise or something needs the host at every step
  const bool deviceLoop = run.stepControl == 0 && !replay && !out && !hook && !pl && r.ord == 3 && may_defer &&
                          desc.solve.reuse > 0;
  while (r.stepCount < maxSteps && r.currT < finalT) {
    r.stepCount++;
    int attempt = 0;
    bool checked = false;  // this step's solves are checked on the host
    if (deviceLoop && (always || defer_ok()) && ctx->c.projCount > 0) {
      bool finished = false, halted = false;
      if ((e = device_steps(r, finalT, maxSteps, newDt, dts, dtsCap, nOut, &finished, &halted))) return e;
      if (finished) break;
      // a missed guess halted it: this step on the host with checked solves,
      // its starting rates first
      checked = true;
      defer = false;
      if ((e = d_dt_vars(r.currT))) return e;
    }
    while (attempt < 100) {
      attempt++;
      defer = may_defer && !checked && (always || defer_ok());
      if (replay) r.deltaT = run.replayDt_h[r.stepCount - stepBase - 1];
      if (r.currT + r.deltaT > finalT) r.deltaT = finalT - r.currT;
      const double h = r.deltaT;
      bool missed = false;
      if (r.ord == 3) {
        { const int fi[1] = {0}; const double c[1] = {0.5 * h}; if ((e = stage_comb(1, r.currT + 0.5 * h, 1, fi, c))) return e; }
        { const int fi[2] = {0, 1}; const double c[2] = {-h, 2 * h}; if ((e = stage_comb(2, r.currT + h, 2, fi, c))) return e; }
        const int flo[2] = {0, 2}, fhi[3] = {0, 1, 2};
        const double clo[2] = {0.5 * h, 0.5 * h}, chi[3] = {h / 6.0, 2 * h / 3.0, h / 6.0};
        if ((e = lohi_error(r, 2, flo, clo, 3, fhi, chi, &totErr, &missed))) return e;
      } else {
        const double c2 = 1. / 2., c3 = 83. / 250., c4 = 31. / 50., c5 = 17. / 20., c6 = 1.;
        const double b1 = 82889. / 524892., b3 = 15625. / 83664., b4 = 69875. / 102672.,
                     b5 = -2260. / 8211., b6 = 1. / 4.;
        const double hb1 = 4586570599. / 29645900160., hb3 = 178811875. / 945068544.,
                     hb4 = 814220225. / 1159782912., hb5 = -3700637. / 11593932.,
                     hb6 = 61727. / 225920.;
        const double a21 = 1. / 2., a31 = 13861. / 62500., a32 = 6889. / 62500.;
        const double a41 = -116923316275. / 2393684061468., a42 = -2731218467317. / 15368042101831.,
                     a43 = 9408046702089. / 11113171139209.;
        const double a51 = -451086348788. / 2902428689909., a52 = -2682348792572. / 7519795681897.,
                     a53 = 12662868775082. / 11960479115383., a54 = 3355817975965. / 11060851509271.;
        const double a61 = 647845179188. / 3216320057751., a62 = 73281519250. / 8382639484533.,
                     a63 = 552539513391. / 3454668386233., a64 = 3354512671639. / 8306763924573.,
                     a65 = 4040. / 17871.;
        { const int fi[1] = {0}; const double c[1] = {a21 * h}; if ((e = stage_comb(2, r.currT + c2 * h, 1, fi, c))) return e; }
        { const int fi[2] = {0, 2}; const double c[2] = {a31 * h, a32 * h}; if ((e = stage_comb(3, r.currT + c3 * h, 2, fi, c))) return e; }
        { const int fi[3] = {0, 2, 3}; const double c[3] = {a41 * h, a42 * h, a43 * h}; if ((e = stage_comb(4, r.currT + c4 * h, 3, fi, c))) return e; }
        { const int fi[4] = {0, 2, 3, 4}; const double c[4] = {a51 * h, a52 * h, a53 * h, a54 * h}; if ((e = stage_comb(5, r.currT + c5 * h, 4, fi, c))) return e; }
        { const int fi[5] = {0, 2, 3, 4, 5}; const double c[5] = {a61 * h, a62 * h, a63 * h, a64 * h, a65 * h}; if ((e = stage_comb(6, r.currT + c6 * h, 5, fi, c))) return e; }
        const int fi[5] = {0, 3, 4, 5, 6};
        const double clo[5] = {hb1 * h, hb3 * h, hb4 * h, hb5 * h, hb6 * h};
        const double chi[5] = {b1 * h, b3 * h, b4 * h, b5 * h, b6 * h};
        if ((e = lohi_error(r, 5, fi, clo, 5, fi, chi, &totErr, &missed))) return e;
      }
      if (missed) {
        // a projected guess since the last check (this attempt's stages or the
        // rates at the step start) missed: redo both with checked solves
        defer = false;
        checked = true;
        if ((e = d_dt_vars(r.currT))) return e;
        attempt--;
        continue;
      }
      if (replay || (r.ord == 3 ? totErr <= r.tol : totErr < r.tol)) break;
      r.deltaT = step_size(r, totErr);
      counters.rejected_steps++;
      if (r.minDt == r.deltaT) break;
    }
    r.currT = r.currT + r.deltaT;
    // accept: var <- hi (the buffers trade places)
    for (auto& kv : vars) std::swap(kv.second.var, kv.second.hi);
    defer = may_defer && (always || defer_ok());
    if ((e = d_dt_vars(r.currT))) return e;
    if (totErr != 0.0) newDt = step_size(r, totErr);
    r.errA[1] = r.errA[0];
    r.errA[0] = totErr;
    if (dts && nOut < dtsCap) dts[nOut++] = r.deltaT;
    if (out) {
      int stop = 0;
      if ((e = monitor(r.currT, r.deltaT, r.stepCount, r.errA, newDt, run.maxSteps, finalT, false, &stop)))
        return e;
      if (stop) break;
    }
    if (hook) {
      int stop = 0;
      if ((e = hook->after_step(*this, r.currT, r.deltaT, r.stepCount, &stop))) return e;
      if (stop) break;
    }
    r.deltaT = newDt;
  }
  defer = false;
  // the rates at the final state: check their guesses too
  if ((e = settle(r.currT))) return e;
  if ((e = check_status())) return e;
  *tEnd = r.currT;
  *steps = r.stepCount;
  return SC_OK;
}

int QdDriver::max_ratio(double t, double* R) {
  int e = settle(t);
  if (e) return e;
  hipStream_t s = ctx->c.stream;
  double* slot = err_d + kErrSlots - 1;
  if ((e = launch_max_ratio(s, eta_d, V_d, tauQS_d, Nz, slot))) return e;
  SC_HIP(hipMemcpyAsync(host_scratch, slot, sizeof(double), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  *R = ((const double*)host_scratch)[0];
  return SC_OK;
}

int QdDriver::settle(double t) {
  hipStream_t s = ctx->c.stream;
  const int ef = flush_finish();
  if (ef) return ef;
  SC_HIP(hipMemcpyAsync(host_scratch, status_d, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  const bool missed = ((const int*)host_scratch)[1] != 0;
  note_checked(missed);
  if (missed) {
    counters.missed_guesses++;
    SC_HIP(hipMemsetAsync(status_d, 0, kFlagInts * sizeof(int), s));
    const bool d = defer;
    defer = false;
    const int e = d_dt_vars(t);
    defer = d;
    if (e) return e;
    SC_HIP(hipStreamSynchronize(s));
  }
  return materialize_u();
}

}  // namespace sc

// ---------------------------------------------------------------- C ABI ---
using sc::fail;

extern "C" int sc_qd_create(const sc_qd_desc* desc, const double* mu_h, const double* y_h,
                            const double* z_h, const double* a_h, const double* b_h,
                            const double* Dc_h, const double* sN_h, const double* eta_h,
                            const double* locked_h, const double* prestress_h,
                            const double* bcRShift_h, sc_qd** out) {
  if (!desc || !mu_h || !a_h || !b_h || !Dc_h || !sN_h || !eta_h || !out)
    return fail(SC_ERR_ARG, "sc_qd_create: null argument");
  if (desc->stateLaw != SC_AGING_LAW && desc->stateLaw != SC_SLIP_LAW)
    return fail(SC_ERR_ARG, "stateLaw must be agingLaw or slipLaw");
  if (!(desc->faultTypeScale > 0)) return fail(SC_ERR_ARG, "faultTypeScale must be > 0");
  auto* h = new sc_qd;
  h->q.reset(new sc::QdDriver);
  int e = h->q->init(*desc, mu_h, y_h, z_h, a_h, b_h, Dc_h, sN_h, eta_h, locked_h, prestress_h,
                     bcRShift_h);
  if (e) {
    delete h;
    return e;
  }
  *out = h;
  return SC_OK;
}

extern "C" int sc_qd_create_powerlaw(const sc_qd_desc* desc, const double* mu_h, const double* y_h,
                                     const double* z_h, const double* a_h, const double* b_h,
                                     const double* Dc_h, const double* sN_h, const double* eta_h,
                                     const double* locked_h, const double* prestress_h,
                                     const double* bcRShift_h, const double* T_h,
                                     const double* A_h, const double* B_h, const double* n_h,
                                     double effViscCap, int errMask, sc_qd** out) {
  if (!T_h || !A_h || !B_h || !n_h || !out) return fail(SC_ERR_ARG, "sc_qd_create_powerlaw: null argument");
  if (!(effViscCap > 0)) return fail(SC_ERR_ARG, "effViscCap must be > 0");
  if (errMask <= 0 || errMask > 15) return fail(SC_ERR_ARG, "errMask must select 1..4 fields");
  if (desc && desc->grid.laplace != 0) return fail(SC_ERR_ARG, "power law needs laplace type yz");
  sc_qd* h = nullptr;
  int e = sc_qd_create(desc, mu_h, y_h, z_h, a_h, b_h, Dc_h, sN_h, eta_h, locked_h, prestress_h,
                       bcRShift_h, &h);
  if (e) return e;
  e = h->q->init_powerlaw(T_h, A_h, B_h, n_h, effViscCap, errMask);
  if (e) {
    delete h;
    return e;
  }
  *out = h;
  return SC_OK;
}

extern "C" int sc_qd_set_visc(sc_qd* h, const double* gxy_h, const double* gxz_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  if (!q.pl) return fail(SC_ERR_ARG, "not a power-law driver");
  hipStream_t s = q.ctx->c.stream;
  const size_t bytes = (size_t)q.N * sizeof(double);
  if (gxy_h) SC_HIP(hipMemcpyAsync(q.vars["gVxy"].var, gxy_h, bytes, hipMemcpyHostToDevice, s));
  if (gxz_h) SC_HIP(hipMemcpyAsync(q.vars["gVxz"].var, gxz_h, bytes, hipMemcpyHostToDevice, s));
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

extern "C" int sc_qd_get_visc(sc_qd* h, double* gxy_h, double* gxz_h, double* dgxy_h,
                              double* dgxz_h, double* effVisc_h, double* sxy_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  if (!q.pl) return fail(SC_ERR_ARG, "not a power-law driver");
  hipStream_t s = q.ctx->c.stream;
  const size_t bytes = (size_t)q.N * sizeof(double);
  SC_HIP(hipStreamSynchronize(s));
  if (gxy_h) SC_HIP(hipMemcpyAsync(gxy_h, q.vars["gVxy"].var, bytes, hipMemcpyDeviceToHost, s));
  if (gxz_h) SC_HIP(hipMemcpyAsync(gxz_h, q.vars["gVxz"].var, bytes, hipMemcpyDeviceToHost, s));
  if (dgxy_h) SC_HIP(hipMemcpyAsync(dgxy_h, q.vars["gVxy"].dvar, bytes, hipMemcpyDeviceToHost, s));
  if (dgxz_h) SC_HIP(hipMemcpyAsync(dgxz_h, q.vars["gVxz"].dvar, bytes, hipMemcpyDeviceToHost, s));
  if (effVisc_h) SC_HIP(hipMemcpyAsync(effVisc_h, q.pl->effVisc, bytes, hipMemcpyDeviceToHost, s));
  if (sxy_h) SC_HIP(hipMemcpyAsync(sxy_h, q.pl->sxy, bytes, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

extern "C" int sc_qd_destroy(sc_qd* q) {
  delete q;
  return SC_OK;
}

extern "C" sc_ctx* sc_qd_ctx(sc_qd* q) { return q ? q->q->ctx : nullptr; }

extern "C" int sc_qd_integrate_timed(sc_qd* h, const sc_qd_run* run, double* tEnd_h, int* steps_h,
                                     double* dts_h, int dtsCap, double* ms_stage, double* ms_end,
                                     long long* attempts) {
  if (!h || !run || !ms_stage || !ms_end || !attempts) return fail(SC_ERR_ARG, "null argument");
  sc::QdDriver& q = *h->q;
  sc::QdDriver::StageTiming t;
  q.timing = &t;
  const int e = sc_qd_integrate(h, run, tEnd_h, steps_h, dts_h, dtsCap);
  q.timing = nullptr;
  *attempts = t.attempts;
  *ms_stage = t.attempts > 0 ? (t.ms[0] + t.ms[1]) / (2.0 * t.attempts) : 0.0;
  *ms_end = t.attempts > 0 ? t.ms[2] / t.attempts : 0.0;
  return e;
}

extern "C" int sc_qd_guess_residual(sc_qd* h, double t, double* rr_bb) {
  if (!h || !rr_bb) return fail(SC_ERR_ARG, "null argument");
  return h->q->guess_residual(t, rr_bb);
}

extern "C" int sc_qd_bench_stage(sc_qd* h, double t, int reps, double* ms_per_launch, double* words) {
  if (!h || !ms_per_launch || !words) return fail(SC_ERR_ARG, "null argument");
  return h->q->bench_stage(t, reps, ms_per_launch, words);
}

extern "C" int sc_qd_set_state(sc_qd* h, const double* psi_h, const double* slip_h,
                               const double* V_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  const size_t bytes = (size_t)q.Nz * sizeof(double);
  if (psi_h) SC_HIP(hipMemcpy(q.vars["psi"].var, psi_h, bytes, hipMemcpyHostToDevice));
  if (slip_h) SC_HIP(hipMemcpy(q.vars["slip"].var, slip_h, bytes, hipMemcpyHostToDevice));
  if (V_h) SC_HIP(hipMemcpy(q.V_d, V_h, bytes, hipMemcpyHostToDevice));
  return SC_OK;
}

extern "C" int sc_qd_get_state(sc_qd* h, double* psi_h, double* slip_h, double* V_h,
                               double* tau_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  SC_HIP(hipStreamSynchronize(q.ctx->c.stream));
  const size_t bytes = (size_t)q.Nz * sizeof(double);
  if (psi_h) SC_HIP(hipMemcpy(psi_h, q.vars["psi"].var, bytes, hipMemcpyDeviceToHost));
  if (slip_h) SC_HIP(hipMemcpy(slip_h, q.vars["slip"].var, bytes, hipMemcpyDeviceToHost));
  if (V_h) SC_HIP(hipMemcpy(V_h, q.V_d, bytes, hipMemcpyDeviceToHost));
  if (tau_h) SC_HIP(hipMemcpy(tau_h, q.tau_d, bytes, hipMemcpyDeviceToHost));
  return SC_OK;
}

extern "C" int sc_qd_get_u(sc_qd* h, double* u_h) {
  if (!h || !u_h) return fail(SC_ERR_ARG, "null argument");
  auto& q = *h->q;
  const int e = q.materialize_u();
  if (e) return e;
  SC_HIP(hipStreamSynchronize(q.ctx->c.stream));
  SC_HIP(hipMemcpy(u_h, q.u_d, (size_t)q.N * sizeof(double), hipMemcpyDeviceToHost));
  return SC_OK;
}

extern "C" int sc_qd_ddt(sc_qd* h, double t, double* dpsi_h, double* dslip_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  auto& P = q.vars["psi"];
  auto& S = q.vars["slip"];
  int e = q.d_dt_vars(t);
  if (e) return e;
  SC_HIP(hipStreamSynchronize(q.ctx->c.stream));
  if ((e = q.check_status())) return e;
  const size_t bytes = (size_t)q.Nz * sizeof(double);
  if (dpsi_h) SC_HIP(hipMemcpy(dpsi_h, P.dvar, bytes, hipMemcpyDeviceToHost));
  if (dslip_h) SC_HIP(hipMemcpy(dslip_h, S.dvar, bytes, hipMemcpyDeviceToHost));
  return SC_OK;
}

extern "C" int sc_qd_integrate(sc_qd* h, const sc_qd_run* run, double* tEnd_h, int* steps_h,
                               double* dts_h, int dtsCap) {
  if (!h || !run) return fail(SC_ERR_ARG, "null argument");
  if (run->integrator != 0 && run->integrator != 1) return fail(SC_ERR_ARG, "integrator must be RK32 (0) or RK43 (1)");
  if (!(run->tol >= 1e-14)) return fail(SC_ERR_ARG, "timeStepTol must be >= 1e-14");
  if (run->deferChecks < 0 || run->deferChecks > 2) return fail(SC_ERR_ARG, "deferChecks must be 0, 1 or 2");
  if (run->stepControl != 0 && run->stepControl != 1) return fail(SC_ERR_ARG, "stepControl must be 0 or 1");
  double tEnd = 0;
  int steps = 0;
  int e = h->q->integrate(*run, &tEnd, &steps, dts_h, dtsCap);
  if (tEnd_h) *tEnd_h = tEnd;
  if (steps_h) *steps_h = steps;
  return e;
}

extern "C" int sc_qd_get_counters(sc_qd* h, sc_qd_counters* c) {
  if (!h || !c) return fail(SC_ERR_ARG, "null argument");
  *c = h->q->counters;
  c->graph_builds = h->q->graphBuilds;
  return SC_OK;
}
