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

// Kernel arguments of the fused field kernels (qd_kernels.hip)
struct VarsLinComb {
  int nv = 0, nt = 0;
  int n[kMaxVars] = {};
  double* out[kMaxVars] = {};
  const double* base[kMaxVars] = {};
  const double* f[kMaxVars][6] = {};
  double c[6] = {};
  // under a step gate c[i] is a template: the coefficient is (c[i] h) / cd[i]
  // with h the control block's step (the order of the host's h/6.0, 2*h/3.0)
  double cd[6] = {1, 1, 1, 1, 1, 1};
};
// The end of an RK32 attempt in one launch of one block (k_step_end): the
// residual test the last stage's deferred guess left (fin), the lo/hi pair
// and its error sums, the step control: error norm, acceptance, the PID
// step size (odeSolver.cpp:395-445, 449-501, 505-627), the accepted
// solution into var; flags = (fault status, missed guess)
struct StepEnd {
  PendingFinish fin{};
  VarsLoHi L{};
  StepCtl* ctl = nullptr;
  double* err = nullptr;
  const int* flags = nullptr;
  VarsCopy cp{};
};
int launch_step_end(hipStream_t s, const StepEnd& e, StepGate gate);

// One kernel per RK stage of the device loop (qd_kernels.hip, k_qd_stage):
// the stage combination, the projections of the stage's rhs from the
// per-basis plan (k_stage_plan), the fault rates and the guess's residual,
// with b formed on the band rows; no rhs launch.  Buffers: the plan (PL, PR,
// sums), and per stage position (stage 2, stage 3, rates at the accepted
// state) the residual partials and scalars, so that a launch can test the
// previous stage's residual while it writes its own.
struct QdStageBufs {
  double* plan = nullptr;
  double* parts = nullptr;
  void* scal = nullptr;
  // the second state of the device loop's ping-pong: psi, slip, their rates
  // (Nz each)
  double* alt = nullptr;
  // double-double sums of the guess residual (QL, QR, S3, G; the Gram
  // matrix's block partials)
  double* gram = nullptr;
  int nz = 0, nstream = 0;
};
struct QdStageIn {
  double tol2 = 0;
  QdFuse fuse{};
  VarsLinComb L{};
  int psiIdx = -1, slipIdx = -1;
  const double* slipVar = nullptr;
  double fts = 1, vL = 0;
  int remote = 0;
  const double* rfix = nullptr;
  double* bcL = nullptr;
  double* bcR = nullptr;
  StepGate gate{};
  double tfrac = 0;
  double t0 = 0;  // the time without a gate
  PendingFinish fin{};
  // the attempt's end in the same launch (the rates at the accepted state):
  // gate.ctl is the attempt's control block, ctlOut the next attempt's;
  // fields 0 psi, 1 slip: var, the three stage rates, the pair's templates
  // and the error slots (-1: not in the norm)
  bool end = false;
  StepCtl* ctlOut = nullptr;
  int* flags = nullptr;
  // the previous attempt's end's root-solve status (read); this launch's
  // own root solves report into the other word of the pair (fuse.a.status_d)
  const int* flagEnd = nullptr;
  PendingFinish fin3{};
  const double* pv[2] = {};
  double* pvOut[2] = {};
  const double* pf[2][3] = {};
  double clo[2] = {}, cdlo[2] = {1, 1}, chi[3] = {}, cdhi[3] = {1, 1, 1};
  int selSlot[2] = {-1, -1};
};
// buffers of the stage kernel for a fault of Nz nodes and nBandRows band rows
int stage_bufs_ensure(QdStageBufs& b, int Nz, int nBandRows);
void stage_bufs_free(QdStageBufs& b);
// a basis and Nz within the kernel's LDS tables
bool stage_fits(const Ctx& c);
// r = b - Au over the grid: block partials of r.r and b.b (2 per block, nb
// blocks of 256 threads; nb <= 1024)
int launch_resid_norms(const double* b, const double* Au, size_t N, double* parts, int nb, hipStream_t s);
// the residual test a stage launch at position pos leaves
PendingFinish stage_finish(const QdStageBufs& b, int pos, int* miss);
int launch_stage_plan(const Ctx& c, QdStageBufs& b, const double* rfix, hipStream_t s);
// rec != nullptr: append the launch's arguments (kernel, grid, argument
// bytes) to *rec instead of launching
// evStart / evStop (optional): events that take the kernel's own start and
// end (hipExtLaunchKernelGGL), for per-launch timing
int launch_qd_stage(const Ctx& c, const QdStageBufs& b, int pos, const QdStageIn& in, hipStream_t s,
                    std::vector<unsigned char>* rec = nullptr, hipEvent_t evStart = nullptr,
                    hipEvent_t evStop = nullptr);
// fin (optional): a pending residual test run as one more block
int launch_vars_lohi_err(hipStream_t s, const VarsLoHi& L, const PendingFinish* fin = nullptr,
                         StepGate gate = StepGate{});
// plan (optional, from proj_frame_plan): also write the projections of the
// rhs on the basis; *mnb gets the number of partial blocks (0 without).
// pre (optional): the Runge-Kutta combination that forms this stage's state
// (done in the same launch with a plan, before it without), field preSlip
// being the slip the rhs is formed from; fin (optional): a pending residual
// test, run in the same launch with a plan, before it without.
int launch_qd_rhs_frame(const Ctx& c, double* rhs, const double* slip, const double* shift, double t,
                        double vL, double fts, int remote, double* bcL, double* bcR, const double* bcT,
                        const double* bcB, hipStream_t s, const ProjFramePlan* plan = nullptr,
                        int* mnb = nullptr, const VarsLinComb* pre = nullptr, int preSlip = -1,
                        const PendingFinish* fin = nullptr, StepGate gate = StepGate{}, double tfrac = 0.0);
int launch_qd_fault(const Ctx& c, const double* u, const sc_fault_qd_args& a, double* dslip, hipStream_t s);
int launch_vars_lincomb(hipStream_t s, const VarsLinComb& L);
// out[0] = max over k of (eta[k] V[k]) / tauQS[k] (the switching ratio of the
// QD-FD mediator); NaN entries are ignored
int launch_max_ratio(hipStream_t s, const double* eta, const double* V, const double* tauQS, int n,
                     double* out);

struct RkState {
  int ord = 3, control = 1, normType = 0, stepCount = 0;
  int pidBase = 0;  // PID needs four steps after this one (the controller's start)
  double tol = 1e-8, kappa = 0.9, minDt = 0, maxDt = 1e30, deltaT = 0, currT = 0;
  double errA[2] = {0, 0};
};

// Power-law viscoelastic terms (StrikeSlip_PowerLaw_qd with dislocation creep)
struct PowerLawState {
  double *T = nullptr, *A = nullptr, *B = nullptr, *n = nullptr;  // N each
  double* K = nullptr;  // 1e3 A exp(-B/T): the creep law's factor that does not depend on stress
  double cap = 1e30;
  double *effVisc = nullptr, *sxy = nullptr;                       // last d_dt
};

// Output series and checkpoints (output.cpp)
struct QdOutput {
  std::string dir;
  int stride1D = 0, stride2D = 0, ckpt = 0, interval = 1;
  int ckptNumber = 0;
  bool append = false;                  // series start appended (a restart), else truncated
  bool started1D = false, started2D = false;
  double* sxy_d = nullptr;              // muxDy u for the 2-D series (N)
  double* plf_d = nullptr;              // the power law's eight 2-D fields (8 N)
};

struct QdDriver;
// Called by integrate after every accepted step, with the rates at the new
// state computed (the mediator's timeMonitor); *stop ends the integration.
struct QdStepHook {
  virtual ~QdStepHook() {}
  virtual int after_step(QdDriver& q, double t, double dt, int step, int* stop) = 0;
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
  static constexpr int kErrSlots = 16;  // 3 error sums per selected field (<= 4), then 3 scratch
  // + fault status, missed-guess flag and the attempt's-end status pair (4 ints)
  static constexpr int kChkDoubles = 18;
  static constexpr int kFlagInts = 4;
  static constexpr int kShortField = 16384;  // fields up to this length: one block each
  double* err_d = nullptr;              // the read-back block
  double* err_part_d = nullptr;         // block partials of the error sums
  static constexpr int kErrBlocks = 512;
  int* status_d = nullptr;             // inside the read-back block
  int* miss_d = nullptr;               // set by a deferred guess check that missed
  // status_d[2 + p]: root solves of the attempt's end launched at parity 1 - p
  // (read by the next attempt's end, never by the launch that writes it)
  int* endStatus_d = nullptr;
  // the fault status of a read-back of the flag words: stage launches', and
  // both attempt's-end words
  static int fault_status(const int* w) { return w[0] | w[2] | w[3]; }
  bool defer = false;                  // solves may defer their check (integrate only)
  // Deferring pays only while the projected guesses keep meeting the
  // tolerance: after a miss the driver checks every solve again until
  // kDeferAfter guesses in a row have hit.
  static constexpr int kDeferAfter = 8;
  int hits = 0;                        // consecutive guesses that met the tolerance
  int pendingDeferred = 0;             // deferred solves since the last read-back
  bool defer_ok() const { return desc.solve.reuse > 0 && hits >= kDeferAfter; }
  void note_checked(bool missed) {     // a read-back of the missed-guess flag
    hits = missed ? 0 : hits + pendingDeferred;
    pendingDeferred = 0;
  }
  void* host_scratch = nullptr;
  std::map<std::string, VecSet> vars;
  int errMask = 3;                      // RK error fields: 1 psi, 2 slip, 4 gVxy, 8 gVxz
  std::unique_ptr<PowerLawState> pl;    // set for the power-law mediator
  sc_qd_counters counters{};
  std::vector<double> mu_host, y_host, z_host;  // grid and modulus, for a second context
  sc_qd_run file_run{};
  bool have_file_run = false;
  std::unique_ptr<QdOutput> out;       // set: integrate calls monitor()
  QdStepHook* hook = nullptr;           // set: integrate calls hook->after_step()
  // step count integrate starts from; run.maxSteps counts from it
  // (OdeSolver::setInitialStepCount, odeSolver.cpp:25-29)
  int stepBase = 0;
  bool pidFresh = false;  // the PID controller counts its four P steps from stepBase
  // the discrete adjoint's second context (its own solve workspace and
  // projection basis) and the boundary operator (1/fts) R_L^T A^-1 S^T of
  // its reverse pass (Nz x Nz, row-major), kept across gradients; the keys
  // say which operator (ctx opGen) and adjoint solve options they belong to
  sc_ctx* adjCtx = nullptr;
  int adjCtxGen = -1;
  double* adjB = nullptr;
  int adjBGen = -1, adjBPrec = -1, adjBMode = -1;
  double adjBRtol = 0.0;

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
  // pre (linear-elastic mediator only): the Runge-Kutta combination that
  // forms psi and slip (field preSlip) of this stage, done inside the rhs launch
  int d_dt(double t, const double* psi, const double* slip, double* dpsi, double* dslip,
           const double* gxy = nullptr, const double* gxz = nullptr, double* dgxy = nullptr,
           double* dgxz = nullptr, const VarsLinComb* pre = nullptr, int preSlip = -1,
           StepGate gate = StepGate{}, double tfrac = 0.0);
  int d_dt_vars(double t);
  // the fault-rate arguments for a stage's psi and psi rate
  sc_fault_qd_args fault_args(const double* psi, double* dpsi) const;
  // a stage kernel's inputs: L the combination of stage st (nullptr: the
  // rates at the current state)
  QdStageIn stage_input(const VarsLinComb* L, int st, StepGate g, double tfrac, PendingFinish f);
  // times the stage kernel in its accepted-state form (sc_qd_bench_stage)
  int bench_stage(double t, int reps, double* ms, double* words);              // on var -> dvar of every integrated variable
  // the stage kernel's guess residual at time t (r.r, b.b from its Gram
  // form) and the same two sums by an explicit pass over the grid
  // (sc_qd_guess_residual)
  int guess_residual(double t, double* out4);
  int check_status();
  // fault status and missed-guess flag in one read-back (the flags are left set)
  int read_flags(int* status, int* miss);
  double step_size(const RkState& r, double totErr) const;
  // lo and hi of the embedded pair (lo = var + sum clo f, hi likewise) and the error
  int lohi_error(const RkState& r, int ntlo, const int* filo, const double* clo, int nthi,
                 const int* fihi, const double* chi, double* totErr, bool* missed);
  int lincomb_all(const char* out_sel, int stage, int nt, const int* fidx, const double* c);
  int stage(int st, double t);
  // k[st] = var + sum_i c[i] f_i, then the rates there (one launch fewer
  // than lincomb_all + stage for the linear-elastic mediator)
  int stage_comb(int st, double t, int nt, const int* fidx, const double* c);
  // The residual test of the last deferred guess, postponed into the next
  // rhs or lo/hi launch (PendingFinish); flush_finish launches it on its own
  // before anything that reads the flags.
  PendingFinish fin{};
  bool finPending = false;
  int flush_finish();
  // u_d of the last deferred solve is formed only on demand (solve's x_lazy)
  bool uLazy = false;
  int materialize_u();
  // rows next to the y-sides where the QD rhs is nonzero (qd_band_rows)
  int band_rows() const {
    int lo, hi;
    qd_band_rows(Ny, &lo, &hi);
    return lo + hi;
  }
  // RK32 step control on the device (linear-elastic mediator, deferred
  // checks, no replay / output / hook): the host enqueues attempts in
  // batches and reads the control block once per batch.
  StepCtl* ctl_d = nullptr;
  StepCtl* ctl_h = nullptr;           // pinned
  double* ctlDts_d = nullptr;
  int ctlDtsCap = 0;
  QdStageBufs stageBufs;  // the device loop's one-kernel stages
  static constexpr int kSlotsPerBatch = 24;
  // attempts captured in one graph: a relaunch of a graph waits for its
  // previous run (about 8 us between the two on the device), so a full batch
  // is one launch of kAttemptsPerGraph attempts and only a batch's tail is
  // made of launches of the short graph (kAttemptsPerTail attempts)
  static constexpr int kAttemptsPerGraph = kSlotsPerBatch;
  static constexpr int kAttemptsPerTail = 4;
  // The instantiated graphs of the one-kernel stage loop (the full batch and
  // the short tail), kept across device_steps calls; stageKey holds the
  // arguments of every launch of the short graph (the full one repeats them),
  // and both are captured again only when they change (a new basis size or
  // new basis vectors).
  hipGraph_t stageGraph = nullptr;
  hipGraphExec_t stageExec = nullptr;
  hipGraph_t tailGraph = nullptr;
  hipGraphExec_t tailExec = nullptr;
  std::vector<unsigned char> stageKey;
  long long graphBuilds = 0;
  void drop_stage_graph();
  // Set during sc_qd_integrate_timed: the device loop launches its attempts
  // one by one (no graph) with HIP events around each stage launch and sums
  // their durations over the attempts that ran (positions 0, 1: the plain
  // stage kernels; 2: the attempt's end).
  struct StageTiming {
    double ms[3] = {0, 0, 0};
    long long attempts = 0;
  };
  StageTiming* timing = nullptr;
  // runs steps from the current one (r.stepCount, started) on the device
  // until the run ends (*finished) or a missed guess / fault failure halts
  // it (*halted: the host redoes the current step with checked solves)
  int device_steps(RkState& r, double finalT, int maxSteps, double& newDt, double* dts, int dtsCap, int& nOut,
                   bool* finished, bool* halted);
  int integrate(const sc_qd_run& run, double* tEnd, int* steps, double* dts, int dtsCap);
  // the rates at time t were computed with deferred guess checks: read the
  // flags and redo them with checked solves if a guess missed
  int settle(double t);
  // max over the fault of eta V / tauQS of the last d_dt (settled first)
  int max_ratio(double t, double* R);
  // timeMonitor (output.cpp): writes what is due at this step; *stop set by
  // a checkpoint.  initial: the call before the first step (no rates to settle)
  int monitor(double t, double dt, int step, const double* errA, double newDt, int maxSteps,
              double maxTime, bool initial, int* stop);
  int set_output(const sc_qd_output* o);
  int load_checkpoint(sc_qd_run* run, int* ckptNumber);
};

}  // namespace sc

struct sc_qd {
  std::unique_ptr<sc::QdDriver> q;
};
