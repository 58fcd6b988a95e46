// Internal declarations of libscycle_hip (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/scycle_hip.h"

namespace sc {

struct FdDiag;
constexpr int kBand = 9;     // band width of every 1-D table (offsets -4..+4)
constexpr int kFrame = 8;    // frame width handled by the generic kernel

// -------------------------------------------------------- error handling --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define SC_HIP(call)                                                          \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::sc::fail(SC_ERR_HIP, std::string(#call " -> ") + hipGetErrorString(_e)); \
  } while (0)

// ------------------------------------------------------------ 1-D tables --
// One direction of the SBP operator, in banded form: entry (i, c) of a 1-D
// matrix sits at [i*kBand + (c - i + 4)].  Built from the reference's 1-D
// stencil definitions (spmat.cpp:310-627).
struct Dir1D {
  int n = 0, order = 4;
  double h = 1, alphaD = 0, h11 = 0;
  std::vector<double> H, Hinv;       // n
  std::vector<double> D1;            // n*9
  std::vector<double> DR, DRT;       // D3 (order 4) / D2 (order 2) and transpose
  std::vector<double> D4, D4T;       // order 4 only (zeros otherwise)
  std::vector<double> wR, w4;        // C3/(18h) or C2*h^3/4 ; C4/(144h)
  std::vector<double> BS0, BSN;      // boundary rows of BS: cols 0..8 / n-9..n-1
};
int build_dir(int order, int n, double h, bool compatible, Dir1D& d);

// ---------------------------------------------------- device parameters --
// Raw pointers into context-owned device memory, passed by value to kernels.
template <typename T>
struct DevDir {
  int n;
  const T *H, *Hinv, *D1, *DR, *DRT, *D4, *D4T, *wR, *w4, *BS0, *BSN;
  const T* met;    // yq (y) or zr (z), n values
  const T* imet;   // qy (y) or rz (z), n values
  T alphaD;
  // interior constants (row n/2 of the tables)
  T d1[5], dr[5], d4[5], drt[5], d4t[5], wr, w4c, Hc;
  // the interior rows in the factored form the tile kernel evaluates (the
  // stencils' integer parts are literals there): D1 u = ia (u[+1] - u[-1]) +
  // ib (u[+2] - u[-2]); the same with H: Ha, Hb; the DR flux scale g3s (order
  // 4: wR/2, the c3 average's 1/2 folded in; order 2: wR times the D2 scale),
  // the D4 flux scale g4s = w4, and rts, the scale of the transposed D2 of
  // order 2
  T ia, ib, Ha, Hb, g3s, g4s, rts;
};

// Boundary-closure tables of one direction and one end, in the layout the
// frame-coefficient kernel stages into LDS: flux rows mb..mb+11 (D1, DR, D4 and their
// weights) and output rows jb..jb+7 (D1, DR^T, D4^T, H), with mb = jb = 0 at
// the low end and mb = n-12, jb = n-8 at the high end, plus the SAT row of BS.
template <typename T>
struct ClosureTab {
  T fD1[12 * kBand], fDR[12 * kBand], fD4[12 * kBand];
  T oD1[kFrame * kBand], oDRT[kFrame * kBand], oD4T[kFrame * kBand];
  T wR[12], w4[12], H[kFrame], BS[kBand];
};
static_assert((4 * sizeof(ClosureTab<float>)) % 16 == 0, "closure tables copy as 16-byte words");

template <typename T>
struct DevOp {
  int Ny, Nz, order;
  int doY, doZ;                  // laplace type "yz" / "y" / "z"
  int bcR, bcT, bcL, bcB;
  const T* mu;                   // N
  DevDir<T> y, z;
  const ClosureTab<T>* ctab;     // [y low, y high, z low, z high] (n > 2*kFrame only)
};

// ----------------------------------------------------------- context -----
struct DevTables {                 // device copies of one direction (one precision)
  void* base = nullptr;
  size_t bytes = 0;
};

struct Ctx {
  sc_grid_desc d{};
  int N = 0;
  hipStream_t stream = nullptr;
  Dir1D ty, tz;
  std::vector<double> yq, zr, qy, rz;  // 1-D metric (Ny / Nz)
  std::vector<double> mu_h;            // host copy
  // device memory
  double* mu64 = nullptr;
  float* mu32 = nullptr;
  void* tab64 = nullptr;
  void* tab32 = nullptr;
  void* fcoef64 = nullptr;             // assembled frame operator, 34 x frame nodes
  void* fcoef32 = nullptr;
  void* ctab64 = nullptr;              // packed closure tables (ClosureTab x4)
  void* ctab32 = nullptr;
  DevOp<double> op64{};
  DevOp<float> op32{};
  double* diag64 = nullptr;            // diag(A)
  double* dinv64 = nullptr;            // 1/diag(A)
  float* dinv32 = nullptr;
  // solver workspace (lazily allocated)
  double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *d32 = nullptr;
  double* z64 = nullptr;               // preconditioned residual
  double* zold64 = nullptr;            // previous z (flexible CG)
  struct FdDiag* fd = nullptr;         // fast-diagonalisation factors (lazy)
  bool fd_stale = false;
  int opGen = 0;                       // bumped whenever mu or the BC types change
  double* partials = nullptr;          // reduction scratch
  // previous solutions kept for the initial guess (ring buffer of A-orthogonal
  // vectors y_i with weights w_i = 1 / (y_i^T (-A) y_i) on the device)
  double* projU[SC_REUSE_MAX] = {};    // ring of the last projCap solutions that iterated
  double* projZU[SC_REUSE_MAX] = {};   // A u for each of them
  double* projT[SC_REUSE_MAX] = {};    // fault traction mu Dy y on row j = 0 (Nz each)
  double* projY[SC_REUSE_MAX] = {};    // A-orthogonal basis of the ring's span
  double* projZ[SC_REUSE_MAX] = {};    // A y for each basis vector
  double* projW = nullptr;             // w[SC_REUSE_MAX], coef[SC_REUSE_MAX]
  int projCap = 0, projCount = 0, projNext = 0;
  void* scal = nullptr;                // CgScalars (device)
  void* scal_h = nullptr;              // pinned host mirror
  int nPartial = 0;
};

int ctx_upload_tables(Ctx& c);
int ctx_upload_mu(Ctx& c);
int ctx_compute_diag(Ctx& c);

// ------------------------------------------------------ kernel launchers --
int launch_apply64(const Ctx& c, const double* in, double* out, hipStream_t s);
int launch_apply32(const Ctx& c, const float* in, float* out, hipStream_t s);
int launch_diag(const Ctx& c, double* out, hipStream_t s);
// one explicit wave-equation step of the body (propagateWaves): uNext from u,
// uPrev; D2u on the fault column into d2uFault (Nz)
template <typename T>
struct WaveArgs {
  const T* u;
  const T* uPrev;
  T* uNext;
  const T* rho;  // rho, or the per-node scale of launch_wave_c1 when rhoIsC1
  const T* ay;
  T* d2uFault;
  T dt;
  int rhoIsC1 = 0;
  // the fault's step fused into the launch (Fault_fd::d_dt at j = 0; needs
  // rhoIsC1), or null
  const sc_fault_fd_args* fault = nullptr;
};
// c1 = dt^2/rho * Hy^-1 Hz^-1 (* qy rz on a variable grid) per node: the
// wave update's whole scale of A u, formed once for a run at constant dt
int launch_wave_c1(const Ctx& c, const double* rho, double* c1, double dt, hipStream_t s);
int launch_wave_c1(const Ctx& c, const float* rho, float* c1, float dt, hipStream_t s);
int launch_wave64(const Ctx& c, const WaveArgs<double>& a, hipStream_t s);
int launch_wave32(const Ctx& c, const WaveArgs<float>& a, hipStream_t s);
// frame_only: write only the nodes within kFrame of a boundary, for a
// caller whose rhs is zero elsewhere
// forms the frame coefficients of the current operator (after any change of mu or BC types)
int ctx_build_frame(Ctx& c);
int launch_set_rhs(const Ctx& c, double* rhs, const double* bcL, const double* bcR,
                   const double* bcT, const double* bcB, hipStream_t s, bool frame_only = false);
// which: 0 Dy, 1 muxDy, 2 Dz, 3 muxDz, 4 H, 5 Hinv, 6 Dyxmu, 7 Dzxmu
int launch_deriv(const Ctx& c, int which, const double* in, double* out, hipStream_t s);
// the boundary restrictions scaled by the 1D inverse norm, out on the grid:
// 0 Hy^-1 e0y, 1 Hy^-1 eNy (in: Nz values of a face), 2 Hy^-1 E0y,
// 3 Hy^-1 ENy, 4 Hz^-1 E0z, 5 Hz^-1 ENz (in: a grid field)
int launch_bnd(const Ctx& c, int which, const double* in, double* out, hipStream_t s);
// sxy on the fault line j = 0 only: out[k] = mu(0,k)*qy[0]*(D1y u)(0,k)
int launch_fault_sxy(const Ctx& c, const double* u, double* out, hipStream_t s);

// Fault rates fused into a deferred projected guess (the QD step): with the
// guess accepted, the fault traction is sum_m c_m (mu Dy y_m) on row 0, so
// the root solves need only the guess coefficients and run in blocks of
// the guess kernel beside the streaming ones.  dslip gets V.
struct QdFuse {
  sc_fault_qd_args a;
  double* dslip;
};

// The fields of an RK pair (kernel arguments of qd_kernels.hip): lo and hi = base + sum_t c[t] f[t], their error sums
// [sum (lo-hi)^2, sum hi^2, max |lo-hi|/hi] into out[v][0..2]
constexpr int kMaxVars = 4;
struct VarsLoHi {
  int nv = 0, ntlo = 0, nthi = 0;
  int n[kMaxVars] = {};
  const double* base[kMaxVars] = {};
  double* lo[kMaxVars] = {};
  double* hi[kMaxVars] = {};
  const double* flo[kMaxVars][6] = {};
  const double* fhi[kMaxVars][6] = {};
  double clo[6] = {}, chi[6] = {};
  double cdlo[6] = {1, 1, 1, 1, 1, 1}, cdhi[6] = {1, 1, 1, 1, 1, 1};  // as VarsLinComb::cd
  double* out[kMaxVars] = {};
};
// dst_v = src_v for every field (the accepted solution into var)
struct VarsCopy {
  int nv = 0;
  int n[kMaxVars] = {};
  double* dst[kMaxVars] = {};
  const double* src[kMaxVars] = {};
};

// Runge-Kutta step control on the device (the QD driver's device loop, see
// qd.cpp): the state of odeSolver.cpp's RK32 loop between attempts, updated
// by one control kernel per attempt.  The step kernels read t and dt from it.
struct StepCtl {
  double t, dt, newDt, tol, minDt, maxDt, finalT, kappa, errA0, errA1;
  int stepCount, maxSteps, pidBase, control, normType, ord;
  int attempts, halt, accepted, done, rejected, nOut, dtsCap, nsel;
  int selN[4];   // lengths of the selected error fields, in error-slot order
  double* dts;   // accepted step sizes, indexed by nOut (dtsCap entries)
};
// Which step kernels run: mode 0 always; 1 a stage of the current attempt
// (not once halted or done); 2 the rates after an accepted attempt.
struct StepGate {
  const StepCtl* ctl = nullptr;
  int mode = 0;
};

// Rows of the QD driver's rhs band next to each y-side (the rhs vanishes
// elsewhere: its top and bottom data are zero)
inline void qd_band_rows(int Ny, int* nLow, int* nHigh) {
  *nLow = Ny < kBand ? Ny : kBand;
  *nHigh = Ny > kBand ? (Ny - kBand < kBand ? Ny - kBand : kBand) : 0;
}

// The residual test of a deferred projected guess, left for a later launch
// (cg_dev.hpp): the partials of nb blocks, the scalars, the missed-guess flag.
struct CgScal;
struct PendingFinish {
  CgScal* sc = nullptr;
  const double* partials = nullptr;
  int nb = 0;
  int* flag = nullptr;
};
// launches a pending test on its own (one block)
int launch_pending_finish(hipStream_t s, const PendingFinish& f);

// miss_flag/deferred: see cg.hip (deferred residual check of a projected
// guess); rhs_frame: rhs is zero off the nodes within kFrame of a boundary;
// pre_mnb > 0: the caller already wrote the projections y_m . rhs as pre_mnb
// block partials (see proj_frame_plan); fuse (deferred solves only): also
// compute the fault rates from the guess, *fused tells whether it did;
// postpone (deferred solves only): the residual test is not launched but
// described there, for the caller to run in a later launch before anything
// else writes the reduction partials.
int solve(Ctx& c, const double* rhs, double* u, const sc_solve_opts& o, sc_solve_stats& st,
          int* miss_flag = nullptr, bool* deferred = nullptr, bool rhs_frame = false, int pre_mnb = 0,
          const QdFuse* fuse = nullptr, bool* fused = nullptr, PendingFinish* postpone = nullptr,
          StepGate gate = StepGate{}, bool* x_lazy = nullptr);
// x_lazy (deferred, fused solves): u is left unwritten (*x_lazy = true) and
// proj_materialize writes it from the guess's coefficients when needed;
// valid until the next solve on the context
int proj_materialize(Ctx& c, double* u);

// For a caller that forms a frame-supported rhs and its projections on the
// basis in one kernel: M basis vectors Y (M = 0: no projected guess for this
// solve), where to write the block partials (M per block, block order) and
// the most blocks that may write them.
// Up to SC_REUSE_MAX device vectors (kernel argument)
struct VecList {
  const double* v[SC_REUSE_MAX];
};
struct ProjFramePlan {
  int M = 0;
  VecList Y{};
  double* mpart = nullptr;
  int maxBlocks = 0;
};
int proj_frame_plan(Ctx& c, const sc_solve_opts& o, ProjFramePlan* plan);
int bench_proj_guess(Ctx& c, int reps, double* ms, double* words);
// z = M^-1 r with one of the solve's preconditioners (SC_PREC_*)
int precond_apply(Ctx& c, int precond, const double* r, double* z);
// power law: rhs += B gxy + C gxz; strains -> stresses -> viscosity -> rates
// the boundary data of the power-law d_dt, formed in the source kernel:
// bcL = slip/fts; bcR = vL t/fts + shift when remote, else bcR as given
struct PlRhsArgs {
  const double *slip, *shift;
  double t, vL, fts;
  int remote;
  double* bcR;  // in (fixed datum) or out (remote)
  const double *bcT, *bcB;
  double* bcLOut;
};
// rhs = setRhs + the viscous source (the boundary data, setRhs and
// computeViscStrainSourceTerms in one launch)
int launch_pl_source(const Ctx& c, const double* gxy, const double* gxz, double* rhs, const PlRhsArgs& b,
                     hipStream_t s);
// the power law's 2-D output fields at the current state (N each)
struct PlFields {
  const double *u, *gVxy, *gVxz, *T, *A, *B, *n;
  double cap;
  double *gTxy, *gTxz, *sxy, *sxz, *sdev, *effVisc, *invDisl, *dgVdev;
};
int launch_pl_fields(const Ctx& c, const PlFields& f, hipStream_t s);
// the creep law's stress-free factor K = 1e3 A exp(-B/T) per node, formed once
int launch_pl_factor(const Ctx& c, const double* A, const double* B, const double* T, double* K, hipStream_t s);
int launch_pl_strain(const Ctx& c, const double* u, const double* gVxy, const double* gVxz,
                     const double* K, const double* n, const double* bcL, const double* bcR, double cap,
                     double* dgxy, double* dgxz, double* tauQS, double* effVisc, double* sxyOut,
                     const sc_fault_qd_args* fault, double* dslip, const PendingFinish* fin, hipStream_t s);
int fault_qd_rates(hipStream_t s, const sc_fault_qd_args& a);
// FD fault step; uBody (may be null) receives the fault displacement
int fault_fd_step64(hipStream_t s, const sc_fault_fd_args& a, double* uBody);

}  // namespace sc

struct sc_ctx {
  sc::Ctx c;
};
