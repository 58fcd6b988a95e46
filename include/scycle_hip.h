/*
 * scycle_hip.h -- C ABI of the MI355X (gfx950) SCycle step library,
 * libscycle_hip.so.
 *
 * Every entry point returns int: 0 on success, a non-zero SC_ERR_* code
 * otherwise (the reference propagates PetscErrorCode the same way), and
 * sc_last_error() returns a message for the last failure on the calling
 * thread.  No C++ exception crosses this boundary.
 *
 * Pointers named *_d are device pointers (hipMalloc'd, FP64 unless the
 * argument says otherwise); *_h are host pointers.  All work is enqueued on
 * the context's own HIP stream; sc_ctx_sync() waits for it.  A context must
 * not be used from two host threads at once (the reference is single-threaded
 * SPMD per rank, too).
 *
 * Layout of every 2-D field: index I = j*Nz + k (z fastest, j along y); the
 * fault is j = 0, the surface k = 0 (domain.cpp:379-383, 449-501).
 * Units are the reference's: km, GPa, g/cm^3, MPa, m, m/s, s.
 */
#ifndef SCYCLE_HIP_H
#define SCYCLE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SC_OK 0
#define SC_ERR_ARG 1        /* invalid argument / unsupported option      */
#define SC_ERR_HIP 2        /* HIP runtime error                          */
#define SC_ERR_NOCONV 3     /* iterative solve or root solve not converged */
#define SC_ERR_NONFINITE 4  /* NaN/Inf produced (reference asserts)       */
#define SC_ERR_IO 5         /* file / parse error                         */

#define SC_DIRICHLET 0
#define SC_NEUMANN 1

#define SC_FP64 0
#define SC_FP32 1

typedef struct sc_ctx sc_ctx;

/* Grid and operator options; fields mirror the setters of SbpOps
 * (sbpOps.hpp:42-104) and the Domain keys that select the operator
 * (domain.cpp:167-188). */
typedef struct sc_grid_desc {
  int Ny, Nz;
  double Ly, Lz;
  int order;           /* 2 or 4                                          */
  int var_grid;        /* 1 = variableGridSpacing, 0 = constantGridSpacing */
  int compat;          /* 0 = fullyCompatible, 1 = compatible             */
  int laplace;         /* 0 = "yz", 1 = "y", 2 = "z" (setLaplaceType)      */
  int bcR, bcT, bcL, bcB; /* SC_DIRICHLET / SC_NEUMANN (setBCTypes order R,T,L,B) */
} sc_grid_desc;

const char* sc_last_error(void);
int sc_version(int* major, int* minor);

/* Device memory helpers (the caller may also pass its own allocations). */
int sc_malloc(void** ptr_d, unsigned long long bytes);
int sc_free(void* ptr_d);
int sc_memcpy_h2d(void* dst_d, const void* src_h, unsigned long long bytes);
int sc_memcpy_d2h(void* dst_h, const void* src_d, unsigned long long bytes);

/* Creates the operator context: replaces the SbpOps constructor +
 * setGrid/setBCTypes/setMultiplyByH(1)/setLaplaceType/computeMatrices
 * sequence of LinearElastic::setUpSBPContext (linearElastic.cpp:426-461).
 * y_h, z_h: N-length host coordinate fields (NULL => y = Ly*q, z = Lz*r);
 * mu_h: N-length shear modulus.  The grid must be a tensor-product grid
 * (y depends on j only, z on k only); otherwise SC_ERR_ARG. */
int sc_ctx_create(const sc_grid_desc* desc, const double* y_h, const double* z_h,
                  const double* mu_h, sc_ctx** out);
int sc_ctx_destroy(sc_ctx* ctx);
int sc_ctx_sync(sc_ctx* ctx);
/* Returns the context's hipStream_t (as void*). */
int sc_ctx_stream(sc_ctx* ctx, void** stream);

/* SbpOps::updateVarCoeff (sbpOps.hpp:63; sbpOps_m_varGrid.cpp:531-560). */
int sc_update_var_coeff(sc_ctx* ctx, const double* mu_h);
/* SbpOps::changeBCTypes (sbpOps.hpp:58; sbpOps_m_varGrid.cpp:187-213). */
int sc_change_bc_types(sc_ctx* ctx, int bcR, int bcT, int bcL, int bcB);
/* SbpOps::geth11 (sbpOps.hpp:66). */
int sc_geth11(sc_ctx* ctx, double* h11y, double* h11z);
/* Diagonals of SbpOps::getCoordTrans yq (Ny values) and zr (Nz values)
 * (sbpOps.hpp:79; sbpOps_m_varGrid.cpp:333-385). */
int sc_get_metric(sc_ctx* ctx, double* yq_h, double* zr_h);

/* out = A*in, A = getA (sbpOps.hpp:69) as used by KSPSolve and by
 * propagateWaves (strikeSlip_linearElastic_fd.cpp:459-460).
 * prec = SC_FP64 (double buffers) or SC_FP32 (float buffers). */
int sc_sbp_apply(sc_ctx* ctx, int prec, const void* in_d, void* out_d);
/* diag(A) (FP64). */
int sc_sbp_diag(sc_ctx* ctx, double* out_d);
/* SbpOps::setRhs(rhs, bcL, bcR, bcT, bcB) (sbpOps.hpp:82;
 * sbpOps_m_varGrid.cpp:1046-1083).  bcL/bcR: Nz values, bcT/bcB: Ny. */
int sc_sbp_set_rhs(sc_ctx* ctx, double* rhs_d, const double* bcL_d, const double* bcR_d,
                   const double* bcT_d, const double* bcB_d);
/* SbpOps::Dy, muxDy, Dz, muxDz, H, Hinv (sbpOps.hpp:85-94). */
int sc_sbp_dy(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_mux_dy(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_dz(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_mux_dz(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_h(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_hinv(sc_ctx* ctx, const double* in_d, double* out_d);
/* SbpOps::Dyxmu, Dzxmu: out = Dy (mu in), Dz (mu in) (sbpOps.hpp:87,90;
 * sbpOps_m_varGrid.cpp:1641-1660, 1704-1723; sbpOps_m_constGrid.cpp:1503+). */
int sc_sbp_dyxmu(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_dzxmu(sc_ctx* ctx, const double* in_d, double* out_d);
/* SbpOps::Hyinvxe0y, HyinvxeNy (sbpOps.hpp:95-96; sbpOps_m_varGrid.cpp
 * :1760-1801; sbpOps_m_constGrid.cpp:1622+): in_d holds Nz values (one
 * face), out_d the grid: Hy^-1 e0y in is in / (h11 dy) on the face j = 0 and
 * zero elsewhere (j = Ny-1 for eNy). */
int sc_sbp_hyinvx_e0y(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_hyinvx_eNy(sc_ctx* ctx, const double* in_d, double* out_d);
/* SbpOps::HyinvxE0y, HyinvxENy, HzinvxE0z, HzinvxENz (sbpOps.hpp:97-100;
 * sbpOps_m_varGrid.cpp:1804-1889): in_d and out_d are grid fields; the face
 * j = 0 (j = Ny-1, k = 0, k = Nz-1) of in scaled by the 1D inverse norm's
 * weight there, zero elsewhere. */
int sc_sbp_hyinvx_E0y(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_hyinvx_ENy(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_hzinvx_E0z(sc_ctx* ctx, const double* in_d, double* out_d);
int sc_sbp_hzinvx_ENz(sc_ctx* ctx, const double* in_d, double* out_d);

/* Measurement: mean duration (ms) of `reps` back-to-back launches timed with
 * HIP events on the context stream, and the nodes they write.  which:
 * 0 FP64 apply, 1 FP32 apply, 2 FP64 interior kernel, 3 FP32 interior kernel,
 * 4 / 5 the interior kernel with the plain (not XCD-grouped) tile order,
 * 6 FP64 frame kernel, 7 FP32 frame kernel, 8 the projected-guess kernel of
 * sc_solve on the context's current basis (needs a solve with reuse > 0;
 * `nodes` then returns the 8-byte words it moves per launch), 9 / 10 the
 * FP64 / FP32 wave step of sc_wave_step (laplace type yz), 11 / 12 / 13 one
 * application of the fast-diagonalisation preconditioner with FP64 / FP32 /
 * bf16 GEMMs on FP64 vectors (sc_precond_apply; `nodes` then returns the
 * GEMM flops of one application, 4 Ny Nz (Ny + Nz)).
 * No counterpart in the reference (its timers are MPI_Wtime accumulators,
 * e.g. linearElastic.cpp:474-477). */
int sc_bench_apply(sc_ctx* ctx, int which, int reps, double* ms_per_launch, double* nodes);

/* ---------------------------------------------------------------- solve */
#define SC_SOLVE_CG64 0  /* FP64 Jacobi-PCG                                */
#define SC_SOLVE_IR32 1  /* FP32 PCG inside FP64 iterative refinement       */

typedef struct sc_solve_opts {
  int mode;            /* SC_SOLVE_CG64 / SC_SOLVE_IR32                   */
  double rtol;         /* ||b - A u||_2 <= rtol * ||b||_2 (FP64 residual)  */
  double inner_rtol;   /* IR32: relative reduction per FP32 inner solve    */
  int max_iters;       /* cap on (inner) CG iterations in total           */
  int max_outer;       /* IR32: cap on refinement sweeps                  */
  int check_every;     /* iterations between host convergence checks      */
  int precond;         /* SC_PREC_*                                         */
  int reuse;           /* 0: start from u as given.  1..SC_REUSE_MAX: keep
                          the last `reuse` solutions of this context as an
                          A-orthogonal basis and start from the projection
                          of the new solution onto it (Fischer's method for
                          successive right-hand sides); the basis is reset
                          when the operator or its BC types change.        */
} sc_solve_opts;

#define SC_REUSE_MAX 8

/* Preconditioners.  JACOBI: 1/diag(A).  FDIAG*: exact inverse of the
 * separable part of A (the operator with a depth-only shear modulus and
 * without the reference's two-point mu average), applied as four dense
 * GEMMs in FP64, FP32 or bf16 (FP32 accumulation).  SC_SOLVE_CG64 takes any
 * of them (flexible CG); SC_SOLVE_IR32 takes JACOBI, FDIAG32 or FDIAG_BF16
 * for its FP32 inner solves (flexible PCG in FP32, FP64 refinement). */
#define SC_PREC_JACOBI 0
#define SC_PREC_FDIAG64 1
#define SC_PREC_FDIAG32 2
#define SC_PREC_FDIAG_BF16 3

typedef struct sc_solve_stats {
  int iters;           /* total CG iterations (FP64 or FP32)              */
  int outer;           /* refinement sweeps (IR32)                        */
  int converged;
  double relres;       /* final FP64 relative residual                    */
} sc_solve_stats;

/* LinearElastic::computeU -> KSPSolve(A, rhs, u) (linearElastic.cpp:465-489)
 * with u as the initial guess (KSPSetInitialGuessNonzero) and the solution. */
int sc_solve(sc_ctx* ctx, const double* rhs_d, double* u_d, const sc_solve_opts* opts,
             sc_solve_stats* stats);

/* PCApply of the solve's preconditioner (KSPGetPC / PCApply, the PC of the
 * reference's KSP set up in linearElastic.cpp:251-289): z = M^-1 r on the
 * device for precond SC_PREC_FDIAG64 / FDIAG32 / FDIAG_BF16 (the
 * fast-diagonalisation inverse of the separable operator, built on first use
 * and after the operator changes) or SC_PREC_JACOBI (z = r / diag(A)). */
int sc_precond_apply(sc_ctx* ctx, int precond, const double* r_d, double* z_d);

/* sc_solve without any device allocation: FP64 Jacobi-PCG (SC_SOLVE_CG64,
 * SC_PREC_JACOBI, reuse 0; other options give SC_ERR_ARG) from u as the
 * initial guess, the same kernels and results as sc_solve.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (the CG vectors, reduction partials and scalars)
 * is one static device array of 25.2 MB per GPU, held until the process
 * ends, which limits it to grids of at most 1048576 nodes (larger grids:
 * SC_ERR_ARG).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU race.  The context itself (made by sc_ctx_create) is
 * the caller's, as for sc_solve. */
int sc_solve_noalloc(sc_ctx* ctx, const double* rhs_d, double* u_d, const sc_solve_opts* opts,
                     sc_solve_stats* stats);

/* ---------------------------------------------------------------- fault */
#define SC_AGING_LAW 0
#define SC_SLIP_LAW 1

typedef struct sc_fault_qd_args {
  int N;                     /* fault nodes (Nz)                          */
  double v0, vL, f0, rootTol;
  int maxNumIts;
  int stateLaw;              /* SC_AGING_LAW / SC_SLIP_LAW               */
  int jac_mode;              /* 0: reference's Newton slope (fault.cpp:1051), 1: exact */
  const double *eta_d, *sN_d, *a_d, *b_d, *Dc_d, *locked_d, *prestress_d;
  const double *psi_d;
  double *tauQS_d;           /* in: sxy on the fault; out: + prestress    */
  double *V_d;               /* in: previous slip velocity (Newton start); out: V */
  double *dpsi_d, *tau_d, *strength_d;
  int* status_d;             /* device int: set non-zero on failure (may be NULL) */
} sc_fault_qd_args;

/* Fault_qd::d_dt minus the scatter (fault.cpp:861-918): tauQS += prestress,
 * BracketedNewton slip-velocity solve (fault.cpp:954-1059,
 * rootFinder.cpp:174-257), aging/slip state rate (fault.cpp:1984-2088),
 * tau = tauQS - eta*V, strength (fault.cpp:2246-2283). */
int sc_fault_qd_rates(sc_ctx* ctx, const sc_fault_qd_args* args);

typedef struct sc_fault_fd_args {
  int N;                     /* fault nodes (Nz)                          */
  double dt;                 /* time step                                 */
  double v0, vL, f0, rootTol, faultTypeScale;
  int maxNumIts;
  int stateLaw;              /* SC_AGING_LAW / SC_SLIP_LAW               */
  const double *a_d, *b_d, *Dc_d, *sNEff_d, *rho_d, *alphay_d, *tau0_d, *slip0_d;
  const double* locked_d;    /* may be NULL                               */
  const double* d2u_d;       /* D2u on the fault from sc_wave_step        */
  const double *psi_d, *psiPrev_d, *slip_d, *slipPrev_d;
  double* V_d;               /* in: previous |V| (Newton start); out: V   */
  double *psiNext_d, *slipNext_d, *strength_d;
  double* uFault_d;          /* out: fault displacement (may be NULL)     */
  int* status_d;             /* device int: set non-zero on failure (may be NULL) */
} sc_fault_fd_args;

/* Fault_fd::d_dt (fault.cpp:1436-1523): u, uPrev from slip, setPhi
 * (fault.cpp:1527-1575), slip-velocity root solve (ComputeVel_fd,
 * fault.cpp:1588-1685), fault displacement update, implicit state update
 * (ComputeAging_fd / ComputeSlipLaw_fd, fault.cpp:1697-1879), slip and
 * strength.  One thread per fault node, FP64. */
int sc_fault_fd_step(sc_ctx* ctx, const sc_fault_fd_args* args);

/* ------------------------------------------------------------ power law */
typedef struct sc_powerlaw_args {
  int N;
  double effViscCap;         /* viscosity cap (PowerLaw _effViscCap, default 1e30) */
  const void *sxy_d, *sxz_d; /* stresses (sxz may be NULL: 0)              */
  const void *T_d, *A_d, *B_d, *n_d;  /* temperature (K), A, Q/R, n           */
  const void *mu_d;
  const void *sat_d;         /* SAT term of the strain rate (may be NULL: 0) */
  void *effVisc_d;           /* out (may be NULL)                           */
  void *gxy_t_d, *gxz_t_d;   /* out: viscous strain rates (gxz may be NULL) */
} sc_powerlaw_args;

/* DislocationCreep::computeInvEffVisc + PowerLaw::computeViscosity +
 * computeViscStrainRates, pointwise (powerLaw.cpp:438-474, 1674-1743):
 * 1/eta = 1/cap + 1e3 A s^(n-1) exp(-B/T), s = |(sxy, sxz)|,
 * d/dt gxy = (mu*sat + sxy)/eta, d/dt gxz = sxz/eta.  prec SC_FP64: double
 * arrays; SC_FP32: float arrays (the law itself is evaluated in FP64). */
int sc_powerlaw_rates(sc_ctx* ctx, int prec, const sc_powerlaw_args* args);

/* ------------------------------------------------------------ wave step */
/* strikeSlip_linearElastic_fd::propagateWaves (strikeSlip_linearElastic_fd.cpp:
 * 445-516): D2u = Jinv*Hinv*(A u), uNext = (dt^2/rho*D2u + 2u +
 * (dt*ay-1)*uPrev)/(dt*ay+1), fused into the matrix-free apply.  prec selects
 * the arrays' type (SC_FP64: double*, SC_FP32: float*); d2uFault receives
 * D2u on the fault column (Nz values, same type).  ay must vanish off the
 * domain boundary (it does: strikeSlip_linearElastic_fd.cpp:645-665). */
int sc_wave_step(sc_ctx* ctx, int prec, double dt, const void* u_d, const void* uPrev_d,
                 void* uNext_d, const void* rho_d, const void* ay_d, void* d2uFault_d);

/* -------------------------------------------------------- FD integration */
typedef struct sc_fd sc_fd;

typedef struct sc_fd_desc {
  sc_grid_desc grid;          /* material SAT types: Neumann on all four sides
                                 (strikeSlip_linearElastic_fd.cpp:22)        */
  int outR, outT, outB;       /* outgoing-characteristics sides (a_y terms) */
  int faultL;                 /* a_y term on the fault side (symmFault /
                                 rigidFault / outgoing)                     */
  double faultTypeScale, vL, f0, v0, rootTol;
  int maxNumIts, stateLaw;
  int bulkPrec;               /* SC_FP64, or SC_FP32: body arrays and update in
                                 FP32, fault state and root solves in FP64  */
} sc_fd_desc;

/* StrikeSlip_LinearElastic_fd with its Fault_fd, one constant-dt step per
 * OdeSolver_WaveEq::integrate iteration (odeSolver_WaveEq.cpp:147-186).
 * Host arrays: mu, rho, cs (N, z-fast), y, z (N, may be NULL for a constant
 * grid), fault arrays a, b, Dc, sNEff, tau0, prestress, slip0 (Nz), locked
 * (Nz, may be NULL).  The fault density is rho on the fault column and
 * alpha_y = h11y*yq (computePenaltyVectors, :608-632). */
int sc_fd_create(const sc_fd_desc* desc, const double* mu, const double* rho, const double* cs,
                 const double* y, const double* z, const double* a, const double* b,
                 const double* Dc, const double* sNEff, const double* tau0,
                 const double* prestress, const double* slip0, const double* locked,
                 sc_fd** out);
int sc_fd_destroy(sc_fd* fd);
sc_ctx* sc_fd_ctx(sc_fd* fd);
/* state at the previous and current time levels (host arrays; u: N, the
 * rest Nz; V may be NULL: zero) */
int sc_fd_set_state(sc_fd* fd, const double* uPrev, const double* u, const double* psiPrev,
                    const double* psi, const double* slipPrev, const double* slip,
                    const double* V);
/* nSteps steps of size dt from time t0 (d_dt, then prev <- cur <- next) */
int sc_fd_steps(sc_fd* fd, double t0, double dt, int nSteps);
/* current level (host arrays, any may be NULL): u (N), psi, slip, V, tau,
 * strength (Nz); tau = sxy on the fault + prestress + tau0 (d_dt, :421-434) */
int sc_fd_get_state(sc_fd* fd, double* u, double* psi, double* slip, double* V, double* tau,
                    double* strength);

/* ------------------------------------------------------- QD integration */
typedef struct sc_qd sc_qd;

typedef struct sc_qd_desc {
  sc_grid_desc grid;          /* material SAT types: typically R=Dirichlet (remoteLoading),
                                 T=Neumann, L=Dirichlet (symmFault), B=Neumann */
  int remoteLoadingR;         /* bcR = vL*t/faultTypeScale + bcRShift      */
  double faultTypeScale, vL, f0, v0, rootTol;
  int maxNumIts, stateLaw, jac_mode;
  sc_solve_opts solve;
} sc_qd_desc;

typedef struct sc_qd_run {
  int integrator;             /* 0 RK32, 1 RK43 (odeSolver.cpp)           */
  int control;                /* 0 P, 1 PID                               */
  int normType;               /* 0 L2_absolute, 1 L2_relative             */
  int maxSteps;
  double initTime, maxTime, initDt, minDt, maxDt, tol;
  int nReplay;                /* >0: force the step sizes in replayDt       */
  const double* replayDt_h;
  /* With sc_solve_opts.reuse > 0, when a projected initial guess may be
   * accepted on the device and checked with the next error estimate:
   * 0 after a run of guesses that met the tolerance, 1 never (every solve
   * is checked on the host), 2 always. */
  int deferChecks;
  /* PID error history at the start, errA[0] and errA[1] of the reference's
   * OdeSolver (odeSolver.cpp:231-234, 610): 0 and 0 for a fresh start, the
   * checkpointed chkpt_currErr / chkpt_prevErr on a restart
   * (strikeSlip_linearElastic_qd.cpp:734-737). */
  double currErr, prevErr;
  /* 0: with deferred checks (and no replay, output or power law), RK32's
   * step control runs on the device and the host reads it back once per
   * batch of attempts; 1: the host controls every step (the reference's
   * loop as written).  Both take the same steps up to the last bit of pow. */
  int stepControl;
} sc_qd_run;

typedef struct sc_qd_counters {
  long long d_dt_calls, solves, cg_iters, outer_sweeps, rejected_steps;
  double solve_seconds, total_seconds;
  /* solves inside sc_qd_integrate whose projected guess was accepted on the
   * device, and the attempts repeated because such a guess missed */
  long long deferred_checks, missed_guesses;
  /* accepted steps whose step control ran on the device (sc_qd_run.stepControl) */
  long long device_steps;
  /* captures of the device loop's attempt graph (it is kept across
   * sc_qd_integrate calls and captured again only when its launches'
   * arguments change, e.g. after the solution basis grew) */
  long long graph_builds;
} sc_qd_counters;

/* StrikeSlip_LinearElastic_qd (strikeSlip_linearElastic_qd.cpp:7-78) with
 * Fault_qd and LinearElastic, over host arrays: mu/y/z (N), fault a, b, Dc,
 * sN, eta, locked, prestress, bcRShift (Nz; locked/prestress/bcRShift may be
 * NULL = zero). */
int sc_qd_create(const sc_qd_desc* desc, const double* mu_h, const double* y_h,
                 const double* z_h, const double* a_h, const double* b_h, const double* Dc_h,
                 const double* sN_h, const double* eta_h, const double* locked_h,
                 const double* prestress_h, const double* bcRShift_h, sc_qd** out);
/* Reads a reference .in deck (Domain/LinearElastic/Fault/mediator keys,
 * binary Vec overrides from inputDir) and builds the same mediator.  A deck
 * that names outputDir also sets the output (sc_qd_set_output with its
 * stride1D, stride2D, enableCheckpointing, interval; reference defaults 1, 1,
 * 0, 10000); when that directory holds a checkpoint the mediator starts from
 * it (sc_qd_load_checkpoint into the deck's run, no solveSS), as the
 * reference's constructor does. */
int sc_qd_create_from_file(const char* in_path, sc_qd** out);
int sc_qd_destroy(sc_qd* q);
sc_ctx* sc_qd_ctx(sc_qd* q);
/* Integrated state (psi, slip) and the persistent slip velocity (Nz each). */
int sc_qd_set_state(sc_qd* q, const double* psi_h, const double* slip_h, const double* V_h);
int sc_qd_get_state(sc_qd* q, double* psi_h, double* slip_h, double* V_h, double* tau_h);
int sc_qd_get_u(sc_qd* q, double* u_h);
/* One StrikeSlip_LinearElastic_qd::d_dt (strikeSlip_linearElastic_qd.cpp:756-798)
 * at time t on the current state: rates dpsi_h, dslip_h (= V). */
int sc_qd_ddt(sc_qd* q, double t, double* dpsi_h, double* dslip_h);
/* OdeSolver::integrate (odeSolver.cpp:505-627 RK32, :973-1175 RK43).
 * Returns accepted step sizes in dts_h[0..min(steps,dtsCap)). */
int sc_qd_integrate(sc_qd* q, const sc_qd_run* run, double* tEnd_h, int* steps_h,
                    double* dts_h, int dtsCap);
int sc_qd_get_counters(sc_qd* q, sc_qd_counters* c);
/* Measurement: mean duration (ms) of `reps` back-to-back launches of the
 * device loop's stage kernel (k_qd_stage) in its accepted-state form — the
 * rates at the current state and time t, which it leaves as d_dt would up
 * to rounding — timed with HIP events on the context stream, and the 8-byte
 * words it moves per launch.  Needs the linear-elastic mediator, a solution
 * basis (an integration with reuse > 0) and Nz <= 2048.  No counterpart in
 * the reference (its timers are MPI_Wtime accumulators). */
int sc_qd_bench_stage(sc_qd* q, double t, int reps, double* ms_per_launch, double* words);
/* Diagnostic: the squared residual of the projected guess at the current
 * state and time t, r = b - A x, x = sum c_m y_m, as the stage kernel's
 * residual test forms it (rr_bb[0] = r.r, rr_bb[1] = b.b: node by node on
 * the band rows where b is nonzero, and c^T G c with the off-band Gram
 * matrix G = <A y, A y> in double-double elsewhere), and
 * the same two sums by an explicit pass (x formed, b from the boundary data,
 * A x applied: rr_bb[2], rr_bb[3]).  The KSP convergence test on a guess
 * (||r|| < rtol ||b||) is what the stage kernel decides with these.  Same
 * requirements as sc_qd_bench_stage; no counterpart in the reference. */
int sc_qd_guess_residual(sc_qd* q, double t, double* rr_bb);
/* Measurement: sc_qd_integrate with the device loop's attempts launched one
 * by one (no graph) and HIP events around each stage launch on the context
 * stream; returns the mean duration (ms) of the plain stage kernels
 * (k_qd_stage, two per attempt) and of the attempt-end kernel over the
 * attempts that ran (*attempts; 0 when the run never entered the device
 * loop).  The steps taken are those of sc_qd_integrate.  No counterpart in
 * the reference. */
int sc_qd_integrate_timed(sc_qd* q, const sc_qd_run* run, double* tEnd_h, int* steps_h,
                          double* dts_h, int dtsCap, double* ms_stage, double* ms_end,
                          long long* attempts);
/* Time-integration settings read from the deck (timeIntegrator,
 * timeControlType, maxStepCount, init/maxTime, min/max/initDeltaT with
 * computeMinTimeStep, timeStepTol, normType), for sc_qd_create_from_file
 * drivers. */
int sc_qd_file_run(sc_qd* q, sc_qd_run* run_out);
/* StrikeSlip_LinearElastic_qd::solveSS (strikeSlip_linearElastic_qd.cpp:895-1006):
 * steady-state guess at vL (Fault::guessSS fault.cpp:633-709; psi recomputed
 * unless have_state_vals), Neumann-fault solve, shift u >= 0, bcRShift and
 * slip from the solution, back to the configured BC types.  psi_h in/out,
 * slip_h out (Nz); also sets the driver state (psi, slip, V = vL). */
int sc_qd_solve_ss(sc_qd* q, double* psi_h, int have_state_vals, double* slip_h);

/* Output series and checkpoints of the linear-elastic QD mediator
 * (StrikeSlip_LinearElastic_qd::timeMonitor, strikeSlip_linearElastic_qd.cpp
 * :417-463).  With an output set, sc_qd_integrate calls the monitor before
 * its first step (step 0, with the fields as they are) and after every
 * accepted step:
 *   stride1D > 0, step % stride1D == 0 or t == maxTime: one record appended
 *     to each PETSc-binary series outputDir+name: slip, slipVel, tauP, tauQSP,
 *     strength, psi, sNEff (Fault::writeStep fault.cpp:512-557), surfDisp,
 *     momBal_bcR/T/L/B and momBal_bcRShift (LinearElastic::writeStep1D
 *     linearElastic.cpp:623-652; as there, the bcRShift series holds bcRShift
 *     in its first record and bcR after it), and a line "%.15e" to
 *     med_time1D.txt and med_dt1D.txt (:466-495);
 *   stride2D > 0, likewise: momBal_u, momBal_sxy (N, z-fast;
 *     linearElastic.cpp:656-686), med_time2D.txt, med_dt2D.txt;
 *   enableCheckpointing > 0, step % interval == 0, step >= maxSteps or
 *     t >= maxTime: ckptNumber += 1 and the checkpoint files (ASCII
 *     ckptNumber, chkpt_currT, chkpt_stepCount, chkpt_prevErr, chkpt_currErr,
 *     chkpt_deltaT; PETSc-binary chkpt_slip, chkpt_slipVel, chkpt_psi,
 *     chkpt_sNEff, chkpt_sN, chkpt_tau, chkpt_tauQS, chkpt_prestress,
 *     chkpt_momBal_bcR, chkpt_momBal_bcRShift, chkpt_momBal_bcT,
 *     chkpt_momBal_bcL, chkpt_momBal_bcB; :527-557, fault.cpp:561-584,
 *     linearElastic.cpp:689-705), after which the integration stops (not at
 *     step 0, whose stop request the reference ignores).
 * Each record is of the rates at the accepted state, with any deferred guess
 * check settled first.  Series start truncated, or appended when
 * enableCheckpointing > 0 and outputDir+ckptNumber holds a number > 0
 * (Domain, domain.cpp:26-28).  o = NULL switches output off.
 * The power-law mediator (StrikeSlip_PowerLaw_qd::timeMonitor,
 * strikeSlip_powerLaw_qd.cpp:352-408) writes the same 1-D series and, as the
 * 2-D series, PowerLaw::writeStep2D's files (powerLaw.cpp:2312-2372):
 * momBal_u, momBal_sxy, momBal_sxz, momBal_sdev, momBal_gTxy, momBal_gTxz,
 * momBal_gxy, momBal_gxz, momBal_dgVxy, momBal_dgVxz, momBal_effVisc,
 * momBal_T, disl_invEffVisc, disl_dgVdev; like the reference it replaces the
 * strides by 1 below step 50 and by 100 from there on (a stride <= 0 still
 * switches a series off), and it writes no checkpoints (enableCheckpointing
 * must be 0). */
typedef struct sc_qd_output {
  const char* outputDir;      /* prefix of every file name                   */
  int stride1D, stride2D;     /* <= 0: that series is not written            */
  int enableCheckpointing;    /* > 0: checkpoints every `interval` steps     */
  int interval;               /* >= 1 with checkpoints                       */
} sc_qd_output;
int sc_qd_set_output(sc_qd* q, const sc_qd_output* o);
/* The restart half of the checkpoint (the mediator's constructor,
 * strikeSlip_linearElastic_qd.cpp:32-39, Fault::loadCheckpoint
 * fault.cpp:144-176, LinearElastic::loadCheckpoint linearElastic.cpp:396-420):
 * with an output set whose ckptNumber is > 0, loads psi, slip = bcL *
 * faultTypeScale (initiateIntegrand takes slip from bcL, :391-395), sN(Eff),
 * prestress, tauQS, tau, bcRShift and bcR/T/L/B, and sets run->initTime,
 * initDt, currErr, prevErr from the ASCII files.  The saved slip velocity is
 * taken as the first root-solve start (the reference restarts that from its
 * set-up value; either start gives the root to rootTol).  The step count
 * restarts at 0: the reference's step-0 monitor call overwrites the loaded
 * count before it reaches the solver (:732-737).  *ckptNumber gets the number
 * (0: nothing loaded). */
int sc_qd_load_checkpoint(sc_qd* q, sc_qd_run* run, int* ckptNumber);

/* Output series of the fully dynamic mediator
 * (strikeSlip_linearElastic_fd::timeMonitor, strikeSlip_linearElastic_fd.cpp
 * :218-251, writeStep1D/2D :254-305): the same files as the QD mediator's
 * series (med_time1D.txt, med_dt1D.txt, surfDisp, momBal_bcR/T/L/B,
 * momBal_bcRShift, slip, slipVel, tauP, tauQSP, strength, psi, sNEff;
 * med_time2D.txt, med_dt2D.txt, momBal_u, momBal_sxy).  As in the reference,
 * surfDisp keeps the surface of the state given to sc_fd_set_state (the FD
 * mediator never updates it), the boundary vectors stay zero, and tauQSP
 * equals tauP.  enableCheckpointing must be 0 (the reference's FD mediator
 * writes no checkpoints); o = NULL switches output off. */
int sc_fd_set_output(sc_fd* fd, const sc_qd_output* o);
/* OdeSolver_WaveEq::integrate (odeSolver_WaveEq.cpp:147-186): the output of
 * the initial state (step 0), then steps of deltaT (deltaT = 0: (maxTime -
 * initTime) / maxSteps) while stepCount < maxSteps and t < maxTime, the time
 * clamped to maxTime, the output after each step; *tEnd, *steps at the end.
 * Steps between two written steps run as one batch of launches. */
int sc_fd_integrate(sc_fd* fd, double initTime, double deltaT, int maxSteps, double maxTime,
                    double* tEnd, int* steps);

/* ------------------------------------------------ QD <-> FD switching */
/* strikeSlip_linearElastic_qd_fd (strikeSlip_linearElastic_qd_fd.cpp): QD
 * phases (the QD mediator's d_dt, RK32/RK43) and FD phases (propagateWaves +
 * Fault_fd::d_dt at the constant deltaT_fd) in turn, switched on
 * R = max_k eta V / tauQS (checkSwitchRegime, :590-640): a QD phase may end
 * once R has been below limit_qd and then exceeds trigger_qd2fd; an FD phase
 * once R has been above limit_fd and then falls below trigger_fd2qd.
 * Switching to FD takes one QD step of deltaT_fd (prepare_qd2fd, :807-889);
 * the FD phase then integrates the displacement change du from the QD
 * displacement u0 at that step's start, starting from rest (uPrev = u, as
 * OdeSolver_WaveEq::setInitialConds(var) sets it), with slip0 = slip and
 * tau0 = strength + prestress; switching back hands psi, slip and V to the
 * QD phase (prepare_fd2qd, :747-805).  Each QD phase starts at dt =
 * deltaT_fd; each phase may take maxStepCount more steps
 * (setInitialStepCount).  Two reference behaviours are kept by default:
 *   fdStop = 0: OdeSolver_WaveEq::integrate never reads the stop request of
 *     the monitor (odeSolver_WaveEq.cpp:166-182), so an FD phase runs to
 *     maxTime or maxStepCount; fdStop = 1 ends it on the switching criterion
 *     (full cycles) and starts each QD phase's step-size controller afresh
 *     (P control for its first four steps, as in a first phase): the
 *     reference's new RK object would keep the global step count with an
 *     empty error history, for which the PID ratio is NaN and a rejected
 *     step grows fivefold (odeSolver.cpp:415-433);
 *   ayCompat = 1: computePenaltyVectors of this mediator (:486-500) writes
 *     a_y through an index that advances only on the fault row, so a_y holds
 *     the top/bottom terms of the fault row's nodes, the last node's value at
 *     index Nz, and zeros elsewhere (and no fault-side term); ayCompat = 0
 *     uses the FD mediator's a_y (strikeSlip_linearElastic_fd.cpp:640-665).
 * deltaT_fd: computeTimeStep (:356-443) from deltaT_fd and CFL (<= 0:
 * unset; with CFL alone dt = CFL g_cfl^2 min(d/cs)). */
typedef struct sc_qdfd sc_qdfd;
typedef struct sc_qdfd_desc {
  sc_qd_desc qd;              /* grid with the QD-phase material SAT types, friction, solve */
  int fdBcR, fdBcT, fdBcL, fdBcB;  /* material SAT types of the FD phase (reference: Neumann) */
  int outR, outT, outB, faultL;    /* a_y sides (momBal_bc*_fd outGoingCharacteristics; fault) */
  int ayCompat;               /* 1: the reference's a_y of this mediator (see above) */
  int bulkPrec;               /* SC_FP64 / SC_FP32 body in FD phases     */
  double deltaT_fd, CFL;      /* <= 0: unset                              */
  double trigger_qd2fd, trigger_fd2qd, limit_qd, limit_fd;  /* reference defaults 1e-3, 1e-3, 10 vL, 0.1 */
  int maxNumCycles;
  int fdStop;                 /* 0 as the reference, 1: FD phases end on trigger_fd2qd */
} sc_qdfd_desc;
typedef struct sc_qdfd_run {
  int integrator, control, normType;  /* QD phases, as sc_qd_run */
  int maxStepCount;           /* the mediator's maxStepCount             */
  double initTime, maxTime, minDt, maxDt, tol;
  int deferChecks;
} sc_qdfd_run;
/* one phase of a run: regime 0 QD, 1 FD, 2 the QD step of deltaT_fd that
 * starts an FD phase; R the switching ratio at its last step */
typedef struct sc_qdfd_phase {
  int regime, stepStart, stepEnd;
  double tStart, tEnd, R;
} sc_qdfd_phase;
/* host arrays: mu, rho, cs, y, z (N; y, z may be NULL for a constant grid),
 * a, b, Dc, sN, eta (radiation damping), locked, prestress, bcRShift (Nz;
 * the last three may be NULL = zero) */
int sc_qdfd_create(const sc_qdfd_desc* desc, const double* mu_h, const double* rho_h,
                   const double* cs_h, const double* y_h, const double* z_h, const double* a_h,
                   const double* b_h, const double* Dc_h, const double* sN_h, const double* eta_h,
                   const double* locked_h, const double* prestress_h, const double* bcRShift_h,
                   sc_qdfd** out);
int sc_qdfd_destroy(sc_qdfd* h);
int sc_qdfd_time_step(sc_qdfd* h, double* deltaT_fd);
/* the initial state (a QD state): psi, slip, V (Nz; V may be NULL = 0) */
int sc_qdfd_set_state(sc_qdfd* h, const double* psi_h, const double* slip_h, const double* V_h);
/* the current state of the current regime: psi, slip, V (Nz), the total
 * displacement u (N) and the regime (0 QD, 1 FD); any may be NULL */
int sc_qdfd_get_state(sc_qdfd* h, double* psi_h, double* slip_h, double* V_h, double* u_h, int* regime);
/* strikeSlip_linearElastic_qd_fd::integrate (:511-580) from run->initTime;
 * the phases go to log[0..logCap) and their number to *nPhases */
int sc_qdfd_integrate(sc_qdfd* h, const sc_qdfd_run* run, double* tEnd, int* steps, int* cycles,
                      sc_qdfd_phase* log, int logCap, int* nPhases);

/* Power-law viscoelastic QD: StrikeSlip_PowerLaw_qd (strikeSlip_powerLaw_qd.cpp
 * :633-686, 787-823) with PowerLaw in its dislocation-creep form.  Same
 * arguments as sc_qd_create plus T, A, B (= Q/R), n (N each, z-fast) and the
 * viscosity cap; the integrated variables are psi, slip (Nz) and the viscous
 * strains gVxy, gVxz (N, zero initially).  errMask picks the fields of the
 * RK error estimate (timeIntInds): 1 psi, 2 slip, 4 gVxy, 8 gVxz. */
int sc_qd_create_powerlaw(const sc_qd_desc* desc, const double* mu_h, const double* y_h,
                          const double* z_h, const double* a_h, const double* b_h,
                          const double* Dc_h, const double* sN_h, const double* eta_h,
                          const double* locked_h, const double* prestress_h,
                          const double* bcRShift_h, const double* T_h, const double* A_h,
                          const double* B_h, const double* n_h, double effViscCap, int errMask,
                          sc_qd** out);
/* viscous strains in/out (host arrays, N each; any may be NULL) */
int sc_qd_set_visc(sc_qd* q, const double* gVxy_h, const double* gVxz_h);
/* state of the last d_dt: strains, their rates, effective viscosity, sxy */
int sc_qd_get_visc(sc_qd* q, double* gVxy_h, double* gVxz_h, double* dgVxy_h, double* dgVxz_h,
                   double* effVisc_h, double* sxy_h);

/* ------------------------------------- discrete adjoint of the QD step --
 * No counterpart in the reference: the gradient of a surface-displacement
 * misfit with respect to the fault's friction parameters a, b, Dc (per fault
 * node), through nSteps RK32 steps of the linear-elastic, aging-law QD
 * mediator with the step sizes held fixed (as sc_qd_run.replayDt).
 *   J = 1/2 sum_s (u(j_s, k = 0, t_N) - d_s)^2,
 * u the off-fault displacement of the last d_dt (at the final state and
 * time).  The forward starts from the mediator's current psi, slip, V and
 * both calls leave them there.  The reverse pass recomputes the forward from
 * checkpoints taken every ckptEvery steps; the elasticity solves of the
 * adjoint use the same symmetric operator apply (options adjSolve, their own
 * projection basis), the root solve is differentiated through the implicit
 * function theorem. */
typedef struct sc_adj_desc {
  int nSteps;
  const double* dts_h;        /* nSteps step sizes                          */
  double t0;                  /* time of the initial state                  */
  int ckptEvery;              /* steps between checkpoints (>= 1)           */
  int nStations;
  const int* stationJ_h;      /* station y-indices j_s (node j_s*Nz, z = 0) */
  const double* data_h;       /* d_s (NULL: zero)                           */
  sc_solve_opts adjSolve;
  /* How the reverse pass applies (1/fts) R_L^T A^-1 S^T to a stage's
   * fault-traction cotangent.  1: as the Nz x Nz boundary operator formed
   * once by Nz adjoint solves of unit tractions (kept by the mediator and
   * reused by later gradients until mu, the BC types or the adjoint solve
   * options change); 2: one adjoint solve per stage; 0: the operator when
   * it is already formed or the run has at least Nz / 3 steps. */
  int greenMode;
} sc_adj_desc;

typedef struct sc_adj_stats {
  double J;
  double forward_seconds;     /* checkpointed forward                        */
  double reverse_seconds;     /* recomputation + adjoint steps (+ forming the
                                 boundary operator when this call formed it) */
  long long fwd_solves, adj_solves, adj_cg_iters, recomputed_steps;
  long long green_solves;     /* adjoint solves that formed the boundary
                                 operator in this call (0: reused or unused)  */
  double green_seconds;       /* their wall clock, inside reverse_seconds    */
} sc_adj_stats;

/* Replace the friction parameters a, b, Dc (host arrays, Nz; NULL keeps one). */
int sc_qd_set_friction(sc_qd* q, const double* a_h, const double* b_h, const double* Dc_h);
/* Forward only: J and the station displacements u(j_s, 0, t_N) (may be NULL). */
int sc_qd_forward_misfit(sc_qd* q, const sc_adj_desc* d, double* J, double* uStations_h);
/* J and dJ/da, dJ/db, dJ/dDc per fault node (host arrays, Nz each). */
int sc_qd_adjoint(sc_qd* q, const sc_adj_desc* d, double* ga_h, double* gb_h, double* gDc_h,
                  sc_adj_stats* stats);

/* --------------------------------------------- host-only input utilities */
/* Value of `key` in a .in deck, cut at its first space like every
 * loadSettings (e.g. domain.cpp:151-165); the last occurrence wins. */
int sc_in_get(const char* in_path, const char* key, char* buf, int buflen);
/* Vector value "[v1 v2 ...]" of `key` (loadVectorFromInputFile,
 * genFuncs.cpp:413-440; a one-element list yields 0 values). */
int sc_in_get_vector(const char* in_path, const char* key, double* vals, int cap, int* n);
/* PETSc binary Vec files (big-endian, classid 1211214; genFuncs.cpp:74-94,
 * 380-410, 1024-1042): number of appended records, read record `rec`,
 * write or append one record. */
int sc_vec_count(const char* path, int* nrec);
int sc_vec_read(const char* path, int rec, double* vals, int cap, int* n);
int sc_vec_write(const char* path, const double* vals, int n, int append);

#ifdef __cplusplus
}
#endif

#endif /* SCYCLE_HIP_H */
