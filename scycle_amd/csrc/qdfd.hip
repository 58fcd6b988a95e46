// QD <-> FD switching mediator (strikeSlip_linearElastic_qd_fd.cpp): the QD
// driver (qd.cpp) and the FD driver (fd_internal.hpp), each with its own
// operator context (the two phases use different material SAT types), and
// the state handed between them on the device at each switch.  The phase
// control flow, the switching test and the hand-over follow integrate()
// (:511-580), checkSwitchRegime (:585-640), prepare_fd2qd (:747-805),
// prepare_qd2fd (:807-889), integrate_singleQDTimeStep (:891-957),
// integrate_qd (:1436-1509), integrate_fd (:1511-1546) and the FD-phase
// d_dt (:1676-1759); the time step and a_y follow computeTimeStep (:356-443)
// and computePenaltyVectors (:446-509).
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "fd_internal.hpp"
#include "qd_internal.hpp"
#include "sbp_dev.hpp"

namespace sc {

// du = a - b into both FD time levels (the FD phase starts from rest:
// OdeSolver_WaveEq::setInitialConds(var) copies var into varPrev)
template <typename T>
__global__ __launch_bounds__(256) void k_start_levels(const double* __restrict__ a, const double* __restrict__ b,
                                                      T* __restrict__ cur, T* __restrict__ prev, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = (T)(a[i] + -1.0 * b[i]);
  cur[i] = d;
  prev[i] = d;
}

// u = du + u0 (the total displacement)
template <typename T>
__global__ __launch_bounds__(256) void k_total_u(const T* __restrict__ du, const double* __restrict__ u0,
                                                 double* __restrict__ u, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u[i] = (double)du[i] + 1.0 * u0[i];
}

// tau0 = strength + prestress (prepare_qd2fd: VecCopy, then VecAXPY 1.0)
__global__ __launch_bounds__(256) void k_tau0(const double* __restrict__ strength,
                                              const double* __restrict__ prestress, double* __restrict__ tau0,
                                              int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) tau0[k] = strength[k] + 1.0 * prestress[k];
}

// The FD phase's switching ratio: tauQS = mu qy Dy (du + u0) on the fault +
// prestress (:1696-1708), with the u0 part sxy0 formed once per phase, then
// max_k eta V / tauQS.  One block.
constexpr int kFdRatioThreads = 1024;
template <typename T>
__global__ __launch_bounds__(kFdRatioThreads) void k_fd_ratio(DevOp<double> p, const T* __restrict__ du,
                                                              const double* __restrict__ sxy0,
                                                              const double* __restrict__ prestress,
                                                              const double* __restrict__ eta,
                                                              const double* __restrict__ V, double* out) {
  __shared__ double sh[kFdRatioThreads];
  const int Nz = p.Nz;
  double m = -INFINITY;
  for (int k = threadIdx.x; k < Nz; k += kFdRatioThreads) {
    double s = 0.0;
#pragma unroll
    for (int q = 4; q < kBand; ++q) s += p.y.D1[q] * (double)du[(size_t)min(q - 4, p.Ny - 1) * Nz + k];
    const double tq = p.mu[k] * (p.y.imet[0] * s) + sxy0[k] + 1.0 * prestress[k];
    m = fmax(m, (eta[k] * V[k]) / tq);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kFdRatioThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

static unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

struct QdFdDriver;

struct QdPhaseHook : QdStepHook {
  QdFdDriver* d = nullptr;
  int after_step(QdDriver& q, double t, double dt, int step, int* stop) override;
};

struct QdFdDriver {
  sc_qdfd_desc desc{};
  sc_qd* qh = nullptr;
  sc_fd* fh = nullptr;
  QdDriver* q = nullptr;
  FdDriver* f = nullptr;
  int Ny = 0, Nz = 0, N = 0;
  double dtFd = 0;
  double trig_qd2fd = 1e-3, trig_fd2qd = 1e-3, limit_qd = 1e-8, limit_fd = 1e-1;
  double *u0 = nullptr, *uStart = nullptr, *sxy0 = nullptr, *ratio = nullptr;
  double* hostR = nullptr;
  bool inDynamic = false, allowed = false, started = false;
  double currT = 0, lastR = 0;
  int stepCount = 0, cycleCount = 0;
  QdPhaseHook hook;
  std::vector<sc_qdfd_phase> log;

  ~QdFdDriver() {
    if (q) (void)hipStreamSynchronize(q->ctx->c.stream);
    if (f) (void)hipStreamSynchronize(f->ctx->c.stream);
    for (double* p : {u0, uStart, sxy0, ratio})
      if (p) (void)hipFree(p);
    if (hostR) (void)hipHostFree(hostR);
    if (qh) sc_qd_destroy(qh);
    if (fh) sc_fd_destroy(fh);
  }

  // checkSwitchRegime (:585-640) on the ratio R
  bool check_switch(double R) {
    bool must = false;
    if (inDynamic && !allowed && R > limit_fd) allowed = true;
    if (!inDynamic && !allowed && R < limit_qd) allowed = true;
    if (inDynamic && allowed && R < trig_fd2qd) must = true;
    if (!inDynamic && allowed && R > trig_qd2fd) must = true;
    return must;
  }

  int init(const sc_qdfd_desc& d, const double* mu, const double* rho, const double* cs, const double* y,
           const double* z, const double* a, const double* b, const double* Dc, const double* sN,
           const double* eta, const double* locked, const double* prestress, const double* bcRShift);
  int sync_both() {
    SC_HIP(hipStreamSynchronize(q->ctx->c.stream));
    SC_HIP(hipStreamSynchronize(f->ctx->c.stream));
    return SC_OK;
  }
  int qd_phase(const sc_qdfd_run& run, sc_qd_run r, int regime);
  int prepare_fd2qd();
  int prepare_qd2fd(const sc_qdfd_run& run);
  int fd_phase(const sc_qdfd_run& run);
  int fd_ratio(double* R);
  int integrate(const sc_qdfd_run& run);
};

int QdPhaseHook::after_step(QdDriver& q, double t, double dt, int step, int* stop) {
  (void)dt;
  (void)step;
  double R = 0;
  int e = q.max_ratio(t, &R);
  if (e) return e;
  d->lastR = R;
  if (d->check_switch(R)) *stop = 1;
  return SC_OK;
}

int QdFdDriver::init(const sc_qdfd_desc& d, const double* mu, const double* rho, const double* cs,
                     const double* y, const double* z, const double* a, const double* b, const double* Dc,
                     const double* sN, const double* eta, const double* locked, const double* prestress,
                     const double* bcRShift) {
  desc = d;
  const sc_grid_desc& g = d.qd.grid;
  Ny = g.Ny;
  Nz = g.Nz;
  N = Ny * Nz;
  if (g.laplace != 0) return fail(SC_ERR_ARG, "sc_qdfd_create: laplace type must be yz");
  if (d.maxNumCycles < 1) return fail(SC_ERR_ARG, "sc_qdfd_create: maxNumCycles must be >= 1");
  int e = sc_qd_create(&d.qd, mu, y, z, a, b, Dc, sN, eta, locked, prestress, bcRShift, &qh);
  if (e) return e;
  q = qh->q.get();
  // the FD phase: same fault, FD material SAT types; tau0 and slip0 are set
  // at each switch
  sc_fd_desc fd{};
  fd.grid = g;
  fd.grid.bcR = d.fdBcR;
  fd.grid.bcT = d.fdBcT;
  fd.grid.bcL = d.fdBcL;
  fd.grid.bcB = d.fdBcB;
  fd.outR = d.outR;
  fd.outT = d.outT;
  fd.outB = d.outB;
  fd.faultL = d.faultL;
  fd.faultTypeScale = d.qd.faultTypeScale;
  fd.vL = d.qd.vL;
  fd.f0 = d.qd.f0;
  fd.v0 = d.qd.v0;
  fd.rootTol = d.qd.rootTol;
  fd.maxNumIts = d.qd.maxNumIts;
  fd.stateLaw = d.qd.stateLaw;
  fd.bulkPrec = d.bulkPrec;
  std::vector<double> zeros((size_t)Nz, 0.0);
  e = sc_fd_create(&fd, mu, rho, cs, y, z, a, b, Dc, sN, zeros.data(), prestress ? prestress : zeros.data(),
                   zeros.data(), locked, &fh);
  if (e) return e;
  f = fh->d.get();
  const Ctx& c = q->ctx->c;
  if (d.ayCompat) {
    // computePenaltyVectors (:486-500) as it runs: the write index Jj advances
    // only on the fault row (the fault-side term's if-statement has Jj++ as
    // its body)
    const double h11y = c.ty.h11, h11z = c.tz.h11;
    std::vector<double> ay((size_t)N, 0.0);
    int Jj = 0;
    for (int I = 0; I < N; ++I) {
      ay[Jj] = 0;
      if ((I / Nz == Ny - 1) && d.outR) ay[Jj] += 0.5 / h11y;
      if ((I % Nz == 0) && d.outT) ay[Jj] += 0.5 / h11z;
      if (((I + 1) % Nz == 0) && d.outB) ay[Jj] += 0.5 / h11z;
      if (I / Nz == 0) Jj++;  // the fault side is symmFault or rigidFault
    }
    for (int I = 0; I < N; ++I) ay[I] = ay[I] * cs[I];
    if ((e = f->set_ay(ay.data()))) return e;
  }
  // computeTimeStep (:356-443)
  const double gcfl = g.order == 4 ? 0.7071 / std::sqrt(1.4498) : 0.7071;
  double minY = INFINITY, minZ = INFINITY;
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k) {
      const double dy = g.var_grid ? c.yq[j] / (Ny - 1) : g.Ly / (Ny - 1.0);
      const double dz = g.var_grid ? c.zr[k] / (Nz - 1) : g.Lz / (Nz - 1.0);
      const double csI = cs[(size_t)j * Nz + k];
      minY = std::min(minY, dy / csI);
      minZ = std::min(minZ, dz / csI);
    }
  const double maxDt = gcfl * std::min(std::fabs(minY), std::fabs(minZ));
  const double cflDt = d.CFL * gcfl * maxDt;
  const double req = d.deltaT_fd;
  dtFd = maxDt;
  if (req <= 0. && cflDt <= 0.) dtFd = maxDt;
  else if (req > 0. && cflDt <= 0.) dtFd = req;
  else if (req <= 0. && cflDt > 0.) dtFd = cflDt;
  else dtFd = req;
  // checkInput (:282-293)
  trig_qd2fd = d.trigger_qd2fd;
  trig_fd2qd = d.trigger_fd2qd;
  limit_qd = d.limit_qd;
  limit_fd = d.limit_fd;
  if (limit_fd < trig_qd2fd) limit_fd = 10 * trig_qd2fd;
  if (limit_qd > trig_fd2qd) limit_qd = trig_qd2fd / 10.0;
  for (double** p : {&u0, &uStart})
    if (hipMalloc(p, (size_t)N * sizeof(double)) != hipSuccess) return fail(SC_ERR_HIP, "hipMalloc failed");
  if (hipMalloc(&sxy0, (size_t)Nz * sizeof(double)) != hipSuccess ||
      hipMalloc(&ratio, sizeof(double)) != hipSuccess ||
      hipHostMalloc(&hostR, sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(SC_ERR_HIP, "allocation failed");
  SC_HIP(hipMemset(u0, 0, (size_t)N * sizeof(double)));
  hook.d = this;
  return SC_OK;
}

// integrate_qd (:1436-1509) or integrate_singleQDTimeStep (:891-957): a QD
// RK run from the current time and step count with the switching monitor
int QdFdDriver::qd_phase(const sc_qdfd_run& run, sc_qd_run r, int regime) {
  r.integrator = run.integrator;
  r.control = run.control;
  r.normType = run.normType;
  r.tol = run.tol;
  r.deferChecks = run.deferChecks;
  r.initTime = currT;
  r.initDt = dtFd;
  q->stepBase = stepCount;
  q->pidFresh = desc.fdStop != 0;
  q->hook = &hook;
  q->hits = 0;  // the basis of earlier solutions may no longer fit: check again
  double tEnd = currT;
  int steps = stepCount;
  const int e = q->integrate(r, &tEnd, &steps, nullptr, 0);
  q->hook = nullptr;
  q->stepBase = 0;
  q->pidFresh = false;
  if (e) return e;
  log.push_back(sc_qdfd_phase{regime, stepCount, steps, currT, tEnd, lastR});
  currT = tEnd;
  stepCount = steps;
  return SC_OK;
}

// prepare_fd2qd (:747-805): psi, slip and V of the FD phase start the QD
// phase; the total displacement warm-starts its first solve
int QdFdDriver::prepare_fd2qd() {
  if (!started) {  // first cycle: the FD fault was set from the QD state
    started = true;
    return SC_OK;
  }
  int e = sync_both();
  if (e) return e;
  hipStream_t s = q->ctx->c.stream;
  const size_t bz = (size_t)Nz * sizeof(double);
  SC_HIP(hipMemcpyAsync(q->vars["psi"].var, f->psi[f->cur], bz, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(q->vars["slip"].var, f->slip[f->cur], bz, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(q->V_d, f->V, bz, hipMemcpyDeviceToDevice, s));
  if (f->f32)
    hipLaunchKernelGGL(k_total_u<float>, dim3(nblk(N)), dim3(256), 0, s, (const float*)f->u[f->cur], u0, q->u_d,
                       (size_t)N);
  else
    hipLaunchKernelGGL(k_total_u<double>, dim3(nblk(N)), dim3(256), 0, s, (const double*)f->u[f->cur], u0,
                       q->u_d, (size_t)N);
  SC_HIP(hipGetLastError());
  q->uLazy = false;  // u_d now holds the total displacement
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

// prepare_qd2fd (:807-889)
int QdFdDriver::prepare_qd2fd(const sc_qdfd_run& run) {
  hipStream_t sq = q->ctx->c.stream;
  // u at time n-1, then one QD step of deltaT_fd to time n (with the
  // switching monitor in the QD regime, as there)
  if (int e0 = q->materialize_u()) return e0;
  SC_HIP(hipMemcpyAsync(uStart, q->u_d, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, sq));
  inDynamic = false;
  sc_qd_run r{};
  r.maxSteps = 1;
  r.maxTime = currT + dtFd;
  r.minDt = dtFd;
  r.maxDt = dtFd;
  int e = qd_phase(run, r, 2);
  inDynamic = true;
  if (e) return e;
  if ((e = q->settle(currT))) return e;
  if ((e = sync_both())) return e;
  hipStream_t s = f->ctx->c.stream;
  const size_t bz = (size_t)Nz * sizeof(double);
  f->prev = 0;
  f->cur = 1;
  f->next = 2;
  // body: du = u_n - u_{n-1} on both levels, u0 = u_{n-1}
  if (f->f32)
    hipLaunchKernelGGL(k_start_levels<float>, dim3(nblk(N)), dim3(256), 0, s, (const double*)q->u_d,
                       (const double*)uStart, (float*)f->u[f->cur], (float*)f->u[f->prev], (size_t)N);
  else
    hipLaunchKernelGGL(k_start_levels<double>, dim3(nblk(N)), dim3(256), 0, s, (const double*)q->u_d,
                       (const double*)uStart, (double*)f->u[f->cur], (double*)f->u[f->prev], (size_t)N);
  SC_HIP(hipGetLastError());
  SC_HIP(hipMemcpyAsync(u0, uStart, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
  // fault: psi, slip at both levels, V, slip0 = slip, strength, tau0 = strength + prestress
  for (int lv : {f->cur, f->prev}) {
    SC_HIP(hipMemcpyAsync(f->psi[lv], q->vars["psi"].var, bz, hipMemcpyDeviceToDevice, s));
    SC_HIP(hipMemcpyAsync(f->slip[lv], q->vars["slip"].var, bz, hipMemcpyDeviceToDevice, s));
  }
  SC_HIP(hipMemcpyAsync(f->V, q->V_d, bz, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(f->slip0, q->vars["slip"].var, bz, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(f->strength, q->strength_d, bz, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_tau0, dim3(nblk(Nz)), dim3(256), 0, s, (const double*)q->strength_d,
                     (const double*)q->prestress_d, f->tau0, Nz);
  SC_HIP(hipGetLastError());
  // the u0 part of the fault stress, for the FD phase's switching ratio
  if ((e = launch_fault_sxy(f->ctx->c, u0, sxy0, s))) return e;
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

int QdFdDriver::fd_ratio(double* R) {
  hipStream_t s = f->ctx->c.stream;
  const Ctx& c = f->ctx->c;
  if (f->f32)
    hipLaunchKernelGGL(k_fd_ratio<float>, dim3(1), dim3(kFdRatioThreads), 0, s, c.op64,
                       (const float*)f->u[f->cur], (const double*)sxy0, (const double*)f->prestress,
                       (const double*)q->eta_d, (const double*)f->V, ratio);
  else
    hipLaunchKernelGGL(k_fd_ratio<double>, dim3(1), dim3(kFdRatioThreads), 0, s, c.op64,
                       (const double*)f->u[f->cur], (const double*)sxy0, (const double*)f->prestress,
                       (const double*)q->eta_d, (const double*)f->V, ratio);
  SC_HIP(hipGetLastError());
  SC_HIP(hipMemcpyAsync(hostR, ratio, sizeof(double), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  *R = *hostR;
  return SC_OK;
}

// integrate_fd (:1511-1546) with OdeSolver_WaveEq::integrate
// (odeSolver_WaveEq.cpp:147-186): constant deltaT_fd from the current time,
// at most maxStepCount more steps; the monitor's switching test is honoured
// only with fdStop (see the header)
int QdFdDriver::fd_phase(const sc_qdfd_run& run) {
  const double finalT = run.maxTime;
  const int maxNum = stepCount + run.maxStepCount;
  const int s0 = stepCount;
  const double t0 = currT;
  if (finalT == currT) {
    log.push_back(sc_qdfd_phase{1, s0, stepCount, t0, currT, lastR});
    return SC_OK;
  }
  int e;
  if (!desc.fdStop) {
    // the steps the loop takes, then all of them in one go
    int n = 0;
    double t = currT;
    while (stepCount + n < maxNum && t < finalT) {
      t = t + dtFd;
      if (t > finalT) t = finalT;
      ++n;
    }
    if ((e = f->steps(currT, dtFd, n))) return e;
    currT = t;
    stepCount += n;
    if ((e = fd_ratio(&lastR))) return e;
  } else {
    while (stepCount < maxNum && currT < finalT) {
      currT = currT + dtFd;
      if (currT > finalT) currT = finalT;
      stepCount++;
      if ((e = f->steps(currT, dtFd, 1))) return e;
      if ((e = fd_ratio(&lastR))) return e;
      if (check_switch(lastR)) break;
    }
  }
  log.push_back(sc_qdfd_phase{1, s0, stepCount, t0, currT, lastR});
  return SC_OK;
}

// integrate (:511-580)
int QdFdDriver::integrate(const sc_qdfd_run& run) {
  log.clear();
  currT = run.initTime;
  stepCount = 0;
  cycleCount = 0;
  started = false;
  const int maxStep = run.maxStepCount;
  auto qd = [&]() -> int {
    allowed = false;
    inDynamic = false;
    int e = prepare_fd2qd();
    if (e) return e;
    sc_qd_run r{};
    r.maxSteps = run.maxStepCount;
    r.maxTime = run.maxTime;
    r.minDt = run.minDt;
    r.maxDt = run.maxDt;
    return qd_phase(run, r, 0);
  };
  auto fd = [&]() -> int {
    allowed = false;
    inDynamic = true;
    int e = prepare_qd2fd(run);
    if (e) return e;
    return fd_phase(run);
  };
  int e = qd();
  if (e) return e;
  if (currT >= run.maxTime || stepCount >= maxStep) return SC_OK;
  if ((e = fd())) return e;
  if (currT >= run.maxTime || stepCount >= maxStep || desc.maxNumCycles <= 1) return SC_OK;
  cycleCount++;
  while (cycleCount < desc.maxNumCycles && stepCount <= maxStep && currT <= run.maxTime) {
    if ((e = qd())) return e;
    if ((e = fd())) return e;
    cycleCount++;
  }
  return SC_OK;
}

}  // namespace sc

struct sc_qdfd {
  std::unique_ptr<sc::QdFdDriver> d;
};

using sc::fail;

extern "C" int sc_qdfd_create(const sc_qdfd_desc* desc, const double* mu_h, const double* rho_h,
                              const double* cs_h, const double* y_h, const double* z_h, const double* a_h,
                              const double* b_h, const double* Dc_h, const double* sN_h, const double* eta_h,
                              const double* locked_h, const double* prestress_h, const double* bcRShift_h,
                              sc_qdfd** out) {
  if (!desc || !mu_h || !rho_h || !cs_h || !a_h || !b_h || !Dc_h || !sN_h || !eta_h || !out)
    return fail(SC_ERR_ARG, "sc_qdfd_create: null argument");
  auto h = std::make_unique<sc_qdfd>();
  h->d = std::make_unique<sc::QdFdDriver>();
  const int e = h->d->init(*desc, mu_h, rho_h, cs_h, y_h, z_h, a_h, b_h, Dc_h, sN_h, eta_h, locked_h,
                           prestress_h, bcRShift_h);
  if (e) return e;
  *out = h.release();
  return SC_OK;
}

extern "C" int sc_qdfd_destroy(sc_qdfd* h) {
  delete h;
  return SC_OK;
}

extern "C" int sc_qdfd_time_step(sc_qdfd* h, double* dt) {
  if (!h || !dt) return fail(SC_ERR_ARG, "null argument");
  *dt = h->d->dtFd;
  return SC_OK;
}

extern "C" int sc_qdfd_set_state(sc_qdfd* h, const double* psi_h, const double* slip_h, const double* V_h) {
  if (!h || !psi_h || !slip_h) return fail(SC_ERR_ARG, "sc_qdfd_set_state: null argument");
  auto& d = *h->d;
  std::vector<double> zeros((size_t)d.Nz, 0.0);
  d.inDynamic = false;
  return sc_qd_set_state(d.qh, psi_h, slip_h, V_h ? V_h : zeros.data());
}

extern "C" int sc_qdfd_get_state(sc_qdfd* h, double* psi_h, double* slip_h, double* V_h, double* u_h,
                                 int* regime) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& d = *h->d;
  int e = d.sync_both();
  if (e) return e;
  if (regime) *regime = d.inDynamic ? 1 : 0;
  if (!d.inDynamic) {
    if ((psi_h || slip_h || V_h) && (e = sc_qd_get_state(d.qh, psi_h, slip_h, V_h, nullptr))) return e;
    if (u_h && (e = sc_qd_get_u(d.qh, u_h))) return e;
    return SC_OK;
  }
  sc::FdDriver& f = *d.f;
  hipStream_t s = f.ctx->c.stream;
  const size_t bz = (size_t)d.Nz * sizeof(double);
  if (psi_h) SC_HIP(hipMemcpyAsync(psi_h, f.psi[f.cur], bz, hipMemcpyDeviceToHost, s));
  if (slip_h) SC_HIP(hipMemcpyAsync(slip_h, f.slip[f.cur], bz, hipMemcpyDeviceToHost, s));
  if (V_h) SC_HIP(hipMemcpyAsync(V_h, f.V, bz, hipMemcpyDeviceToHost, s));
  if (u_h) {
    double* tmp = d.uStart;  // free between switches
    if (f.f32)
      hipLaunchKernelGGL(sc::k_total_u<float>, dim3(sc::nblk(d.N)), dim3(256), 0, s, (const float*)f.u[f.cur],
                         (const double*)d.u0, tmp, (size_t)d.N);
    else
      hipLaunchKernelGGL(sc::k_total_u<double>, dim3(sc::nblk(d.N)), dim3(256), 0, s,
                         (const double*)f.u[f.cur], (const double*)d.u0, tmp, (size_t)d.N);
    SC_HIP(hipGetLastError());
    SC_HIP(hipMemcpyAsync(u_h, tmp, (size_t)d.N * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

extern "C" int sc_qdfd_integrate(sc_qdfd* h, const sc_qdfd_run* run, double* tEnd, int* steps, int* cycles,
                                 sc_qdfd_phase* log, int logCap, int* nPhases) {
  if (!h || !run) return fail(SC_ERR_ARG, "sc_qdfd_integrate: null argument");
  if (run->integrator != 0 && run->integrator != 1) return fail(SC_ERR_ARG, "integrator must be RK32 (0) or RK43 (1)");
  if (!(run->tol >= 1e-14)) return fail(SC_ERR_ARG, "timeStepTol must be >= 1e-14");
  if (run->maxStepCount < 0) return fail(SC_ERR_ARG, "maxStepCount must be >= 0");
  if (run->deferChecks < 0 || run->deferChecks > 2) return fail(SC_ERR_ARG, "deferChecks must be 0, 1 or 2");
  auto& d = *h->d;
  const int e = d.integrate(*run);
  if (tEnd) *tEnd = d.currT;
  if (steps) *steps = d.stepCount;
  if (cycles) *cycles = d.cycleCount;
  if (nPhases) *nPhases = (int)d.log.size();
  if (log)
    for (int i = 0; i < (int)d.log.size() && i < logCap; ++i) log[i] = d.log[i];
  return e;
}
