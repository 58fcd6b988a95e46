// Discrete adjoint of the quasi-dynamic step (linear elastic, aging law) with
// respect to the friction parameters a, b, Dc of every fault node.
//
// Forward map of one d_dt at state (psi, slip), time t (qd.cpp, d_dt):
//   bcL = slip / fts, rhs = R_L bcL + R_R bcR(t)       (k_qd_rhs_frame)
//   u = A^{-1} rhs                                     (solve)
//   tau = S u + prestress, S u = mu imet0 D1_row0 u    (k_qd_fault, dy_row0)
//   F(V) = a sN asinh(V e^{psi/a} / (2 v0)) + eta V - tau = 0   (root solve)
//   dpsi = G(psi, V) = (b v0/Dc)(e^{(f0-psi)/b} - V/v0),  dslip = V
// Its vector-Jacobian product with cotangents (pb, sb) of (dpsi, dslip):
//   Vb = sb + G_V pb;  psib = G_psi pb;  bbar += G_b pb;  Dcbar += G_Dc pb
//   unlocked nodes (implicit function theorem): w = Vb / F_V,
//   psib -= w F_psi, abar -= w F_a, taub = w (F_tau = -1)
//   slipb = (1/fts) R_L^T A^{-1} S^T taub   (A = A^T: the adjoint solve uses the
//   same operator apply and solver)
// RK32 with the step sizes held fixed (the tableau of OdeSolver RK32,
// odeSolver.cpp:505-627, as qd.cpp steps it):
//   k1 = y + h/2 f0, k2 = y - h f0 + 2h f1, y' = y + h/6 f0 + 2h/3 f1 + h/6 f2
// reversed stage by stage: fb_i = c_i yb', then the VJP at k2 feeds yb, fb0
// (-h), fb1 (+2h); the VJP at k1 feeds yb, fb0 (+h/2); the VJP at y feeds yb.
// The reverse pass recomputes each checkpoint segment forward, keeping psi
// and V of the three stages of each step (all the VJPs need), then walks the
// segment backwards.
// S^T puts a traction cotangent on the five rows next to the fault and R_L^T
// reads lambda on the nine rows there, both node by node along z, so the
// map taub -> slipb of a stage is one fixed Nz x Nz matrix
//   B = (1/fts) R_L^T A^{-1} S^T,
// column k the adjoint solve of a unit traction at node k.  The reverse pass
// either solves per stage or applies B, formed once by Nz such solves and
// kept by the mediator (A does not depend on a, b, Dc).
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "qd_internal.hpp"
#include "reduce_dev.hpp"
#include "sbp_dev.hpp"

namespace sc {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

struct AdjFault {
  int Nz;
  const double *a, *b, *Dc, *sN, *eta, *locked;
  double f0, v0;
};

// VJP through the fault rates at one stage; writes psi_bar (the direct part),
// rhs rows 0..4 = S^T tau_bar, and adds to the parameter gradients.
__global__ __launch_bounds__(256) void k_adj_fault(DevOp<double> p, AdjFault P, const double* __restrict__ psi,
                                                   const double* __restrict__ V, const double* __restrict__ pb,
                                                   const double* __restrict__ sb, double* __restrict__ psi_bar,
                                                   double* __restrict__ ga, double* __restrict__ gb,
                                                   double* __restrict__ gDc, double* __restrict__ rhs,
                                                   double* __restrict__ taub_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.Nz) return;
  const double ps = psi[k], v = V[k], a = P.a[k], b = P.b[k], Dc = P.Dc[k];
  const double gpo = pb[k];
  // aging law (aging_law_psi: zero where b <= 1e-3 or the exponential overflows)
  const double A = exp((P.f0 - ps) / b);
  double G_psi = 0.0, G_V = 0.0, G_b = 0.0, G_Dc = 0.0;
  if (!isinf(A) && b > 1e-3) {
    const double c = P.v0 / Dc;
    G_psi = -c * A;
    G_V = -b / Dc;
    G_b = c * (A * (1.0 - (P.f0 - ps) / b) - v / P.v0);
    G_Dc = -(b * c) * (A - v / P.v0) / Dc;
  }
  const double Vb = sb[k] + G_V * gpo;
  double psib = G_psi * gpo;
  gb[k] += G_b * gpo;
  gDc[k] += G_Dc * gpo;
  double taub = 0.0;
  const double lk = P.locked ? P.locked[k] : 0.0;
  if (lk <= 0.5 && lk >= -0.5) {
    const double sN = P.sN[k];
    const double E = exp(ps / a);
    const double x = (0.5 * v / P.v0) * E;
    const double hx = hypot(1.0, x);
    const double F_V = a * sN * (0.5 * E / P.v0) / hx + P.eta[k];
    const double r = x / hx;
    const double F_psi = sN * r;
    const double F_a = sN * asinh(x) - sN * r * (ps / a);
    const double w = Vb / F_V;
    psib -= w * F_psi;
    ga[k] -= w * F_a;
    taub = w;
  }
  psi_bar[k] = psib;
  if (taub_out) taub_out[k] = taub;
  if (!rhs) return;
  // S^T: tau_qs[k] = mu_k imet0 sum_{q>=4} D1[q] u[(q-4) Nz + k]
  const double t = p.mu[k] * (p.y.imet[0] * taub);
#pragma unroll
  for (int q = 4; q < kBand; ++q) rhs[(size_t)(q - 4) * P.Nz + k] = p.y.D1[q] * t;
}

// S^T e_kk: the rhs rows 0..4 of a unit traction cotangent at fault node kk
// (k_adj_fault's arithmetic with taub = 1)
__global__ __launch_bounds__(256) void k_green_rhs(DevOp<double> p, int Nz, int kk, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 5 * Nz) return;
  const int q = i / Nz, k = i - q * Nz;
  rhs[i] = k == kk ? p.y.D1[q + 4] * (p.mu[kk] * p.y.imet[0]) : 0.0;
}

// column kk of B: (1/fts) R_L^T lam (k_adj_combine's arithmetic)
__global__ __launch_bounds__(256) void k_green_col(DevOp<double> p, int Nz, int kk, const double* __restrict__ lam,
                                                   double invfts, double* __restrict__ B) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Nz) return;
  const int nj = min(kBand, p.Ny);
  double s = 0.0;
  for (int j = 0; j < nj; ++j) {
    const double c = rhs_node(p, j, k, [](int side, int) { return side == 0 ? 1.0 : 0.0; });
    s += c * lam[(size_t)j * Nz + k];
  }
  B[(size_t)k * Nz + kk] = invfts * s;
}

// kb = (psi_bar, (1/fts) R_L^T lam); targets t_i += c_i kb (psi and slip parts).
struct AdjCombine {
  int n;                       // number of targets
  double* tp[3];
  double* ts[3];
  double c[3];
};
__global__ __launch_bounds__(256) void k_adj_combine(DevOp<double> p, int Nz, const double* __restrict__ lam,
                                                     double invfts, const double* __restrict__ psi_bar,
                                                     AdjCombine T) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Nz) return;
  const int nj = min(kBand, p.Ny);
  double s = 0.0;
  for (int j = 0; j < nj; ++j) {
    const double c = rhs_node(p, j, k, [](int side, int) { return side == 0 ? 1.0 : 0.0; });
    s += c * lam[(size_t)j * Nz + k];
  }
  const double sb = invfts * s;
  const double pbv = psi_bar ? psi_bar[k] : 0.0;
  for (int i = 0; i < T.n; ++i) {
    T.tp[i][k] += T.c[i] * pbv;
    T.ts[i][k] += T.c[i] * sb;
  }
}

// the stage's slip cotangent through B (one wave per fault node: row k of
// B against taub), then the targets as in k_adj_combine
__global__ __launch_bounds__(256) void k_adj_green(int Nz, const double* __restrict__ B,
                                                   const double* __restrict__ taub,
                                                   const double* __restrict__ psi_bar, AdjCombine T) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= Nz) return;  // the whole wave
  const double* b = B + (size_t)row * Nz;
  double s = 0.0;
  for (int k = lane; k < Nz; k += 64) s += b[k] * taub[k];
  s = wave_sum(s);
  if (lane == 0) {
    const double pbv = psi_bar ? psi_bar[row] : 0.0;
    for (int i = 0; i < T.n; ++i) {
      T.tp[i][row] += T.c[i] * pbv;
      T.ts[i][row] += T.c[i] * s;
    }
  }
}

// fb_i = c_i * yb (psi and slip parts), i = 0..2
struct AdjInit {
  const double *yp, *ys;
  double* fp[3];
  double* fs[3];
  double c[3];
};
__global__ __launch_bounds__(256) void k_adj_init(int Nz, AdjInit I) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Nz) return;
  const double a = I.yp[k], b = I.ys[k];
  for (int i = 0; i < 3; ++i) {
    I.fp[i][k] = I.c[i] * a;
    I.fs[i][k] = I.c[i] * b;
  }
}

// per-step record: psi and V of the three stages
struct Copy6 {
  const double* src[6];
  double* dst[6];
};
__global__ __launch_bounds__(256) void k_copy6(int n, Copy6 C) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int i = 0; i < 6; ++i) C.dst[i][k] = C.src[i][k];
}

// us = u at the station nodes j_s * Nz; rhs (if set) = Q^T (us - d)
__global__ void k_station_rhs(int ns, const int* __restrict__ js, int Nz, const double* __restrict__ u,
                              const double* __restrict__ d, double* __restrict__ rhs, double* __restrict__ us) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const size_t I = (size_t)js[i] * Nz;
  const double ui = u[I];
  us[i] = ui;
  if (rhs) rhs[I] = ui - d[i];
}

class AdjDriver {
 public:
  AdjDriver(QdDriver& q, const sc_adj_desc& d) : q_(q), d_(d) {}
  ~AdjDriver();
  int forward_misfit(double* J, double* us_h, bool keep_records);
  int forward_misfit_only(double* J, double* us_h) {
    int e = setup(false);
    if (!e) e = forward_misfit(J, us_h, false);
    if (!e) e = restore(0);
    if (!e) SC_HIP(hipStreamSynchronize(q_.ctx->c.stream));
    return e;
  }
  int gradient(double* ga_h, double* gb_h, double* gDc_h, sc_adj_stats* st);

 private:
  QdDriver& q_;
  const sc_adj_desc& d_;
  int Nz_ = 0, nseg_ = 0;
  std::vector<void*> owned_;
  std::vector<double> tseg_;                 // time at each checkpoint
  double* ckpt_ = nullptr;                   // nseg x 3 x Nz: psi, slip, V
  double* rec_ = nullptr;                    // ckptEvery x 6 x Nz
  int* js_d_ = nullptr;
  double* data_d_ = nullptr;
  double* us_d_ = nullptr;
  sc_ctx* adj_ = nullptr;
  double *rhs_ = nullptr, *rhsQ_ = nullptr, *lam_ = nullptr;
  double *yb_ = nullptr, *fb_ = nullptr, *pbar_ = nullptr, *grad_ = nullptr;  // yb: 2Nz, fb: 3x2Nz, grad: 3Nz
  long long adjSolves_ = 0, adjIters_ = 0, recomputed_ = 0;
  double* taub_ = nullptr;     // a stage's traction cotangent (the B path)
  bool green_ = false;         // the reverse pass applies B
  long long greenSolves_ = 0;
  double greenSec_ = 0.0;

  bool allocFailed_ = false;
  double* dal(size_t n) {
    double* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
      allocFailed_ = true;
      return nullptr;
    }
    owned_.push_back(p);
    (void)hipMemset(p, 0, n * sizeof(double));
    return p;
  }
  int setup(bool reverse);
  int step(int n, double t, double* rec);
  int save_ckpt(int seg);
  int restore(int seg);
  int run_segment(int seg, bool record);
  int adj_vjp(const double* psi, const double* V, const double* pb, const double* sb, const AdjCombine& T);
  int adj_solve(const double* rhs);
  int adj_step(int local, double h);
  bool green_ready() const;
  int build_green();
};

AdjDriver::~AdjDriver() {
  if (adj_) (void)hipStreamSynchronize(adj_->c.stream);  // the context is the mediator's
  (void)hipStreamSynchronize(q_.ctx->c.stream);
  for (void* p : owned_) (void)hipFree(p);
}

int AdjDriver::setup(bool reverse) {
  Nz_ = q_.Nz;
  // a run that ended in an error may have left the device flags set (and a
  // residual test pending)
  q_.finPending = false;
  SC_HIP(hipMemsetAsync(q_.status_d, 0, 2 * sizeof(int), q_.ctx->c.stream));
  const int C = d_.ckptEvery;
  nseg_ = (d_.nSteps + C - 1) / C;
  ckpt_ = dal((size_t)nseg_ * 3 * Nz_);
  rec_ = dal((size_t)C * 6 * Nz_);
  us_d_ = dal((size_t)std::max(d_.nStations, 1));
  data_d_ = dal((size_t)std::max(d_.nStations, 1));
  if (hipMalloc(&js_d_, sizeof(int) * std::max(d_.nStations, 1)) != hipSuccess) return fail(SC_ERR_HIP, "hipMalloc");
  owned_.push_back(js_d_);
  if (allocFailed_) return fail(SC_ERR_HIP, "adjoint: device allocation failed");
  SC_HIP(hipMemcpy(js_d_, d_.stationJ_h, sizeof(int) * d_.nStations, hipMemcpyHostToDevice));
  if (d_.data_h) SC_HIP(hipMemcpy(data_d_, d_.data_h, sizeof(double) * d_.nStations, hipMemcpyHostToDevice));
  tseg_.assign(nseg_, 0.0);
  if (!reverse) return SC_OK;
  const size_t N = (size_t)q_.N;
  rhs_ = dal(N);
  rhsQ_ = dal(N);
  lam_ = dal(N);
  yb_ = dal(2 * (size_t)Nz_);
  fb_ = dal(6 * (size_t)Nz_);
  pbar_ = dal((size_t)Nz_);
  grad_ = dal(3 * (size_t)Nz_);
  taub_ = dal((size_t)Nz_);
  if (allocFailed_) return fail(SC_ERR_HIP, "adjoint: device allocation failed");
  // the adjoint's own context for the mediator's operator as it stands
  const Ctx& c = q_.ctx->c;
  if (!q_.adjCtx || q_.adjCtxGen != c.opGen) {
    if (q_.adjCtx) {
      sc_ctx_destroy(q_.adjCtx);
      q_.adjCtx = nullptr;
    }
    q_.adjCtxGen = -1;
    int e = sc_ctx_create(&c.d, q_.y_host.empty() ? nullptr : q_.y_host.data(),
                          q_.z_host.empty() ? nullptr : q_.z_host.data(), c.mu_h.data(), &q_.adjCtx);
    if (e) return e;
    q_.adjCtxGen = c.opGen;
  }
  adj_ = q_.adjCtx;
  // every gradient starts its adjoint solves from an empty projection basis,
  // so that the same inputs give the same gradient
  adj_->c.projCount = adj_->c.projNext = 0;
  return SC_OK;
}

bool AdjDriver::green_ready() const {
  const sc_solve_opts& o = d_.adjSolve;
  return q_.adjB && q_.adjBGen == q_.ctx->c.opGen && q_.adjBPrec == o.precond && q_.adjBMode == o.mode &&
         q_.adjBRtol == o.rtol;
}

// B column by column: the adjoint solve of S^T e_k from zero with the
// adjoint solve options (no projected guess), then (1/fts) R_L^T of it
int AdjDriver::build_green() {
  if (green_ready()) return SC_OK;
  const double t0 = now_s();
  const size_t N = (size_t)q_.N;
  if (!q_.adjB) SC_HIP(hipMalloc(&q_.adjB, (size_t)Nz_ * Nz_ * sizeof(double)));
  q_.adjBGen = -1;  // incomplete until the last column
  hipStream_t s = adj_->c.stream;
  sc_solve_opts o = d_.adjSolve;
  o.reuse = 0;
  const double invfts = 1.0 / q_.desc.faultTypeScale;
  for (int kk = 0; kk < Nz_; ++kk) {
    hipLaunchKernelGGL(k_green_rhs, dim3((5 * Nz_ + 255) / 256), dim3(256), 0, s, adj_->c.op64, Nz_, kk, rhs_);
    SC_HIP(hipMemsetAsync(lam_, 0, N * sizeof(double), s));
    sc_solve_stats st;
    int e = solve(adj_->c, rhs_, lam_, o, st, nullptr, nullptr, true);
    if (e) return e;
    greenSolves_ += 1;
    hipLaunchKernelGGL(k_green_col, dim3((Nz_ + 255) / 256), dim3(256), 0, s, adj_->c.op64, Nz_, kk,
                       (const double*)lam_, invfts, q_.adjB);
  }
  SC_HIP(hipGetLastError());
  // the next adjoint solve starts from zero, as without the operator
  SC_HIP(hipMemsetAsync(lam_, 0, N * sizeof(double), s));
  SC_HIP(hipStreamSynchronize(s));
  q_.adjBGen = q_.ctx->c.opGen;
  q_.adjBPrec = d_.adjSolve.precond;
  q_.adjBMode = d_.adjSolve.mode;
  q_.adjBRtol = d_.adjSolve.rtol;
  greenSec_ = now_s() - t0;
  return SC_OK;
}

// One RK32 step of size h from the current state at time t (the stages of
// QdDriver::integrate); rec: psi and V of the three stages.
int AdjDriver::step(int n, double t, double* rec) {
  (void)n;
  const double h = d_.dts_h[n];
  int e;
  { const int fi[1] = {0}; const double c[1] = {0.5 * h}; if ((e = q_.lincomb_all("k", 1, 1, fi, c))) return e; }
  if ((e = q_.stage(1, t + 0.5 * h))) return e;
  { const int fi[2] = {0, 1}; const double c[2] = {-h, 2 * h}; if ((e = q_.lincomb_all("k", 2, 2, fi, c))) return e; }
  if ((e = q_.stage(2, t + h))) return e;
  VecSet& P = q_.vars["psi"];
  VecSet& S = q_.vars["slip"];
  if (rec) {
    Copy6 C{{P.var, S.dvar, P.k[1], S.f[1], P.k[2], S.f[2]}, {}};
    for (int i = 0; i < 6; ++i) C.dst[i] = rec + (size_t)i * Nz_;
    hipLaunchKernelGGL(k_copy6, dim3((Nz_ + 255) / 256), dim3(256), 0, q_.ctx->c.stream, Nz_, C);
    SC_HIP(hipGetLastError());
  }
  { const int fi[3] = {0, 1, 2}; const double c[3] = {h / 6.0, 2 * h / 3.0, h / 6.0}; if ((e = q_.lincomb_all("hi", 0, 3, fi, c))) return e; }
  for (auto& kv : q_.vars) std::swap(kv.second.var, kv.second.hi);
  return q_.d_dt_vars(t + h);
}

int AdjDriver::save_ckpt(int seg) {
  hipStream_t s = q_.ctx->c.stream;
  const size_t b = (size_t)Nz_ * sizeof(double);
  double* c = ckpt_ + (size_t)seg * 3 * Nz_;
  SC_HIP(hipMemcpyAsync(c, q_.vars["psi"].var, b, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(c + Nz_, q_.vars["slip"].var, b, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(c + 2 * Nz_, q_.V_d, b, hipMemcpyDeviceToDevice, s));
  return SC_OK;
}

int AdjDriver::restore(int seg) {
  hipStream_t s = q_.ctx->c.stream;
  const size_t b = (size_t)Nz_ * sizeof(double);
  const double* c = ckpt_ + (size_t)seg * 3 * Nz_;
  SC_HIP(hipMemcpyAsync(q_.vars["psi"].var, c, b, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(q_.vars["slip"].var, c + Nz_, b, hipMemcpyDeviceToDevice, s));
  SC_HIP(hipMemcpyAsync(q_.V_d, c + 2 * Nz_, b, hipMemcpyDeviceToDevice, s));
  return SC_OK;
}

// Steps of segment seg from its checkpoint (state already there, rates at the
// state computed): solves with a projected guess are checked once at the
// end; a miss repeats the segment with checked solves.
int AdjDriver::run_segment(int seg, bool record) {
  const int C = d_.ckptEvery;
  const int n0 = seg * C, n1 = std::min(d_.nSteps, n0 + C);
  for (int attempt = 0; attempt < 2; ++attempt) {
    double t = tseg_[seg];
    int e = SC_OK;
    for (int n = n0; n < n1 && !e; ++n) {
      q_.defer = attempt == 0 && q_.defer_ok();
      e = step(n, t, record ? rec_ + (size_t)(n - n0) * 6 * Nz_ : nullptr);
      t += d_.dts_h[n];
    }
    q_.defer = false;
    if (e) return e;
    int status, miss;
    if ((e = q_.read_flags(&status, &miss))) return e;
    q_.note_checked(miss != 0);
    if (miss) {
      q_.counters.missed_guesses++;
      SC_HIP(hipMemsetAsync(q_.status_d, 0, 2 * sizeof(int), q_.ctx->c.stream));
      if ((e = restore(seg))) return e;
      if ((e = q_.d_dt_vars(tseg_[seg]))) return e;
      continue;
    }
    if (status) {
      (void)hipMemsetAsync(q_.status_d, 0, sizeof(int), q_.ctx->c.stream);
      return fail(status & SC_ERR_NONFINITE ? SC_ERR_NONFINITE : SC_ERR_NOCONV,
                  "fault root solve failed (status " + std::to_string(status) + ")");
    }
    return SC_OK;
  }
  return fail(SC_ERR_NOCONV, "adjoint: segment recomputation failed");
}

// Forward from the current state: checkpoints, J, station displacements;
// the state is restored afterwards.  keep_records: the records of the last
// segment stay in rec_.
int AdjDriver::forward_misfit(double* J, double* us_h, bool keep_records) {
  int e = q_.d_dt_vars(d_.t0);
  if (e) return e;
  double t = d_.t0;
  for (int seg = 0; seg < nseg_; ++seg) {
    tseg_[seg] = t;
    if ((e = save_ckpt(seg))) return e;
    if ((e = run_segment(seg, keep_records && seg == nseg_ - 1))) return e;
    for (int n = seg * d_.ckptEvery; n < std::min(d_.nSteps, (seg + 1) * d_.ckptEvery); ++n) t += d_.dts_h[n];
  }
  // u of the last d_dt is at the final state
  hipStream_t s = q_.ctx->c.stream;
  if ((e = q_.materialize_u())) return e;
  if (d_.nStations > 0) {
    hipLaunchKernelGGL(k_station_rhs, dim3((d_.nStations + 63) / 64), dim3(64), 0, s, d_.nStations, js_d_, Nz_,
                       (const double*)q_.u_d, (const double*)data_d_, rhsQ_, us_d_);
    SC_HIP(hipGetLastError());
  }
  std::vector<double> us(std::max(d_.nStations, 1));
  SC_HIP(hipMemcpyAsync(us.data(), us_d_, sizeof(double) * d_.nStations, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  double j = 0.0;
  for (int i = 0; i < d_.nStations; ++i) {
    const double r = us[i] - (d_.data_h ? d_.data_h[i] : 0.0);
    j += 0.5 * r * r;
  }
  *J = j;
  if (us_h) std::memcpy(us_h, us.data(), sizeof(double) * d_.nStations);
  return SC_OK;
}

int AdjDriver::adj_solve(const double* rhs) {
  sc_solve_stats st;
  int e = solve(adj_->c, rhs, lam_, d_.adjSolve, st, nullptr, nullptr, true);
  adjSolves_ += 1;
  adjIters_ += st.iters;
  return e;
}

int AdjDriver::adj_vjp(const double* psi, const double* V, const double* pb, const double* sb,
                       const AdjCombine& T) {
  hipStream_t s = adj_->c.stream;
  const AdjFault P{Nz_, q_.a_d, q_.b_d, q_.Dc_d, q_.sN_d, q_.eta_d, q_.locked_d, q_.desc.f0, q_.desc.v0};
  const dim3 g((Nz_ + 255) / 256), b(256);
  if (green_) {
    hipLaunchKernelGGL(k_adj_fault, g, b, 0, s, adj_->c.op64, P, psi, V, pb, sb, pbar_, grad_, grad_ + Nz_,
                       grad_ + 2 * Nz_, (double*)nullptr, taub_);
    hipLaunchKernelGGL(k_adj_green, dim3((Nz_ + 3) / 4), b, 0, s, Nz_, (const double*)q_.adjB,
                       (const double*)taub_, (const double*)pbar_, T);
    SC_HIP(hipGetLastError());
    return SC_OK;
  }
  hipLaunchKernelGGL(k_adj_fault, g, b, 0, s, adj_->c.op64, P, psi, V, pb, sb, pbar_, grad_, grad_ + Nz_,
                     grad_ + 2 * Nz_, rhs_, (double*)nullptr);
  SC_HIP(hipGetLastError());
  int e = adj_solve(rhs_);
  if (e) return e;
  hipLaunchKernelGGL(k_adj_combine, g, b, 0, s, adj_->c.op64, Nz_, (const double*)lam_,
                     1.0 / q_.desc.faultTypeScale, (const double*)pbar_, T);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

// Reverse one RK32 step (record `local` of the segment): yb_ holds the
// cotangent of the state after the step on entry, before it on exit.
int AdjDriver::adj_step(int local, double h) {
  hipStream_t s = adj_->c.stream;
  const double* rec = rec_ + (size_t)local * 6 * Nz_;
  double* ybp = yb_;
  double* ybs = yb_ + Nz_;
  double* fp[3];
  double* fs[3];
  for (int i = 0; i < 3; ++i) {
    fp[i] = fb_ + (size_t)(2 * i) * Nz_;
    fs[i] = fb_ + (size_t)(2 * i + 1) * Nz_;
  }
  AdjInit I{ybp, ybs, {fp[0], fp[1], fp[2]}, {fs[0], fs[1], fs[2]}, {h / 6.0, 2 * h / 3.0, h / 6.0}};
  hipLaunchKernelGGL(k_adj_init, dim3((Nz_ + 255) / 256), dim3(256), 0, s, Nz_, I);
  SC_HIP(hipGetLastError());
  int e;
  {  // stage 2 (k2 = y - h f0 + 2h f1)
    AdjCombine T{3, {ybp, fp[0], fp[1]}, {ybs, fs[0], fs[1]}, {1.0, -h, 2 * h}};
    if ((e = adj_vjp(rec + 4 * (size_t)Nz_, rec + 5 * (size_t)Nz_, fp[2], fs[2], T))) return e;
  }
  {  // stage 1 (k1 = y + h/2 f0)
    AdjCombine T{2, {ybp, fp[0], nullptr}, {ybs, fs[0], nullptr}, {1.0, 0.5 * h, 0.0}};
    if ((e = adj_vjp(rec + 2 * (size_t)Nz_, rec + 3 * (size_t)Nz_, fp[1], fs[1], T))) return e;
  }
  {  // stage 0 (the rates at y)
    AdjCombine T{1, {ybp, nullptr, nullptr}, {ybs, nullptr, nullptr}, {1.0, 0.0, 0.0}};
    if ((e = adj_vjp(rec, rec + (size_t)Nz_, fp[0], fs[0], T))) return e;
  }
  return SC_OK;
}

int AdjDriver::gradient(double* ga_h, double* gb_h, double* gDc_h, sc_adj_stats* st) {
  int e = setup(true);
  if (e) return e;
  const sc_qd_counters c0 = q_.counters;
  const double t0 = now_s();
  double J = 0.0;
  if ((e = forward_misfit(&J, nullptr, true))) return e;
  const double t1 = now_s();
  const long long fwdSolves = q_.counters.solves - c0.solves;
  hipStream_t s = adj_->c.stream;
  // the reverse pass's adjoint solves: through B when asked, when B is
  // already formed, or when the run needs at least Nz of them anyway
  green_ = d_.greenMode == 1 || (d_.greenMode == 0 && (green_ready() || 3LL * d_.nSteps >= Nz_));
  if (green_ && (e = build_green())) return e;
  // terminal cotangent: yb = (0, (1/fts) R_L^T A^{-1} Q^T (Qu - d))
  if ((e = adj_solve(rhsQ_))) return e;
  {
    AdjCombine T{1, {yb_, nullptr, nullptr}, {yb_ + Nz_, nullptr, nullptr}, {1.0, 0.0, 0.0}};
    hipLaunchKernelGGL(k_adj_combine, dim3((Nz_ + 255) / 256), dim3(256), 0, s, adj_->c.op64, Nz_,
                       (const double*)lam_, 1.0 / q_.desc.faultTypeScale, (const double*)nullptr, T);
    SC_HIP(hipGetLastError());
  }
  const int C = d_.ckptEvery;
  for (int seg = nseg_ - 1; seg >= 0; --seg) {
    const int n0 = seg * C, n1 = std::min(d_.nSteps, n0 + C);
    if (seg != nseg_ - 1) {
      SC_HIP(hipStreamSynchronize(s));  // the records are about to be rewritten
      if ((e = restore(seg))) return e;
      if ((e = q_.d_dt_vars(tseg_[seg]))) return e;
      if ((e = run_segment(seg, true))) return e;
      recomputed_ += n1 - n0;
    }
    SC_HIP(hipStreamSynchronize(q_.ctx->c.stream));
    for (int n = n1 - 1; n >= n0; --n)
      if ((e = adj_step(n - n0, d_.dts_h[n]))) return e;
  }
  SC_HIP(hipStreamSynchronize(s));
  const size_t b = (size_t)Nz_ * sizeof(double);
  if (ga_h) SC_HIP(hipMemcpy(ga_h, grad_, b, hipMemcpyDeviceToHost));
  if (gb_h) SC_HIP(hipMemcpy(gb_h, grad_ + Nz_, b, hipMemcpyDeviceToHost));
  if (gDc_h) SC_HIP(hipMemcpy(gDc_h, grad_ + 2 * Nz_, b, hipMemcpyDeviceToHost));
  const double t2 = now_s();
  if ((e = restore(0))) return e;
  SC_HIP(hipStreamSynchronize(q_.ctx->c.stream));
  if (st) {
    st->J = J;
    st->forward_seconds = t1 - t0;
    st->reverse_seconds = t2 - t1;
    st->fwd_solves = fwdSolves;
    st->adj_solves = adjSolves_;
    st->adj_cg_iters = adjIters_;
    st->recomputed_steps = recomputed_;
    st->green_solves = greenSolves_;
    st->green_seconds = greenSec_;
  }
  return SC_OK;
}

int check_adj(const QdDriver& q, const sc_adj_desc* d) {
  if (!d || !d->dts_h || d->nSteps < 1 || d->ckptEvery < 1 || d->nStations < 0 ||
      (d->nStations > 0 && !d->stationJ_h) || d->greenMode < 0 || d->greenMode > 2)
    return fail(SC_ERR_ARG, "adjoint: bad description");
  if (q.pl) return fail(SC_ERR_ARG, "adjoint: the power-law mediator is not supported");
  if (q.desc.stateLaw != SC_AGING_LAW) return fail(SC_ERR_ARG, "adjoint: only the aging law is differentiated");
  if (q.Ny < kBand) return fail(SC_ERR_ARG, "adjoint: Ny must be at least 9");
  for (int i = 0; i < d->nStations; ++i)
    if (d->stationJ_h[i] < 0 || d->stationJ_h[i] >= q.Ny) return fail(SC_ERR_ARG, "adjoint: station index out of range");
  return SC_OK;
}

}  // namespace sc

using sc::fail;

extern "C" int sc_qd_set_friction(sc_qd* h, const double* a_h, const double* b_h, const double* Dc_h) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  auto& q = *h->q;
  SC_HIP(hipStreamSynchronize(q.ctx->c.stream));
  const size_t b = (size_t)q.Nz * sizeof(double);
  if (a_h) SC_HIP(hipMemcpy(q.a_d, a_h, b, hipMemcpyHostToDevice));
  if (b_h) SC_HIP(hipMemcpy(q.b_d, b_h, b, hipMemcpyHostToDevice));
  if (Dc_h) SC_HIP(hipMemcpy(q.Dc_d, Dc_h, b, hipMemcpyHostToDevice));
  return SC_OK;
}

extern "C" int sc_qd_forward_misfit(sc_qd* h, const sc_adj_desc* d, double* J, double* us_h) {
  if (!h || !J) return fail(SC_ERR_ARG, "null argument");
  int e = sc::check_adj(*h->q, d);
  if (e) return e;
  sc::AdjDriver a(*h->q, *d);
  e = a.forward_misfit_only(J, us_h);
  return e;
}

extern "C" int sc_qd_adjoint(sc_qd* h, const sc_adj_desc* d, double* ga_h, double* gb_h, double* gDc_h,
                             sc_adj_stats* st) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  int e = sc::check_adj(*h->q, d);
  if (e) return e;
  sc::AdjDriver a(*h->q, *d);
  return a.gradient(ga_h, gb_h, gDc_h, st);
}
