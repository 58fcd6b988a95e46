// The fully dynamic driver's state and step loop (fd.cpp, and the FD phase
// of the QD-FD switching mediator in qdfd.hip).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sc_internal.hpp"

namespace sc {

struct FdDriver {
  sc_fd_desc desc{};
  sc_ctx* ctx = nullptr;
  int Ny = 0, Nz = 0, N = 0;
  bool f32 = false;
  std::vector<void*> owned;
  // body: three levels (double or float), rho, a_y
  void* u[3] = {};
  void* rho = nullptr;
  void* ay = nullptr;
  void* d2u = nullptr;  // fault column of D2u (body type)
  // fault (FP64)
  double *a = nullptr, *b = nullptr, *Dc = nullptr, *sNEff = nullptr, *tau0 = nullptr,
         *prestress = nullptr, *slip0 = nullptr, *locked = nullptr, *rhoF = nullptr,
         *alphay = nullptr;
  double* psi[3] = {};
  double* slip[3] = {};
  double *V = nullptr, *strength = nullptr, *sxy = nullptr;
  int* status = nullptr;
  int prev = 0, cur = 1, next = 2;
  // output series (strikeSlip_linearElastic_fd::timeMonitor): the files'
  // directory prefix and strides, and whether each series has started
  struct Output {
    std::string dir;
    int stride1D = 0, stride2D = 0;
    bool started1D = false, started2D = false;
  };
  std::unique_ptr<Output> out;
  // LinearElastic's surfDisp is set from u at construction and by computeU,
  // which the FD mediator never calls: it keeps the initial state's surface
  std::vector<double> surf0;
  double* u64 = nullptr;    // FP64 copy of an FP32 body for the 2-D series
  double* sxy2d = nullptr;  // muxDy u over the grid for the 2-D series

  int set_output(const sc_qd_output* o);
  int monitor(double t, double dt, int step, double maxTime);
  int integrate(double t0, double dt, int maxSteps, double maxTime, double* tEnd, int* nSteps);
  // tau = sxy on the fault + prestress + tau0 at the current level
  int fault_tau(double* tau);
  // the current body level in FP64 on the device (the level itself, or a copy)
  int body64(const double** ud);

  ~FdDriver() {
    if (ctx) {
      (void)hipStreamSynchronize(ctx->c.stream);
      for (void* p : owned) (void)hipFree(p);
      sc_ctx_destroy(ctx);
    }
  }
  void* dalloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    owned.push_back(p);
    (void)hipMemsetAsync(p, 0, bytes, ctx->c.stream);
    return p;
  }
  int up64(double*& dst, const double* h, size_t n) {
    dst = (double*)dalloc(n * sizeof(double));
    if (!dst) return fail(SC_ERR_HIP, "hipMalloc failed");
    if (h) SC_HIP(hipMemcpyAsync(dst, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->c.stream));
    SC_HIP(hipStreamSynchronize(ctx->c.stream));
    return SC_OK;
  }
  // body array in the body type from FP64 host data
  int upBody(void*& dst, const double* h, size_t n) {
    dst = dalloc(n * (f32 ? sizeof(float) : sizeof(double)));
    if (!dst) return fail(SC_ERR_HIP, "hipMalloc failed");
    if (!h) return SC_OK;
    if (f32) {
      std::vector<float> t(h, h + n);
      SC_HIP(hipMemcpyAsync(dst, t.data(), n * sizeof(float), hipMemcpyHostToDevice, ctx->c.stream));
      SC_HIP(hipStreamSynchronize(ctx->c.stream));
    } else {
      SC_HIP(hipMemcpyAsync(dst, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->c.stream));
      SC_HIP(hipStreamSynchronize(ctx->c.stream));
    }
    return SC_OK;
  }
  int downBody(double* h, const void* src, size_t n) {
    hipStream_t s = ctx->c.stream;
    if (f32) {
      std::vector<float> t(n);
      SC_HIP(hipMemcpyAsync(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, s));
      SC_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < n; ++i) h[i] = t[i];
    } else {
      SC_HIP(hipMemcpyAsync(h, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
      SC_HIP(hipStreamSynchronize(s));
    }
    return SC_OK;
  }

  int init(const sc_fd_desc& d, const double* mu, const double* rho_h, const double* cs,
           const double* y, const double* z, const double* a_h, const double* b_h,
           const double* Dc_h, const double* sNEff_h, const double* tau0_h,
           const double* prestress_h, const double* slip0_h, const double* locked_h) {
    desc = d;
    if (d.bulkPrec != SC_FP64 && d.bulkPrec != SC_FP32)
      return fail(SC_ERR_ARG, "sc_fd_create: bulkPrec must be SC_FP64 or SC_FP32");
    if (d.grid.laplace != 0) return fail(SC_ERR_ARG, "sc_fd_create: laplace type must be yz");
    if (d.stateLaw != SC_AGING_LAW && d.stateLaw != SC_SLIP_LAW)
      return fail(SC_ERR_ARG, "sc_fd_create: stateLaw must be aging or slip");
    f32 = d.bulkPrec == SC_FP32;
    int e = sc_ctx_create(&d.grid, y, z, mu, &ctx);
    if (e) return e;
    Ny = d.grid.Ny;
    Nz = d.grid.Nz;
    N = Ny * Nz;
    const Ctx& c = ctx->c;
    // a_y = cs * (sum of 0.5/h11 over the absorbing sides the node lies on)
    // (computePenaltyVectors, strikeSlip_linearElastic_fd.cpp:640-665)
    const double h11y = c.ty.h11, h11z = c.tz.h11;
    std::vector<double> ayh(N);
    for (int I = 0; I < N; ++I) {
      double v = 0;
      if ((I / Nz == Ny - 1) && d.outR) v += 0.5 / h11y;
      if ((I % Nz == 0) && d.outT) v += 0.5 / h11z;
      if (((I + 1) % Nz == 0) && d.outB) v += 0.5 / h11z;
      if ((I / Nz == 0) && d.faultL) v += 0.5 / h11y;
      ayh[I] = v * cs[I];
    }
    // fault: rho on the fault column, alpha_y = yq * h11y (:616-632)
    std::vector<double> rf(rho_h, rho_h + Nz), al(Nz, h11y);
    if (d.grid.var_grid)
      for (int k = 0; k < Nz; ++k) al[k] = c.yq[0] * h11y;
    for (int i = 0; i < 3; ++i)
      if ((e = upBody(u[i], nullptr, N))) return e;
    if ((e = upBody(rho, rho_h, N))) return e;
    if ((e = upBody(ay, ayh.data(), N))) return e;
    if ((e = upBody(d2u, nullptr, Nz))) return e;
    if ((e = up64(a, a_h, Nz)) || (e = up64(b, b_h, Nz)) || (e = up64(Dc, Dc_h, Nz)) ||
        (e = up64(sNEff, sNEff_h, Nz)) || (e = up64(tau0, tau0_h, Nz)) ||
        (e = up64(prestress, prestress_h, Nz)) || (e = up64(slip0, slip0_h, Nz)) ||
        (e = up64(rhoF, rf.data(), Nz)) || (e = up64(alphay, al.data(), Nz)))
      return e;
    if (locked_h && (e = up64(locked, locked_h, Nz))) return e;
    for (int i = 0; i < 3; ++i)
      if ((e = up64(psi[i], nullptr, Nz)) || (e = up64(slip[i], nullptr, Nz))) return e;
    if ((e = up64(V, nullptr, Nz)) || (e = up64(strength, nullptr, Nz)) || (e = up64(sxy, nullptr, Nz)))
      return e;
    status = (int*)dalloc(sizeof(int));
    if (!status) return fail(SC_ERR_HIP, "hipMalloc failed");
    SC_HIP(hipStreamSynchronize(ctx->c.stream));
    return SC_OK;
  }

  // replaces a_y (N, FP64 host values, stored in the body type)
  int set_ay(const double* h) {
    hipStream_t s = ctx->c.stream;
    if (f32) {
      std::vector<float> t(h, h + N);
      SC_HIP(hipMemcpyAsync(ay, t.data(), (size_t)N * sizeof(float), hipMemcpyHostToDevice, s));
      SC_HIP(hipStreamSynchronize(s));
    } else {
      SC_HIP(hipMemcpyAsync(ay, h, (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
      SC_HIP(hipStreamSynchronize(s));
    }
    return SC_OK;
  }

  sc_fault_fd_args fault_args(double dt) const {
    sc_fault_fd_args f{};
    f.N = Nz;
    f.dt = dt;
    f.v0 = desc.v0;
    f.vL = desc.vL;
    f.f0 = desc.f0;
    f.rootTol = desc.rootTol;
    f.faultTypeScale = desc.faultTypeScale;
    f.maxNumIts = desc.maxNumIts;
    f.stateLaw = desc.stateLaw;
    f.a_d = a;
    f.b_d = b;
    f.Dc_d = Dc;
    f.sNEff_d = sNEff;
    f.rho_d = rhoF;
    f.alphay_d = alphay;
    f.tau0_d = tau0;
    f.slip0_d = slip0;
    f.locked_d = locked;
    f.d2u_d = f32 ? nullptr : (const double*)d2u;
    f.psi_d = psi[cur];
    f.psiPrev_d = psi[prev];
    f.slip_d = slip[cur];
    f.slipPrev_d = slip[prev];
    f.V_d = V;
    f.psiNext_d = psi[next];
    f.slipNext_d = slip[next];
    f.strength_d = strength;
    f.uFault_d = nullptr;
    f.status_d = status;
    return f;
  }

  // the wave update's per-node scale (dt^2/rho and the grid's H^-1 and
  // metric factors), formed again when dt changes
  void* c1 = nullptr;
  double c1dt = -1.0;
  int ensure_c1(double dt) {
    if (c1 && c1dt == dt) return SC_OK;
    hipStream_t s = ctx->c.stream;
    if (!c1) {
      c1 = dalloc((size_t)N * (f32 ? sizeof(float) : sizeof(double)));
      if (!c1) return fail(SC_ERR_HIP, "hipMalloc failed");
    }
    int e = f32 ? launch_wave_c1(ctx->c, (const float*)rho, (float*)c1, (float)dt, s)
                : launch_wave_c1(ctx->c, (const double*)rho, (double*)c1, dt, s);
    if (e) return e;
    c1dt = dt;
    return SC_OK;
  }

  int steps(double t0, double dt, int n) {
    (void)t0;  // no time-dependent forcing on this path (updatePrestress needs initialConditions "tau")
    const Ctx& c = ctx->c;
    hipStream_t s = c.stream;
    if (n > 0) {
      int e = ensure_c1(dt);
      if (e) return e;
    }
    // one launch a step: the wave update with the fault's step fused in
    for (int it = 0; it < n; ++it) {
      int e;
      const sc_fault_fd_args fa = fault_args(dt);
      if (f32) {
        WaveArgs<float> w{(const float*)u[cur], (const float*)u[prev], (float*)u[next], (const float*)c1,
                          (const float*)ay, (float*)d2u, (float)dt, 1, &fa};
        e = launch_wave32(c, w, s);
      } else {
        WaveArgs<double> w{(const double*)u[cur], (const double*)u[prev], (double*)u[next],
                           (const double*)c1, (const double*)ay, (double*)d2u, dt, 1, &fa};
        e = launch_wave64(c, w, s);
      }
      if (e) return e;
      const int p = prev;
      prev = cur;
      cur = next;
      next = p;
    }
    int st = 0;
    SC_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    if (st) {
      SC_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
      return fail(st, st == SC_ERR_NOCONV ? "FD fault root solve did not converge"
                                          : "FD fault step produced a non-finite value");
    }
    return SC_OK;
  }
};

}  // namespace sc

struct sc_fd {
  std::unique_ptr<sc::FdDriver> d;
};
