// The fully dynamic mediator's integrate loop and output series:
// OdeSolver_WaveEq::integrate (odeSolver_WaveEq.cpp:147-186) with
// strikeSlip_linearElastic_fd::timeMonitor (strikeSlip_linearElastic_fd.cpp
// :218-251), writeStep1D/2D (:254-305), LinearElastic::writeStep1D/2D
// (linearElastic.cpp:623-686) and Fault::writeStep (fault.cpp:512-557).
// The steps between output steps run as one batch of device launches.
#include "fd_internal.hpp"
#include "series_io.hpp"

namespace sc {

int FdDriver::body64(const double** ud) {
  if (!f32) {
    *ud = (const double*)u[cur];
    return SC_OK;
  }
  if (!u64) {
    u64 = (double*)dalloc((size_t)N * sizeof(double));
    if (!u64) return fail(SC_ERR_HIP, "hipMalloc failed");
  }
  std::vector<double> h((size_t)N);
  int e = downBody(h.data(), u[cur], (size_t)N);
  if (e) return e;
  SC_HIP(hipMemcpy(u64, h.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice));
  *ud = u64;
  return SC_OK;
}

// tau = sxy on the fault + prestress + tau0 (strikeSlip_linearElastic_fd.cpp
// :425-434); sxy from the body in FP64
int FdDriver::fault_tau(double* tau) {
  hipStream_t s = ctx->c.stream;
  const size_t nz = (size_t)Nz;
  const double* ud = nullptr;
  int e = body64(&ud);
  if (!e) e = launch_fault_sxy(ctx->c, ud, sxy, s);
  if (e) return e;
  std::vector<double> sx(nz), pre(nz), t0(nz);
  SC_HIP(hipMemcpyAsync(sx.data(), sxy, nz * 8, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(pre.data(), prestress, nz * 8, hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(t0.data(), tau0, nz * 8, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  for (size_t k = 0; k < nz; ++k) tau[k] = sx[k] + 1.0 * pre[k] + 1.0 * t0[k];
  return SC_OK;
}

int FdDriver::set_output(const sc_qd_output* o) {
  if (!o) {
    out.reset();
    return SC_OK;
  }
  if (!o->outputDir) return fail(SC_ERR_ARG, "sc_fd_set_output: outputDir is null");
  if (o->enableCheckpointing > 0) return fail(SC_ERR_ARG, "the fully dynamic mediator writes no checkpoints");
  std::unique_ptr<Output> n(new Output);
  n->dir = o->outputDir;
  n->stride1D = o->stride1D;
  n->stride2D = o->stride2D;
  if (n->stride2D > 0 && !sxy2d) {
    sxy2d = (double*)dalloc((size_t)N * sizeof(double));
    if (!sxy2d) return fail(SC_ERR_HIP, "hipMalloc failed");
  }
  out = std::move(n);
  return SC_OK;
}

int FdDriver::monitor(double t, double dt, int step, double maxTime) {
  Output& o = *out;
  const bool w1 = o.stride1D > 0 && (t == maxTime || step % o.stride1D == 0);
  const bool w2 = o.stride2D > 0 && (t == maxTime || step % o.stride2D == 0);
  if (!(w1 || w2)) return SC_OK;
  hipStream_t s = ctx->c.stream;
  SC_HIP(hipStreamSynchronize(s));
  const size_t nz = (size_t)Nz, ny = (size_t)Ny;
  const std::string& d = o.dir;
  int e;
  if (w1) {
    const bool app = o.started1D;
    std::vector<double> tau(nz), slp, vel, st, ps, sn;
    if ((e = fault_tau(tau.data())) || (e = d2h(slp, slip[cur], nz)) || (e = d2h(vel, V, nz)) ||
        (e = d2h(st, strength, nz)) || (e = d2h(ps, psi[cur], nz)) || (e = d2h(sn, sNEff, nz)))
      return e;
    std::vector<double> surf = surf0;
    surf.resize(ny, 0.0);
    // the momentum balance's boundary vectors: the FD mediator never sets
    // them, so they keep LinearElastic's zero start (linearElastic.cpp:305-319)
    const std::vector<double> zz(nz, 0.0), zy(ny, 0.0);
    if ((e = write_text(d + "med_time1D.txt", fmt_e(t), app)) || (e = write_text(d + "med_dt1D.txt", fmt_e(dt), app)) ||
        (e = write_vec(d + "surfDisp", surf, app)) || (e = write_vec(d + "momBal_bcR", zz, app)) ||
        (e = write_vec(d + "momBal_bcT", zy, app)) || (e = write_vec(d + "momBal_bcL", zz, app)) ||
        (e = write_vec(d + "momBal_bcB", zy, app)) || (e = write_vec(d + "momBal_bcRShift", zz, app)) ||
        (e = write_vec(d + "slip", slp, app)) || (e = write_vec(d + "slipVel", vel, app)) ||
        (e = write_vec(d + "tauP", tau, app)) || (e = write_vec(d + "tauQSP", tau, app)) ||
        (e = write_vec(d + "strength", st, app)) || (e = write_vec(d + "psi", ps, app)) ||
        (e = write_vec(d + "sNEff", sn, app)))
      return e;
    o.started1D = true;
  }
  if (w2) {
    const bool app = o.started2D;
    const double* ud = nullptr;
    if ((e = body64(&ud)) || (e = launch_deriv(ctx->c, 1, ud, sxy2d, s))) return e;
    SC_HIP(hipStreamSynchronize(s));
    std::vector<double> uh, sx;
    if ((e = d2h(uh, ud, (size_t)N)) || (e = d2h(sx, sxy2d, (size_t)N))) return e;
    if ((e = write_text(d + "med_time2D.txt", fmt_e(t), app)) || (e = write_text(d + "med_dt2D.txt", fmt_e(dt), app)) ||
        (e = write_vec(d + "momBal_u", uh, app)) || (e = write_vec(d + "momBal_sxy", sx, app)))
      return e;
    o.started2D = true;
  }
  return SC_OK;
}

int FdDriver::integrate(double t0, double dt, int maxSteps, double maxTime, double* tEnd, int* nSteps) {
  *tEnd = t0;
  *nSteps = 0;
  if (maxTime == t0) return SC_OK;
  if (dt == 0) dt = (maxTime - t0) / maxSteps;
  if (!(dt > 0)) return fail(SC_ERR_ARG, "sc_fd_integrate: the time step must be > 0");
  int e;
  if (out && (e = monitor(t0, dt, 0, maxTime))) return e;  // the initial condition
  double t = t0;
  int step = 0;
  while (step < maxSteps && t < maxTime) {
    // the steps up to the next one that the monitor writes (the time
    // recursion as the reference's, so the last step lands on maxTime alike)
    int n = 0;
    double tb = t;
    for (;;) {
      tb = tb + dt;
      if (tb > maxTime) tb = maxTime;
      ++n;
      const int sc = step + n;
      const bool last = sc >= maxSteps || tb >= maxTime;
      const bool written = out && ((out->stride1D > 0 && (tb == maxTime || sc % out->stride1D == 0)) ||
                                   (out->stride2D > 0 && (tb == maxTime || sc % out->stride2D == 0)));
      if (last || written) break;
    }
    if ((e = steps(t, dt, n))) return e;
    step += n;
    t = tb;
    if (out && (e = monitor(t, dt, step, maxTime))) return e;
  }
  *tEnd = t;
  *nSteps = step;
  return SC_OK;
}

}  // namespace sc

using sc::fail;

extern "C" int sc_fd_set_output(sc_fd* fd, const sc_qd_output* o) {
  if (!fd) return fail(SC_ERR_ARG, "sc_fd_set_output: null handle");
  return fd->d->set_output(o);
}

extern "C" int sc_fd_integrate(sc_fd* fd, double initTime, double deltaT, int maxSteps, double maxTime,
                               double* tEnd, int* steps) {
  if (!fd || !tEnd || !steps) return fail(SC_ERR_ARG, "sc_fd_integrate: null argument");
  if (maxSteps < 0 || deltaT < 0 || (deltaT == 0 && maxSteps == 0))
    return fail(SC_ERR_ARG, "sc_fd_integrate: maxSteps >= 0 and deltaT >= 0 (0 needs maxSteps > 0)");
  return fd->d->integrate(initTime, deltaT, maxSteps, maxTime, tEnd, steps);
}
