// Fully dynamic driver: StrikeSlip_LinearElastic_fd::d_dt
// (strikeSlip_linearElastic_fd.cpp:409-438) stepped by the constant-dt loop of
// OdeSolver_WaveEq::integrate (odeSolver_WaveEq.cpp:147-186).  Three time
// levels of the body displacement and of the fault state live on the device;
// a step is one wave-update launch (the operator apply with the update fused
// into its epilogue) and one fault launch, and the levels rotate by pointer
// swap.
#include "fd_internal.hpp"

using sc::fail;

extern "C" int sc_fd_create(const sc_fd_desc* desc, const double* mu, const double* rho,
                            const double* cs, const double* y, const double* z, const double* a,
                            const double* b, const double* Dc, const double* sNEff,
                            const double* tau0, const double* prestress, const double* slip0,
                            const double* locked, sc_fd** out) {
  if (!desc || !mu || !rho || !cs || !a || !b || !Dc || !sNEff || !tau0 || !prestress || !slip0 || !out)
    return fail(SC_ERR_ARG, "sc_fd_create: null argument");
  auto h = std::make_unique<sc_fd>();
  h->d = std::make_unique<sc::FdDriver>();
  int e = h->d->init(*desc, mu, rho, cs, y, z, a, b, Dc, sNEff, tau0, prestress, slip0, locked);
  if (e) return e;
  *out = h.release();
  return SC_OK;
}

extern "C" int sc_fd_destroy(sc_fd* fd) {
  delete fd;
  return SC_OK;
}

extern "C" sc_ctx* sc_fd_ctx(sc_fd* fd) { return fd ? fd->d->ctx : nullptr; }

extern "C" int sc_fd_set_state(sc_fd* fd, const double* uPrev, const double* u,
                               const double* psiPrev, const double* psi, const double* slipPrev,
                               const double* slip, const double* V) {
  if (!fd || !uPrev || !u || !psiPrev || !psi || !slipPrev || !slip)
    return fail(SC_ERR_ARG, "sc_fd_set_state: null argument");
  auto& d = *fd->d;
  hipStream_t s = d.ctx->c.stream;
  const size_t N = (size_t)d.N, Nz = (size_t)d.Nz;
  d.prev = 0;
  d.cur = 1;
  d.next = 2;
  if (d.f32) {
    std::vector<float> t0(uPrev, uPrev + N), t1(u, u + N);
    SC_HIP(hipMemcpyAsync(d.u[0], t0.data(), N * 4, hipMemcpyHostToDevice, s));
    SC_HIP(hipMemcpyAsync(d.u[1], t1.data(), N * 4, hipMemcpyHostToDevice, s));
    SC_HIP(hipStreamSynchronize(s));
  } else {
    SC_HIP(hipMemcpyAsync(d.u[0], uPrev, N * 8, hipMemcpyHostToDevice, s));
    SC_HIP(hipMemcpyAsync(d.u[1], u, N * 8, hipMemcpyHostToDevice, s));
  }
  SC_HIP(hipMemcpyAsync(d.psi[0], psiPrev, Nz * 8, hipMemcpyHostToDevice, s));
  SC_HIP(hipMemcpyAsync(d.psi[1], psi, Nz * 8, hipMemcpyHostToDevice, s));
  SC_HIP(hipMemcpyAsync(d.slip[0], slipPrev, Nz * 8, hipMemcpyHostToDevice, s));
  SC_HIP(hipMemcpyAsync(d.slip[1], slip, Nz * 8, hipMemcpyHostToDevice, s));
  if (V) SC_HIP(hipMemcpyAsync(d.V, V, Nz * 8, hipMemcpyHostToDevice, s));
  else SC_HIP(hipMemsetAsync(d.V, 0, Nz * 8, s));
  SC_HIP(hipStreamSynchronize(s));
  // the initial state's surface displacement u(y, z = 0)
  d.surf0.resize((size_t)d.Ny);
  for (int j = 0; j < d.Ny; ++j) d.surf0[(size_t)j] = u[(size_t)j * Nz];
  return SC_OK;
}

extern "C" int sc_fd_steps(sc_fd* fd, double t0, double dt, int nSteps) {
  if (!fd) return fail(SC_ERR_ARG, "sc_fd_steps: null handle");
  if (!(dt > 0) || nSteps < 0) return fail(SC_ERR_ARG, "sc_fd_steps: dt must be > 0, nSteps >= 0");
  return fd->d->steps(t0, dt, nSteps);
}

extern "C" int sc_fd_get_state(sc_fd* fd, double* u, double* psi, double* slip, double* V,
                               double* tau, double* strength) {
  if (!fd) return fail(SC_ERR_ARG, "sc_fd_get_state: null handle");
  auto& d = *fd->d;
  hipStream_t s = d.ctx->c.stream;
  const size_t Nz = (size_t)d.Nz;
  SC_HIP(hipStreamSynchronize(s));
  int e = SC_OK;
  if (u) e = d.downBody(u, d.u[d.cur], (size_t)d.N);
  if (e) return e;
  if (psi) SC_HIP(hipMemcpyAsync(psi, d.psi[d.cur], Nz * 8, hipMemcpyDeviceToHost, s));
  if (slip) SC_HIP(hipMemcpyAsync(slip, d.slip[d.cur], Nz * 8, hipMemcpyDeviceToHost, s));
  if (V) SC_HIP(hipMemcpyAsync(V, d.V, Nz * 8, hipMemcpyDeviceToHost, s));
  if (strength) SC_HIP(hipMemcpyAsync(strength, d.strength, Nz * 8, hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  if (tau && (e = d.fault_tau(tau))) return e;
  return SC_OK;
}
