// Output series and checkpoints of the linear-elastic QD mediator, and the
// series of the power-law QD mediator (StrikeSlip_PowerLaw_qd::timeMonitor,
// strikeSlip_powerLaw_qd.cpp:352-408, with PowerLaw::writeStep1D/2D
// powerLaw.cpp:2276-2372): the linear-elastic mediator's timeMonitor
// (strikeSlip_linearElastic_qd.cpp:417-463) with
// writeStep1D / writeStep2D (:466-525) and writeCheckpoint (:527-557), the
// fault's and the momentum balance's series and checkpoints (Fault::writeStep
// fault.cpp:512-557, writeCheckpoint :561-584, loadCheckpoint :144-176;
// LinearElastic::writeStep1D/2D linearElastic.cpp:623-686, writeCheckpoint
// :689-705, loadCheckpoint :396-420) and the restart values the mediator's
// constructor and integrate() read (:32-39, :729-737).  Files are the
// reference's: PETSc-binary Vec records (big-endian, one record appended per
// output step) and one-number ASCII files.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "input.hpp"
#include "qd_internal.hpp"
#include "series_io.hpp"

namespace sc {
namespace {

// loadValueFromCheckpoint (genFuncs.cpp:889-934): the first number of an
// ASCII file; false (value kept) when the file is missing
bool read_value(const std::string& path, double* v) {
  std::FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  double x;
  const bool ok = std::fscanf(f, "%lf", &x) == 1;
  std::fclose(f);
  if (ok) *v = x;
  return ok;
}

// loadVecFromInputFile: the first record when the file exists; n entries
// required
int read_vec(const std::string& path, size_t n, std::vector<double>& out, bool* found) {
  std::vector<std::vector<double>> recs;
  *found = false;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return SC_OK;
  std::fclose(f);
  if (read_petsc_vecs(path, recs) < 1) return fail(SC_ERR_IO, "not a PETSc binary Vec file: " + path);
  if (recs[0].size() != n)
    return fail(SC_ERR_IO, path + ": " + std::to_string(recs[0].size()) + " entries, expected " +
                               std::to_string(n));
  out = recs[0];
  *found = true;
  return SC_OK;
}

int h2d(double* d, const std::vector<double>& h) {
  SC_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return SC_OK;
}

}  // namespace

int QdDriver::set_output(const sc_qd_output* o) {
  if (!o) {
    out.reset();
    return SC_OK;
  }
  if (!o->outputDir) return fail(SC_ERR_ARG, "sc_qd_set_output: outputDir is null");
  // StrikeSlip_PowerLaw_qd writes series only (no writeCheckpoint)
  if (pl && o->enableCheckpointing > 0)
    return fail(SC_ERR_ARG, "the power-law mediator writes no checkpoints");
  if (o->enableCheckpointing > 0 && o->interval < 1)
    return fail(SC_ERR_ARG, "checkpoint interval must be >= 1");
  std::unique_ptr<QdOutput> n(new QdOutput);
  n->dir = o->outputDir;
  n->stride1D = o->stride1D;
  n->stride2D = o->stride2D;
  n->ckpt = o->enableCheckpointing;
  n->interval = o->interval;
  if (n->ckpt > 0) {  // Domain: ckptNumber from the output directory, appended series after one
    double v = 0;
    if (read_value(n->dir + "ckptNumber", &v)) n->ckptNumber = (int)v;
    n->append = n->ckptNumber > 0;
  }
  n->sxy_d = out ? out->sxy_d : nullptr;
  n->plf_d = out ? out->plf_d : nullptr;
  if (n->stride2D > 0 && !pl && !n->sxy_d) {
    n->sxy_d = dalloc((size_t)N, nullptr);
    if (!n->sxy_d) return fail(SC_ERR_HIP, "device allocation failed");
  }
  if (n->stride2D > 0 && pl && !n->plf_d) {
    n->plf_d = dalloc((size_t)8 * N, nullptr);
    if (!n->plf_d) return fail(SC_ERR_HIP, "device allocation failed");
  }
  out = std::move(n);
  return SC_OK;
}

int QdDriver::monitor(double t, double dt, int step, const double* errA, double newDt, int maxSteps,
                      double maxTime, bool initial, int* stop) {
  QdOutput& o = *out;
  // the power-law mediator's timeMonitor replaces both strides by 1 below
  // step 50 and 100 from there on (strikeSlip_powerLaw_qd.cpp:367-368); a
  // stride <= 0 still switches its series off
  const int s1 = pl ? (step < 50 ? 1 : 100) : o.stride1D;
  const int s2 = pl ? (step < 50 ? 1 : 100) : o.stride2D;
  const bool w1 = o.stride1D > 0 && (t == maxTime || step % s1 == 0);
  const bool w2 = o.stride2D > 0 && (t == maxTime || step % s2 == 0);
  const bool ck = o.ckpt > 0 && (step % o.interval == 0 || step >= maxSteps || t >= maxTime);
  if (!(w1 || w2 || ck)) return SC_OK;
  hipStream_t s = ctx->c.stream;
  int e = initial ? materialize_u() : settle(t);
  if (e) return e;
  SC_HIP(hipStreamSynchronize(s));
  const size_t nz = (size_t)Nz, ny = (size_t)Ny;
  std::vector<double> psi, slip, V, tau, tauQS, strength, sN, bcL, bcR, bcT, bcB, shift;
  if (w1 || ck) {
    if ((e = d2h(psi, vars["psi"].var, nz)) || (e = d2h(slip, vars["slip"].var, nz)) || (e = d2h(V, V_d, nz)) ||
        (e = d2h(tau, tau_d, nz)) || (e = d2h(tauQS, tauQS_d, nz)) || (e = d2h(strength, strength_d, nz)) ||
        (e = d2h(sN, sN_d, nz)) || (e = d2h(bcL, bcL_d, nz)) || (e = d2h(bcR, bcR_d, nz)) ||
        (e = d2h(bcT, bcT_d, ny)) || (e = d2h(bcB, bcB_d, ny)) || (e = d2h(shift, shift_d, nz)))
      return e;
  }
  if (w1) {
    const bool first = !o.started1D;
    const bool app = !first || o.append;
    // surface displacement u(y, z = 0): the first entry of every y-row
    std::vector<double> surf(ny);
    SC_HIP(hipMemcpy2D(surf.data(), sizeof(double), u_d, nz * sizeof(double), sizeof(double), ny,
                       hipMemcpyDeviceToHost));
    const std::string& d = o.dir;
    if ((e = write_text(d + "med_time1D.txt", fmt_e(t), app)) || (e = write_text(d + "med_dt1D.txt", fmt_e(dt), app)) ||
        (e = write_vec(d + "surfDisp", surf, app)) || (e = write_vec(d + "momBal_bcR", bcR, app)) ||
        (e = write_vec(d + "momBal_bcT", bcT, app)) || (e = write_vec(d + "momBal_bcL", bcL, app)) ||
        (e = write_vec(d + "momBal_bcB", bcB, app)) ||
        (e = write_vec(d + "momBal_bcRShift", first ? shift : bcR, app)) ||
        (e = write_vec(d + "slip", slip, app)) || (e = write_vec(d + "slipVel", V, app)) ||
        (e = write_vec(d + "tauP", tau, app)) || (e = write_vec(d + "tauQSP", tauQS, app)) ||
        (e = write_vec(d + "strength", strength, app)) || (e = write_vec(d + "psi", psi, app)) ||
        (e = write_vec(d + "sNEff", sN, app)))
      return e;
    o.started1D = true;
  }
  if (w2 && pl) {
    const bool app = o.started2D || o.append;
    const size_t n = (size_t)N;
    double* f = o.plf_d;
    const PlFields pf{u_d, vars["gVxy"].var, vars["gVxz"].var, pl->T, pl->A, pl->B, pl->n, pl->cap,
                      f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n, f + 7 * n};
    if ((e = launch_pl_fields(ctx->c, pf, s))) return e;
    SC_HIP(hipStreamSynchronize(s));
    // PowerLaw::writeStep2D's files, with the dislocation-creep extras
    struct Field {
      const char* file;
      const double* src;
    };
    const Field fields[] = {
        {"momBal_u", u_d},           {"momBal_sxy", f + 2 * n},        {"momBal_sxz", f + 3 * n},
        {"momBal_sdev", f + 4 * n},  {"momBal_gTxy", f},               {"momBal_gTxz", f + n},
        {"momBal_gxy", vars["gVxy"].var}, {"momBal_gxz", vars["gVxz"].var},
        {"momBal_dgVxy", vars["gVxy"].dvar}, {"momBal_dgVxz", vars["gVxz"].dvar},
        {"momBal_effVisc", f + 5 * n}, {"momBal_T", pl->T},            {"disl_invEffVisc", f + 6 * n},
        {"disl_dgVdev", f + 7 * n},
    };
    const std::string& d = o.dir;
    if ((e = write_text(d + "med_time2D.txt", fmt_e(t), app)) || (e = write_text(d + "med_dt2D.txt", fmt_e(dt), app)))
      return e;
    std::vector<double> h;
    for (const Field& fl : fields)
      if ((e = d2h(h, fl.src, n)) || (e = write_vec(d + fl.file, h, app))) return e;
    o.started2D = true;
  } else if (w2) {
    const bool app = o.started2D || o.append;
    if ((e = launch_deriv(ctx->c, 1, u_d, o.sxy_d, s))) return e;
    SC_HIP(hipStreamSynchronize(s));
    std::vector<double> u, sxy;
    if ((e = d2h(u, u_d, (size_t)N)) || (e = d2h(sxy, o.sxy_d, (size_t)N))) return e;
    const std::string& d = o.dir;
    if ((e = write_text(d + "med_time2D.txt", fmt_e(t), app)) || (e = write_text(d + "med_dt2D.txt", fmt_e(dt), app)) ||
        (e = write_vec(d + "momBal_u", u, app)) || (e = write_vec(d + "momBal_sxy", sxy, app)))
      return e;
    o.started2D = true;
  }
  if (ck) {
    o.ckptNumber++;
    const std::string& d = o.dir;
    if ((e = write_text(d + "ckptNumber", fmt_i(o.ckptNumber), false)) ||
        (e = write_text(d + "chkpt_currT", fmt_e(t), false)) ||
        (e = write_text(d + "chkpt_stepCount", fmt_i(step), false)) ||
        (e = write_text(d + "chkpt_prevErr", fmt_e(errA[1]), false)) ||
        (e = write_text(d + "chkpt_currErr", fmt_e(errA[0]), false)) ||
        (e = write_text(d + "chkpt_deltaT", fmt_e(newDt), false)) ||
        (e = write_vec(d + "chkpt_momBal_bcR", bcR, false)) ||
        (e = write_vec(d + "chkpt_momBal_bcRShift", shift, false)) ||
        (e = write_vec(d + "chkpt_momBal_bcT", bcT, false)) || (e = write_vec(d + "chkpt_momBal_bcL", bcL, false)) ||
        (e = write_vec(d + "chkpt_momBal_bcB", bcB, false)) || (e = write_vec(d + "chkpt_slip", slip, false)) ||
        (e = write_vec(d + "chkpt_slipVel", V, false)) || (e = write_vec(d + "chkpt_psi", psi, false)) ||
        (e = write_vec(d + "chkpt_sNEff", sN, false)) || (e = write_vec(d + "chkpt_sN", sN, false)) ||
        (e = write_vec(d + "chkpt_tau", tau, false)) || (e = write_vec(d + "chkpt_tauQS", tauQS, false)))
      return e;
    std::vector<double> pre;
    if ((e = d2h(pre, prestress_d, nz)) || (e = write_vec(d + "chkpt_prestress", pre, false))) return e;
    *stop = 1;
  }
  return SC_OK;
}

int QdDriver::load_checkpoint(sc_qd_run* run, int* ckptNumber) {
  *ckptNumber = 0;
  if (!out || out->ckpt <= 0 || out->ckptNumber <= 0) return SC_OK;
  const std::string& d = out->dir;
  double v;
  if (read_value(d + "chkpt_currT", &v)) run->initTime = v;
  if (read_value(d + "chkpt_deltaT", &v)) run->initDt = v;
  if (read_value(d + "chkpt_currErr", &v)) run->currErr = v;
  if (read_value(d + "chkpt_prevErr", &v)) run->prevErr = v;
  SC_HIP(hipStreamSynchronize(ctx->c.stream));
  const size_t nz = (size_t)Nz, ny = (size_t)Ny;
  struct Item {
    const char* name;
    double* dst;
    size_t n;
  };
  const Item items[] = {
      {"chkpt_psi", vars["psi"].var, nz},   {"chkpt_slipVel", V_d, nz},
      {"chkpt_sN", sN_d, nz},               {"chkpt_sNEff", sN_d, nz},  // the root solve's sigma_N is sNEff
      {"chkpt_prestress", prestress_d, nz}, {"chkpt_tauQS", tauQS_d, nz},
      {"chkpt_tau", tau_d, nz},             {"chkpt_momBal_bcRShift", shift_d, nz},
      {"chkpt_momBal_bcR", bcR_d, nz},      {"chkpt_momBal_bcT", bcT_d, ny},
      {"chkpt_momBal_bcB", bcB_d, ny},
  };
  std::vector<double> h;
  bool found;
  int e;
  for (const Item& it : items) {
    if ((e = read_vec(d + it.name, it.n, h, &found))) return e;
    if (found && (e = h2d(it.dst, h))) return e;
  }
  // initiateIntegrand: slip = bcL * faultTypeScale (:391-395)
  if ((e = read_vec(d + "chkpt_momBal_bcL", nz, h, &found))) return e;
  if (found) {
    if ((e = h2d(bcL_d, h))) return e;
    for (double& x : h) x *= desc.faultTypeScale;
    if ((e = h2d(vars["slip"].var, h))) return e;
  }
  *ckptNumber = out->ckptNumber;
  return SC_OK;
}

}  // namespace sc

using sc::fail;

extern "C" int sc_qd_set_output(sc_qd* h, const sc_qd_output* o) {
  if (!h) return fail(SC_ERR_ARG, "null handle");
  return h->q->set_output(o);
}

extern "C" int sc_qd_load_checkpoint(sc_qd* h, sc_qd_run* run, int* ckptNumber) {
  if (!h || !run || !ckptNumber) return fail(SC_ERR_ARG, "null argument");
  return h->q->load_checkpoint(run, ckptNumber);
}
