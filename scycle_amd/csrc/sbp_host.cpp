// Host-side set-up of the matrix-free SBP-SAT operator: 1-D stencil tables,
// coordinate metric, device upload.  Replaces the one-time matrix assembly of
// SbpOps_m_varGrid::computeMatrices (sbpOps_m_varGrid.cpp:280-307) and
// SbpOps_m_constGrid::computeMatrices (sbpOps_m_constGrid.cpp:261-286).
#include <cmath>
#include <cstring>
#include <mutex>

#include "fdiag.hpp"
#include "sc_internal.hpp"

namespace sc {

namespace {
thread_local std::string g_err;

struct Band {
  int n;
  std::vector<double>& v;
  Band(int n_, std::vector<double>& v_) : n(n_), v(v_) { v.assign((size_t)n * kBand, 0.0); }
  void set(int i, int c, double val) {
    int o = c - i + 4;
    if (o < 0 || o >= kBand || i < 0 || i >= n || c < 0 || c >= n) return;  // never hit for N > 8
    v[(size_t)i * kBand + o] = val;
  }
  double get(int i, int c) const {
    int o = c - i + 4;
    if (o < 0 || o >= kBand || i < 0 || i >= n || c < 0 || c >= n) return 0.0;
    return v[(size_t)i * kBand + o];
  }
};

void transpose_band(int n, const std::vector<double>& a, std::vector<double>& at) {
  at.assign((size_t)n * kBand, 0.0);
  for (int r = 0; r < n; ++r)
    for (int o = 0; o < kBand; ++o) {
      int c = r + o - 4;
      if (c < 0 || c >= n) continue;
      double val = a[(size_t)r * kBand + o];
      if (val == 0.0) continue;
      at[(size_t)c * kBand + (r - c + 4)] = val;  // (c, r) of the transpose
    }
}
}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* last_error() { return g_err.c_str(); }

// Coefficients follow sbp_Spmat / sbp_Spmat2 / sbp_Spmat4 (spmat.cpp:310-627)
// and the penalty constants of the SbpOps constructors
// (sbpOps_m_varGrid.cpp:30-43).
int build_dir(int order, int n, double h, bool compatible, Dir1D& d) {
  if (order != 2 && order != 4) return fail(SC_ERR_ARG, "order must be 2 or 4");
  if (order == 4 && n <= 8) return fail(SC_ERR_ARG, "order 4 needs more than 8 points per direction");
  if (order == 2 && n <= 2) return fail(SC_ERR_ARG, "order 2 needs more than 2 points per direction");
  d.n = n;
  d.order = order;
  d.h = h;
  const double scale = 1.0 / h;
  d.H.assign(n, 1.0);
  if (order == 2) {
    d.H[0] = d.H[n - 1] = 0.5;
  } else {
    const double hb[4] = {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0};
    for (int i = 0; i < 4; ++i) {
      d.H[i] = hb[i];
      d.H[n - 1 - i] = hb[i];
    }
  }
  for (auto& x : d.H) x *= 1 / scale;
  d.Hinv.resize(n);
  for (int i = 0; i < n; ++i) d.Hinv[i] = 1 / d.H[i];

  Band D1(n, d.D1), DR(n, d.DR), D4(n, d.D4);
  std::vector<double> bs0(kBand, 0.0), bsn(kBand, 0.0);  // BS rows
  d.wR.assign(n, 0.0);
  d.w4.assign(n, 0.0);
  auto hi = [&](int i) { return d.Hinv[i]; };
  if (order == 2) {
    D1.set(0, 0, -1.0 * scale);
    D1.set(0, 1, scale);
    for (int i = 1; i < n - 1; ++i) {
      D1.set(i, i - 1, -0.5 * scale);
      D1.set(i, i + 1, 0.5 * scale);
    }
    D1.set(n - 1, n - 1, scale);
    D1.set(n - 1, n - 2, -1 * scale);
    if (!compatible) {
      bs0[0] = -D1.get(0, 0);
      bs0[1] = -D1.get(0, 1);
      bsn[7] = D1.get(n - 1, n - 2);
      bsn[8] = D1.get(n - 1, n - 1);
    } else {
      bs0[0] = 1.5 * scale; bs0[1] = -2.0 * scale; bs0[2] = 0.5 * scale;
      bsn[6] = 0.5 * scale; bsn[7] = -2.0 * scale; bsn[8] = 1.5 * scale;
      D1.set(0, 0, -bs0[0]); D1.set(0, 1, -bs0[1]); D1.set(0, 2, -bs0[2]);
      D1.set(n - 1, n - 3, bsn[6]); D1.set(n - 1, n - 2, bsn[7]); D1.set(n - 1, n - 1, bsn[8]);
    }
    // 2nd-order remainder D2, C2 (spmat.cpp:494-544)
    const double s2 = scale * scale;
    DR.set(0, 0, s2); DR.set(0, 1, -2.0 * s2); DR.set(0, 2, s2);
    for (int i = 1; i < n - 1; ++i) {
      DR.set(i, i - 1, s2); DR.set(i, i, -2.0 * s2); DR.set(i, i + 1, s2);
    }
    DR.set(n - 1, n - 3, s2); DR.set(n - 1, n - 2, -2.0 * s2); DR.set(n - 1, n - 1, s2);
    const double w = 0.25 * std::pow(h, 3);
    for (int i = 1; i < n - 1; ++i) d.wR[i] = 1.0 * w;  // C2 = I with zero corners
    d.alphaD = -4.0 / h;
    d.h11 = 0.5 * h;
  } else {
    for (int i = 4; i < n - 4; ++i) {
      D1.set(i, i - 2, 1.0 / 12.0 * hi(i));
      D1.set(i, i - 1, -2.0 / 3.0 * hi(i));
      D1.set(i, i + 1, 2.0 / 3.0 * hi(i));
      D1.set(i, i + 2, -1.0 / 12.0 * hi(i));
    }
    D1.set(0, 0, -1.0 / 2.0 * hi(0)); D1.set(0, 1, 59.0 / 96.0 * hi(0));
    D1.set(0, 2, -1.0 / 12.0 * hi(0)); D1.set(0, 3, -1.0 / 32.0 * hi(0));
    D1.set(1, 0, -59.0 / 96.0 * hi(1)); D1.set(1, 2, 59.0 / 96.0 * hi(1));
    D1.set(2, 0, 1.0 / 12.0 * hi(2)); D1.set(2, 1, -59.0 / 96.0 * hi(2));
    D1.set(2, 3, 59.0 / 96.0 * hi(2)); D1.set(2, 4, -1.0 / 12.0 * hi(2));
    D1.set(3, 0, 1.0 / 32.0 * hi(3)); D1.set(3, 2, -59.0 / 96.0 * hi(3));
    D1.set(3, 4, 2.0 / 3.0 * hi(3)); D1.set(3, 5, -1.0 / 12.0 * hi(3));
    D1.set(n - 1, n - 1, 1.0 / 2.0 * hi(n - 1)); D1.set(n - 1, n - 2, -59.0 / 96.0 * hi(n - 1));
    D1.set(n - 1, n - 3, 1.0 / 12.0 * hi(n - 1)); D1.set(n - 1, n - 4, 1.0 / 32.0 * hi(n - 1));
    D1.set(n - 2, n - 1, 59.0 / 96.0 * hi(n - 2)); D1.set(n - 2, n - 3, -59.0 / 96.0 * hi(n - 2));
    D1.set(n - 3, n - 1, -1.0 / 12.0 * hi(n - 3)); D1.set(n - 3, n - 2, 59.0 / 96.0 * hi(n - 3));
    D1.set(n - 3, n - 4, -59.0 / 96.0 * hi(n - 3)); D1.set(n - 3, n - 5, 1.0 / 12.0 * hi(n - 3));
    D1.set(n - 4, n - 1, -1.0 / 32.0 * hi(n - 4)); D1.set(n - 4, n - 3, 59.0 / 96.0 * hi(n - 4));
    D1.set(n - 4, n - 5, -2.0 / 3.0 * hi(n - 4)); D1.set(n - 4, n - 6, 1.0 / 12.0 * hi(n - 4));
    if (!compatible) {
      bs0[0] = 24.0 / 17.0 * scale; bs0[1] = -59.0 / 34.0 * scale;
      bs0[2] = 4.0 / 17.0 * scale; bs0[3] = 3.0 / 34.0 * scale;
      bsn[8] = 24.0 / 17.0 * scale; bsn[7] = -59.0 / 34.0 * scale;
      bsn[6] = 4.0 / 17.0 * scale; bsn[5] = 3.0 / 34.0 * scale;
    } else {
      // the reference leaves the last row of the compatible BS unscaled
      // (spmat.cpp:463) -- reproduced.
      bs0[0] = 11.0 / 6.0 * scale; bs0[1] = -3.0 * scale; bs0[2] = 1.5 * scale; bs0[3] = -1.0 / 3.0 * scale;
      bsn[8] = 11.0 / 6.0; bsn[7] = -3.0; bsn[6] = 1.5; bsn[5] = -1.0 / 3.0;
      D1.set(0, 0, -bs0[0]); D1.set(0, 1, -bs0[1]); D1.set(0, 2, -bs0[2]); D1.set(0, 3, -bs0[3]);
      D1.set(n - 1, n - 4, bsn[5]); D1.set(n - 1, n - 2, bsn[7]); D1.set(n - 1, n - 1, bsn[8]);
    }
    // 4th-order remainder D3, D4, C3, C4 (spmat.cpp:547-627)
    const double r2[6] = {-185893.0 / 301051.0, 79000249461.0 / 54642863857.0,
                          -33235054191.0 / 54642863857.0, -36887526683.0 / 54642863857.0,
                          26183621850.0 / 54642863857.0, -4386.0 / 181507.0};
    for (int r = 0; r < 2; ++r) {
      DR.set(r, 0, -1); DR.set(r, 1, 3); DR.set(r, 2, -3); DR.set(r, 3, 1);
    }
    for (int c = 0; c < 6; ++c) DR.set(2, c, r2[c]);
    for (int i = 3; i < n - 4; ++i) {
      DR.set(i, i - 1, -1.0); DR.set(i, i, 3); DR.set(i, i + 1, -3); DR.set(i, i + 2, 1.0);
    }
    for (int c = 0; c < 6; ++c) DR.set(n - 3, n - 1 - c, -r2[c]);
    for (int r = n - 2; r < n; ++r) {
      DR.set(r, n - 4, -1); DR.set(r, n - 3, 3); DR.set(r, n - 2, -3); DR.set(r, n - 1, 1);
    }
    const double d4r[5] = {1, -4, 6, -4, 1};
    for (int c = 0; c < 5; ++c) { D4.set(0, c, d4r[c]); D4.set(1, c, d4r[c]); }
    for (int i = 2; i < n - 2; ++i)
      for (int c = 0; c < 5; ++c) D4.set(i, i - 2 + c, d4r[c]);
    for (int c = 0; c < 5; ++c) { D4.set(n - 2, n - 5 + c, d4r[c]); D4.set(n - 1, n - 5 + c, d4r[c]); }
    std::vector<double> C3(n, 1.0), C4(n, 1.0);
    C3[0] = 0; C3[1] = 0;
    C3[2] = 163928591571.0 / 53268010936.0;
    C3[3] = 189284.0 / 185893.0;
    C3[n - 5] = C3[3];
    C3[n - 4] = 0;
    C3[n - 3] = C3[2];
    C3[n - 2] = 0;
    C3[n - 1] = 0;
    C4[0] = 0; C4[1] = 0;
    C4[2] = 1644330.0 / 301051.0;
    C4[3] = 156114.0 / 181507.0;
    C4[n - 4] = C4[3];
    C4[n - 3] = C4[2];
    C4[n - 2] = 0;
    C4[n - 1] = 0;
    for (int i = 0; i < n; ++i) {
      d.wR[i] = C3[i] * (1.0 / h / 18.0);
      d.w4[i] = C4[i] * (1.0 / h / 144.0);
    }
    d.alphaD = 2.0 * -48.0 / 17.0 / h;
    d.h11 = 17.0 / 48.0 * h;
  }
  d.BS0 = bs0;
  d.BSN = bsn;
  transpose_band(n, d.DR, d.DRT);
  transpose_band(n, d.D4, d.D4T);
  return SC_OK;
}

namespace {

// Size (in elements) of one direction's table block on the device.
size_t dir_elems(int n) { return (size_t)n * (4 + 5 * kBand + 2) + 2 * kBand; }

// false when the interior row of a line with interior tiles does not have the
// structure the tile kernel's factored form assumes (it always has for the
// operators built above; the check guards that form against a table change)
template <typename T>
bool pack_dir(const Dir1D& d, const std::vector<double>& met, const std::vector<double>& imet,
              T* base, DevDir<T>& dd, int order) {
  const int n = d.n;
  std::vector<T> h;
  h.reserve(dir_elems(n));
  auto put = [&](const std::vector<double>& v) {
    size_t off = h.size();
    for (double x : v) h.push_back((T)x);
    return off;
  };
  size_t oH = put(d.H), oHi = put(d.Hinv), oD1 = put(d.D1), oDR = put(d.DR), oDRT = put(d.DRT),
         oD4 = put(d.D4), oD4T = put(d.D4T), owR = put(d.wR), ow4 = put(d.w4), oB0 = put(d.BS0),
         oBN = put(d.BSN), oM = put(met), oIM = put(imet);
  (void)hipMemcpy(base, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  dd.n = n;
  dd.H = base + oH; dd.Hinv = base + oHi; dd.D1 = base + oD1; dd.DR = base + oDR;
  dd.DRT = base + oDRT; dd.D4 = base + oD4; dd.D4T = base + oD4T; dd.wR = base + owR;
  dd.w4 = base + ow4; dd.BS0 = base + oB0; dd.BSN = base + oBN; dd.met = base + oM;
  dd.imet = base + oIM;
  dd.alphaD = (T)d.alphaD;
  const int m = n / 2;
  for (int s = 0; s < 5; ++s) {
    dd.d1[s] = (T)d.D1[(size_t)m * kBand + s + 2];
    dd.dr[s] = (T)d.DR[(size_t)m * kBand + s + 2];
    dd.d4[s] = (T)d.D4[(size_t)m * kBand + s + 2];
    dd.drt[s] = (T)d.DRT[(size_t)m * kBand + s + 2];
    dd.d4t[s] = (T)d.D4T[(size_t)m * kBand + s + 2];
  }
  dd.wr = (T)d.wR[m];
  dd.w4c = (T)d.w4[m];
  dd.Hc = (T)d.H[m];
  auto at = [&](const std::vector<double>& t, int s) { return t[(size_t)m * kBand + s + 2]; };
  const double a = at(d.D1, 3), b = at(d.D1, 4), hc = d.H[m];
  dd.ia = (T)a;
  dd.ib = (T)b;
  dd.Ha = (T)(hc * a);
  dd.Hb = (T)(hc * b);
  bool ok = at(d.D1, 0) == -b && at(d.D1, 1) == -a && at(d.D1, 2) == 0.0;
  if (order == 4) {
    const double r[5] = {0, -1, 3, -3, 1}, rt[5] = {1, -3, 3, -1, 0}, q[5] = {1, -4, 6, -4, 1};
    for (int s = 0; s < 5; ++s)
      ok = ok && at(d.DR, s) == r[s] && at(d.DRT, s) == rt[s] && at(d.D4, s) == q[s] && at(d.D4T, s) == q[s];
    dd.g3s = (T)(0.5 * d.wR[m]);
    dd.g4s = (T)d.w4[m];
    dd.rts = (T)1;
  } else {
    const double s2 = at(d.DR, 1);
    const double r[5] = {0, 1, -2, 1, 0};
    for (int s = 0; s < 5; ++s) ok = ok && at(d.DR, s) == r[s] * s2 && at(d.DRT, s) == r[s] * s2;
    ok = ok && b == 0.0;
    dd.g3s = (T)(d.wR[m] * s2);
    dd.g4s = (T)0;
    dd.rts = (T)s2;
  }
  return ok || n <= 2 * kFrame;
}

}  // namespace

namespace {
template <typename T>
void fill_closure(const Dir1D& d, int hi, ClosureTab<T>& t) {
  const int n = d.n;
  std::memset(&t, 0, sizeof(t));
  if (n <= 2 * kFrame) return;  // the frame kernel for such lines does not use the tables
  const int mb = hi ? n - 12 : 0, jb = hi ? n - kFrame : 0;
  for (int i = 0; i < 12 * kBand; ++i) {
    t.fD1[i] = (T)d.D1[(size_t)mb * kBand + i];
    t.fDR[i] = (T)d.DR[(size_t)mb * kBand + i];
    t.fD4[i] = (T)d.D4[(size_t)mb * kBand + i];
  }
  for (int i = 0; i < kFrame * kBand; ++i) {
    t.oD1[i] = (T)d.D1[(size_t)jb * kBand + i];
    t.oDRT[i] = (T)d.DRT[(size_t)jb * kBand + i];
    t.oD4T[i] = (T)d.D4T[(size_t)jb * kBand + i];
  }
  for (int i = 0; i < 12; ++i) {
    t.wR[i] = (T)d.wR[mb + i];
    t.w4[i] = (T)d.w4[mb + i];
  }
  for (int i = 0; i < kFrame; ++i) t.H[i] = (T)d.H[jb + i];
  for (int i = 0; i < kBand; ++i) t.BS[i] = (T)(hi ? d.BSN[i] : d.BS0[i]);
}

template <typename T>
int upload_closure(const Ctx& c, void*& dev, DevOp<T>& op) {
  ClosureTab<T> h[4];
  fill_closure(c.ty, 0, h[0]);
  fill_closure(c.ty, 1, h[1]);
  fill_closure(c.tz, 0, h[2]);
  fill_closure(c.tz, 1, h[3]);
  if (!dev) SC_HIP(hipMalloc(&dev, sizeof(h)));
  SC_HIP(hipMemcpy(dev, h, sizeof(h), hipMemcpyHostToDevice));
  op.ctab = (const ClosureTab<T>*)dev;
  return SC_OK;
}
}  // namespace

int ctx_upload_tables(Ctx& c) {
  const size_t ey = dir_elems(c.d.Ny), ez = dir_elems(c.d.Nz);
  if (!c.tab64) SC_HIP(hipMalloc(&c.tab64, (ey + ez) * sizeof(double)));
  if (!c.tab32) SC_HIP(hipMalloc(&c.tab32, (ey + ez) * sizeof(float)));
  const int ord = c.d.order;
  bool ok = pack_dir<double>(c.ty, c.yq, c.qy, (double*)c.tab64, c.op64.y, ord);
  ok = pack_dir<double>(c.tz, c.zr, c.rz, (double*)c.tab64 + ey, c.op64.z, ord) && ok;
  ok = pack_dir<float>(c.ty, c.yq, c.qy, (float*)c.tab32, c.op32.y, ord) && ok;
  ok = pack_dir<float>(c.tz, c.zr, c.rz, (float*)c.tab32 + ey, c.op32.z, ord) && ok;
  if (!ok) return fail(SC_ERR_ARG, "SBP interior stencil of an unexpected form");
  // the tile kernel addresses the grid with 32-bit byte offsets
  if ((double)c.d.Ny * c.d.Nz * sizeof(double) >= 2147483648.0)
    return fail(SC_ERR_ARG, "grids of 2^28 nodes or more are not supported");
  int e = upload_closure<double>(c, c.ctab64, c.op64);
  if (!e) e = upload_closure<float>(c, c.ctab32, c.op32);
  if (e) return e;
  SC_HIP(hipGetLastError());
  for (auto* op : {&c.op64.Ny, &c.op32.Ny}) *op = c.d.Ny;
  c.op64.Nz = c.op32.Nz = c.d.Nz;
  c.op64.order = c.op32.order = c.d.order;
  c.op64.doY = c.op32.doY = (c.d.laplace == 0 || c.d.laplace == 1);
  c.op64.doZ = c.op32.doZ = (c.d.laplace == 0 || c.d.laplace == 2);
  c.op64.bcR = c.op32.bcR = c.d.bcR;
  c.op64.bcT = c.op32.bcT = c.d.bcT;
  c.op64.bcL = c.op32.bcL = c.d.bcL;
  c.op64.bcB = c.op32.bcB = c.d.bcB;
  return SC_OK;
}

int ctx_upload_mu(Ctx& c) {
  const size_t N = (size_t)c.N;
  if (!c.mu64) SC_HIP(hipMalloc(&c.mu64, N * sizeof(double)));
  if (!c.mu32) SC_HIP(hipMalloc(&c.mu32, N * sizeof(float)));
  SC_HIP(hipMemcpy(c.mu64, c.mu_h.data(), N * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float> m32(N);
  for (size_t i = 0; i < N; ++i) m32[i] = (float)c.mu_h[i];
  SC_HIP(hipMemcpy(c.mu32, m32.data(), N * sizeof(float), hipMemcpyHostToDevice));
  c.op64.mu = c.mu64;
  c.op32.mu = c.mu32;
  return SC_OK;
}

int ctx_compute_diag(Ctx& c) {
  const size_t N = (size_t)c.N;
  if (!c.diag64) SC_HIP(hipMalloc(&c.diag64, N * sizeof(double)));
  if (!c.dinv64) SC_HIP(hipMalloc(&c.dinv64, N * sizeof(double)));
  if (!c.dinv32) SC_HIP(hipMalloc(&c.dinv32, N * sizeof(float)));
  int e = launch_diag(c, c.diag64, c.stream);
  if (!e) e = ctx_build_frame(c);
  if (e) return e;
  SC_HIP(hipStreamSynchronize(c.stream));
  return SC_OK;
}

}  // namespace sc

// ---------------------------------------------------------------- C ABI ---
using sc::fail;

extern "C" const char* sc_last_error(void) { return sc::last_error(); }

extern "C" int sc_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 2;
  return SC_OK;
}

extern "C" int sc_ctx_create(const sc_grid_desc* desc, const double* y_h, const double* z_h,
                             const double* mu_h, sc_ctx** out) {
  if (!desc || !mu_h || !out) return fail(SC_ERR_ARG, "sc_ctx_create: null argument");
  const sc_grid_desc d = *desc;
  if (d.Ny < 3 || d.Nz < 3) return fail(SC_ERR_ARG, "Ny and Nz must be >= 3 (2-D operator)");
  if (!(d.Ly > 0) || !(d.Lz > 0)) return fail(SC_ERR_ARG, "Ly and Lz must be > 0");
  if (d.laplace < 0 || d.laplace > 2) return fail(SC_ERR_ARG, "laplace must be 0 (yz), 1 (y) or 2 (z)");
  for (int b : {d.bcR, d.bcT, d.bcL, d.bcB})
    if (b != SC_DIRICHLET && b != SC_NEUMANN) return fail(SC_ERR_ARG, "BC type must be Dirichlet or Neumann");
  if ((long long)d.Ny * d.Nz > (1LL << 31) / 8) return fail(SC_ERR_ARG, "grid too large");
  auto* h = new sc_ctx;
  sc::Ctx& c = h->c;
  c.d = d;
  c.N = d.Ny * d.Nz;
  const int Ny = d.Ny, Nz = d.Nz, N = c.N;
  double hy, hz;
  if (d.var_grid) {
    hy = 1. / (Ny - 1.);
    hz = 1. / (Nz - 1.);
  } else {
    hy = d.Ly / (Ny - 1.);
    hz = d.Lz / (Nz - 1.);
  }
  int e = sc::build_dir(d.order, Ny, hy, d.compat != 0, c.ty);
  if (!e) e = sc::build_dir(d.order, Nz, hz, d.compat != 0, c.tz);
  // the kernels apply SAT terms on the kFrame nodes next to each boundary only
  for (const sc::Dir1D* t : {&c.ty, &c.tz}) {
    if (e) break;
    for (int i = 0; i < sc::kBand; ++i)
      if ((i >= sc::kFrame && t->BS0[i] != 0.0) || (i < sc::kBand - sc::kFrame && t->BSN[i] != 0.0))
        e = fail(SC_ERR_ARG, "BS has entries beyond the frame width");
  }
  if (e) {
    delete h;
    return e;
  }
  c.mu_h.assign(mu_h, mu_h + N);
  c.yq.assign(Ny, 1.0);
  c.zr.assign(Nz, 1.0);
  if (d.var_grid) {
    // y = Ly*q, z = Lz*r unless given (Domain::setFields, domain.cpp:352-430)
    std::vector<double> ycol(Ny), zrow(Nz);
    const double dq = 1.0 / (Ny - 1.0), dr = 1.0 / (Nz - 1.0);
    for (int j = 0; j < Ny; ++j) ycol[j] = y_h ? y_h[(size_t)j * Nz] : d.Ly * (dq * j);
    for (int k = 0; k < Nz; ++k) zrow[k] = z_h ? z_h[k] : d.Lz * (dr * k);
    if (y_h || z_h) {
      for (int j = 0; j < Ny; ++j)
        for (int k = 0; k < Nz; ++k) {
          size_t I = (size_t)j * Nz + k;
          if ((y_h && y_h[I] != ycol[j]) || (z_h && z_h[I] != zrow[k])) {
            delete h;
            return fail(SC_ERR_ARG, "grid is not a tensor product (y must depend on j only, z on k only)");
          }
        }
    }
    // yq = D1q * y, zr = D1r * z (constructJacobian, sbpOps_m_varGrid.cpp:356-368)
    for (int j = 0; j < Ny; ++j) {
      double s = 0;
      for (int o = 0; o < sc::kBand; ++o) {
        int cidx = j + o - 4;
        if (cidx < 0 || cidx >= Ny) continue;
        s += c.ty.D1[(size_t)j * sc::kBand + o] * ycol[cidx];
      }
      c.yq[j] = s;
    }
    for (int k = 0; k < Nz; ++k) {
      double s = 0;
      for (int o = 0; o < sc::kBand; ++o) {
        int cidx = k + o - 4;
        if (cidx < 0 || cidx >= Nz) continue;
        s += c.tz.D1[(size_t)k * sc::kBand + o] * zrow[cidx];
      }
      c.zr[k] = s;
    }
  }
  c.qy.resize(Ny);
  c.rz.resize(Nz);
  for (int j = 0; j < Ny; ++j) c.qy[j] = 1.0 / c.yq[j];
  for (int k = 0; k < Nz; ++k) c.rz[k] = 1.0 / c.zr[k];
  for (int j = 0; j < Ny; ++j)
    if (!(c.yq[j] > 0) || !std::isfinite(c.qy[j])) {
      delete h;
      return fail(SC_ERR_ARG, "y coordinate must increase with j");
    }
  for (int k = 0; k < Nz; ++k)
    if (!(c.zr[k] > 0) || !std::isfinite(c.rz[k])) {
      delete h;
      return fail(SC_ERR_ARG, "z coordinate must increase with k");
    }
  // A blocking stream: host-side hipMemcpy / hipMemset on the null stream are
  // ordered with the work queued here, so uploads and clears never race it.
  hipError_t he = hipStreamCreateWithFlags(&c.stream, hipStreamDefault);
  if (he != hipSuccess) {
    delete h;
    return fail(SC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  e = sc::ctx_upload_tables(c);
  if (!e) e = sc::ctx_upload_mu(c);
  if (!e) e = sc::ctx_compute_diag(c);
  if (e) {
    sc_ctx_destroy(h);
    return e;
  }
  *out = h;
  return SC_OK;
}

extern "C" int sc_ctx_destroy(sc_ctx* h) {
  if (!h) return SC_OK;
  sc::Ctx& c = h->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  void* bufs[] = {c.mu64, c.mu32, c.tab64, c.tab32, c.ctab64, c.ctab32, c.fcoef64, c.fcoef32, c.diag64,
                  c.dinv64, c.dinv32, c.r64,
                  c.p64, c.q64, c.r32, c.p32, c.q32, c.d32, c.partials, c.scal, c.z64,
                  c.zold64};
  delete c.fd;
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (double* p : c.projY)
    if (p) (void)hipFree(p);
  for (double* p : c.projZ)
    if (p) (void)hipFree(p);
  for (double* p : c.projU)
    if (p) (void)hipFree(p);
  for (double* p : c.projT)
    if (p) (void)hipFree(p);
  for (double* p : c.projZU)
    if (p) (void)hipFree(p);
  if (c.projW) (void)hipFree(c.projW);
  if (c.scal_h) (void)hipHostFree(c.scal_h);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return SC_OK;
}

extern "C" int sc_ctx_sync(sc_ctx* h) {
  if (!h) return fail(SC_ERR_ARG, "null ctx");
  SC_HIP(hipStreamSynchronize(h->c.stream));
  return SC_OK;
}

extern "C" int sc_ctx_stream(sc_ctx* h, void** stream) {
  if (!h || !stream) return fail(SC_ERR_ARG, "null argument");
  *stream = (void*)h->c.stream;
  return SC_OK;
}

extern "C" int sc_update_var_coeff(sc_ctx* h, const double* mu_h) {
  if (!h || !mu_h) return fail(SC_ERR_ARG, "null argument");
  sc::Ctx& c = h->c;
  SC_HIP(hipStreamSynchronize(c.stream));
  c.mu_h.assign(mu_h, mu_h + c.N);
  c.fd_stale = true;
  ++c.opGen;
  c.projCount = c.projNext = 0;
  int e = sc::ctx_upload_mu(c);
  if (!e) e = sc::ctx_compute_diag(c);
  return e;
}

extern "C" int sc_change_bc_types(sc_ctx* h, int bcR, int bcT, int bcL, int bcB) {
  if (!h) return fail(SC_ERR_ARG, "null ctx");
  for (int b : {bcR, bcT, bcL, bcB})
    if (b != SC_DIRICHLET && b != SC_NEUMANN) return fail(SC_ERR_ARG, "BC type must be Dirichlet or Neumann");
  sc::Ctx& c = h->c;
  SC_HIP(hipStreamSynchronize(c.stream));
  c.d.bcR = bcR;
  c.d.bcT = bcT;
  c.d.bcL = bcL;
  c.d.bcB = bcB;
  c.op64.bcR = c.op32.bcR = bcR;
  c.op64.bcT = c.op32.bcT = bcT;
  c.op64.bcL = c.op32.bcL = bcL;
  c.op64.bcB = c.op32.bcB = bcB;
  c.projCount = c.projNext = 0;
  ++c.opGen;
  return sc::ctx_compute_diag(c);
}

extern "C" int sc_geth11(sc_ctx* h, double* h11y, double* h11z) {
  if (!h) return fail(SC_ERR_ARG, "null ctx");
  if (h11y) *h11y = h->c.ty.h11;
  if (h11z) *h11z = h->c.tz.h11;
  return SC_OK;
}

extern "C" int sc_get_metric(sc_ctx* h, double* yq_h, double* zr_h) {
  if (!h) return fail(SC_ERR_ARG, "null ctx");
  if (yq_h) std::memcpy(yq_h, h->c.yq.data(), sizeof(double) * h->c.d.Ny);
  if (zr_h) std::memcpy(zr_h, h->c.zr.data(), sizeof(double) * h->c.d.Nz);
  return SC_OK;
}

extern "C" int sc_malloc(void** p, unsigned long long bytes) {
  if (!p) return fail(SC_ERR_ARG, "null argument");
  SC_HIP(hipMalloc(p, bytes));
  return SC_OK;
}
extern "C" int sc_free(void* p) {
  SC_HIP(hipFree(p));
  return SC_OK;
}
extern "C" int sc_memcpy_h2d(void* dst, const void* src, unsigned long long bytes) {
  SC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SC_OK;
}
extern "C" int sc_memcpy_d2h(void* dst, const void* src, unsigned long long bytes) {
  SC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SC_OK;
}
