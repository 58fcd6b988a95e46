// Oracle sparse-matrix kernels (test infrastructure only; see oracle.hpp).
// Plays the role of the PETSc Mat operations the reference composes its
// operator with (MatMatMult, MatTranspose, MatAXPY, kronConvert
// spmat.cpp:242-306).
#include <algorithm>
#include <cassert>
#include <cmath>
#include <omp.h>

#include "oracle.hpp"

namespace orc {

double Sp1::get(int r, int c) const {
  auto it = m.find(r);
  if (it == m.end()) return 0.0;
  auto jt = it->second.find(c);
  return jt == it->second.end() ? 0.0 : jt->second;
}
void Sp1::eye() {
  m.clear();
  for (int i = 0; i < nr; ++i) m[i][i] = 1.0;
}
void Sp1::scale(double s) {
  for (auto& r : m)
    for (auto& c : r.second) c.second *= s;
}

Csr to_csr(const Sp1& a) {
  Csr o;
  o.nr = a.nr;
  o.nc = a.nc;
  o.rp.assign(a.nr + 1, 0);
  for (int i = 0; i < a.nr; ++i) {
    auto it = a.m.find(i);
    if (it != a.m.end())
      for (auto& c : it->second) {
        if (c.second == 0.0) continue;  // kronConvert drops zeros
        o.ci.push_back(c.first);
        o.v.push_back(c.second);
      }
    o.rp[i + 1] = (int64_t)o.ci.size();
  }
  return o;
}

Csr diag(const std::vector<double>& d) {
  Csr o;
  o.nr = o.nc = (int)d.size();
  o.rp.resize(d.size() + 1);
  o.ci.resize(d.size());
  o.v = d;
  for (size_t i = 0; i <= d.size(); ++i) o.rp[i] = (int64_t)i;
  for (size_t i = 0; i < d.size(); ++i) o.ci[i] = (int)i;
  return o;
}

Csr eye(int n) { return diag(std::vector<double>(n, 1.0)); }

Csr kron(const Csr& a, const Csr& b) {
  Csr o;
  o.nr = a.nr * b.nr;
  o.nc = a.nc * b.nc;
  o.rp.assign((size_t)o.nr + 1, 0);
  // count
  for (int ra = 0; ra < a.nr; ++ra)
    for (int rb = 0; rb < b.nr; ++rb) {
      int64_t cnt = 0;
      for (int64_t p = a.rp[ra]; p < a.rp[ra + 1]; ++p)
        for (int64_t q = b.rp[rb]; q < b.rp[rb + 1]; ++q)
          if (a.v[p] * b.v[q] != 0.0) ++cnt;
      o.rp[(size_t)ra * b.nr + rb + 1] = cnt;
    }
  for (int i = 0; i < o.nr; ++i) o.rp[i + 1] += o.rp[i];
  o.ci.resize(o.rp[o.nr]);
  o.v.resize(o.rp[o.nr]);
#pragma omp parallel for schedule(static)
  for (int ra = 0; ra < a.nr; ++ra)
    for (int rb = 0; rb < b.nr; ++rb) {
      int64_t w = o.rp[(size_t)ra * b.nr + rb];
      // columns of a row are sorted, and within each a-column the b columns
      // are sorted, so output columns come out sorted.
      for (int64_t p = a.rp[ra]; p < a.rp[ra + 1]; ++p)
        for (int64_t q = b.rp[rb]; q < b.rp[rb + 1]; ++q) {
          double val = a.v[p] * b.v[q];
          if (val == 0.0) continue;
          o.ci[w] = a.ci[p] * b.nc + b.ci[q];
          o.v[w] = val;
          ++w;
        }
    }
  return o;
}

Csr matmul(const Csr& a, const Csr& b) {
  assert(a.nc == b.nr);
  Csr o;
  o.nr = a.nr;
  o.nc = b.nc;
  o.rp.assign((size_t)a.nr + 1, 0);
  int nt = omp_get_max_threads();
  std::vector<std::vector<int>> rcols(nt);
  std::vector<std::vector<double>> rvals(nt);
  std::vector<std::vector<int64_t>> rstart(nt);
  std::vector<int> rlo(nt), rhi(nt);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    int nth = omp_get_num_threads();
    int lo = (int)((int64_t)a.nr * t / nth), hi = (int)((int64_t)a.nr * (t + 1) / nth);
    rlo[t] = lo;
    rhi[t] = hi;
    std::vector<double> acc(b.nc, 0.0);
    std::vector<int> mark(b.nc, -1);
    std::vector<int> cols;
    auto& oc = rcols[t];
    auto& ov = rvals[t];
    auto& os = rstart[t];
    for (int i = lo; i < hi; ++i) {
      cols.clear();
      for (int64_t p = a.rp[i]; p < a.rp[i + 1]; ++p) {
        int k = a.ci[p];
        double av = a.v[p];
        for (int64_t q = b.rp[k]; q < b.rp[k + 1]; ++q) {
          int c = b.ci[q];
          if (mark[c] != i) {
            mark[c] = i;
            acc[c] = 0.0;
            cols.push_back(c);
          }
          acc[c] += av * b.v[q];
        }
      }
      std::sort(cols.begin(), cols.end());
      os.push_back((int64_t)oc.size());
      for (int c : cols) {
        oc.push_back(c);
        ov.push_back(acc[c]);
      }
    }
    os.push_back((int64_t)oc.size());
  }
  int64_t off = 0;
  for (int t = 0; t < nt; ++t) {
    for (int i = rlo[t]; i < rhi[t]; ++i) o.rp[i] = off + rstart[t][i - rlo[t]];
    off += (int64_t)rcols[t].size();
  }
  o.rp[a.nr] = off;
  o.ci.reserve(off);
  o.v.reserve(off);
  for (int t = 0; t < nt; ++t) {
    o.ci.insert(o.ci.end(), rcols[t].begin(), rcols[t].end());
    o.v.insert(o.v.end(), rvals[t].begin(), rvals[t].end());
  }
  return o;
}

Csr matmul3(const Csr& a, const Csr& b, const Csr& c) { return matmul(matmul(a, b), c); }

Csr transpose(const Csr& a) {
  Csr o;
  o.nr = a.nc;
  o.nc = a.nr;
  o.rp.assign((size_t)o.nr + 1, 0);
  for (int64_t p = 0; p < a.rp[a.nr]; ++p) o.rp[a.ci[p] + 1]++;
  for (int i = 0; i < o.nr; ++i) o.rp[i + 1] += o.rp[i];
  o.ci.resize(a.rp[a.nr]);
  o.v.resize(a.rp[a.nr]);
  std::vector<int64_t> w(o.rp.begin(), o.rp.end() - 1);
  for (int i = 0; i < a.nr; ++i)
    for (int64_t p = a.rp[i]; p < a.rp[i + 1]; ++p) {
      int64_t q = w[a.ci[p]]++;
      o.ci[q] = i;
      o.v[q] = a.v[p];
    }
  return o;
}

Csr axpy(const Csr& y, double alpha, const Csr& x) {
  assert(y.nr == x.nr && y.nc == x.nc);
  Csr o;
  o.nr = y.nr;
  o.nc = y.nc;
  o.rp.assign((size_t)y.nr + 1, 0);
  o.ci.reserve(y.ci.size() + x.ci.size());
  o.v.reserve(y.ci.size() + x.ci.size());
  for (int i = 0; i < y.nr; ++i) {
    int64_t p = y.rp[i], pe = y.rp[i + 1], q = x.rp[i], qe = x.rp[i + 1];
    while (p < pe || q < qe) {
      if (q >= qe || (p < pe && y.ci[p] < x.ci[q])) {
        o.ci.push_back(y.ci[p]);
        o.v.push_back(y.v[p]);
        ++p;
      } else if (p >= pe || x.ci[q] < y.ci[p]) {
        o.ci.push_back(x.ci[q]);
        o.v.push_back(alpha * x.v[q]);
        ++q;
      } else {
        o.ci.push_back(y.ci[p]);
        o.v.push_back(y.v[p] + alpha * x.v[q]);
        ++p;
        ++q;
      }
    }
    o.rp[i + 1] = (int64_t)o.ci.size();
  }
  return o;
}

void scale(Csr& a, double s) {
  for (auto& x : a.v) x *= s;
}

void spmv(const Csr& a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < a.nr; ++i) {
    double s = 0.0;
    for (int64_t p = a.rp[i]; p < a.rp[i + 1]; ++p) s += a.v[p] * x[a.ci[p]];
    y[i] = s;
  }
}

void spmv_add(const Csr& a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < a.nr; ++i) {
    double s = 0.0;
    for (int64_t p = a.rp[i]; p < a.rp[i + 1]; ++p) s += a.v[p] * x[a.ci[p]];
    y[i] += s;
  }
}

std::vector<double> diagonal(const Csr& a) {
  std::vector<double> d(a.nr, 0.0);
  for (int i = 0; i < a.nr; ++i)
    for (int64_t p = a.rp[i]; p < a.rp[i + 1]; ++p)
      if (a.ci[p] == i) d[i] = a.v[p];
  return d;
}

// ------------------------------------------------------------- solvers ---
static double dot(int n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

PcgStats pcg(const Csr& A, const double* b, double* u, double rtol, int maxit) {
  const int n = A.nr;
  std::vector<double> r(n), z(n), p(n), q(n), dinv = diagonal(A);
  for (auto& x : dinv) x = 1.0 / x;
  spmv(A, u, q.data());
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  double bn = std::sqrt(dot(n, b, b));
  if (bn == 0.0) bn = 1.0;
  PcgStats st;
  double rn = std::sqrt(dot(n, r.data(), r.data()));
  if (rn <= rtol * bn) {
    st.relres = rn / bn;
    return st;
  }
  for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i], p[i] = z[i];
  double rz = dot(n, r.data(), z.data());
  for (int it = 1; it <= maxit; ++it) {
    spmv(A, p.data(), q.data());
    double alpha = rz / dot(n, p.data(), q.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      u[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rn = std::sqrt(dot(n, r.data(), r.data()));
    st.iters = it;
    st.relres = rn / bn;
    if (rn <= rtol * bn) break;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    double rz1 = dot(n, r.data(), z.data());
    double beta = rz1 / rz;
    rz = rz1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return st;
}

int BandChol::factor(const Csr& A) {
  n = A.nr;
  bw = 0;
  for (int i = 0; i < n; ++i)
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) bw = std::max(bw, std::abs(A.ci[p] - i));
  const int w = bw + 1;
  L.assign((size_t)n * w, 0.0);
  // store -A lower band: L[i*w + (j - i + bw)] for j in [i-bw, i]
  for (int i = 0; i < n; ++i)
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
      int j = A.ci[p];
      if (j <= i) L[(size_t)i * w + (j - i + bw)] = -A.v[p];
    }
  for (int j = 0; j < n; ++j) {
    double* Lj = &L[(size_t)j * w];
    double s = Lj[bw];
    for (int k = std::max(0, j - bw); k < j; ++k) s -= Lj[k - j + bw] * Lj[k - j + bw];
    if (!(s > 0.0)) return 1;
    double djj = std::sqrt(s);
    Lj[bw] = djj;
    int iend = std::min(n - 1, j + bw);
#pragma omp parallel for schedule(static) if (iend - j > 64)
    for (int i = j + 1; i <= iend; ++i) {
      double* Li = &L[(size_t)i * w];
      double t = Li[j - i + bw];
      int k0 = std::max(0, i - bw);
      for (int k = k0; k < j; ++k) t -= Li[k - i + bw] * Lj[k - j + bw];
      Li[j - i + bw] = t / djj;
    }
  }
  return 0;
}

void BandChol::solve(const double* b, double* x) const {
  const int w = bw + 1;
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) {  // L y = -b
    double s = -b[i];
    const double* Li = &L[(size_t)i * w];
    for (int k = std::max(0, i - bw); k < i; ++k) s -= Li[k - i + bw] * y[k];
    y[i] = s / Li[bw];
  }
  for (int i = n - 1; i >= 0; --i) {  // L^T x = y
    double s = y[i];
    for (int k = i + 1; k <= std::min(n - 1, i + bw); ++k) s -= L[(size_t)k * w + (i - k + bw)] * x[k];
    x[i] = s / L[(size_t)i * w + bw];
  }
}

}  // namespace orc
