// Writers of the reference's output series: one-number ASCII lines
// ("%.15e\n", initiateWriteASCII / PetscViewerASCIIPrintf) and PETSc-binary
// Vec records appended per output step (initiate_appendVecToOutput /
// VecView).  Shared by the QD, power-law and FD mediators.
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "input.hpp"
#include "sc_internal.hpp"

namespace sc {

inline std::string fmt_e(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%.15e\n", v);
  return b;
}

inline std::string fmt_i(int v) {
  char b[32];
  std::snprintf(b, sizeof b, "%d\n", v);
  return b;
}

inline int write_text(const std::string& path, const std::string& text, bool append) {
  std::FILE* f = std::fopen(path.c_str(), append ? "a" : "w");
  if (!f) return fail(SC_ERR_IO, "cannot write " + path);
  const size_t w = std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return w == text.size() ? SC_OK : fail(SC_ERR_IO, "short write to " + path);
}

inline int write_vec(const std::string& path, const std::vector<double>& v, bool append) {
  if (write_petsc_vec(path, v.data(), (int)v.size(), append)) return fail(SC_ERR_IO, "cannot write " + path);
  return SC_OK;
}

inline int d2h(std::vector<double>& h, const double* d, size_t n) {
  h.resize(n);
  SC_HIP(hipMemcpy(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost));
  return SC_OK;
}

}  // namespace sc
