// C-ABI exports of the host-side discretization tables (no GPU needed).
#include <cstring>
#include <vector>

#include "host_basis.hpp"

extern "C" {

int fmm3dbie_rv_nodes(int norder, double* uvs, double* wts) {
  return fmmhost::rv_nodes(norder, uvs, wts);
}

int fmm3dbie_rv_umat(int norder, double* umat) {
  std::vector<double> u;
  if (fmmhost::rv_umat(norder, u) != 0) return -1;
  std::memcpy(umat, u.data(), sizeof(double) * u.size());
  return 0;
}

void fmm3dbie_koorn_pols(double u, double v, int nmax, double* pols) {
  fmmhost::koorn_pols(u, v, nmax, pols);
}

int fmm3dbie_xg_rule(int deg, double* uvs, double* wts) {
  std::vector<double> a, b;
  int n = fmmhost::xg_rule(deg, a, b);
  if (n < 0) return n;
  if (uvs) std::memcpy(uvs, a.data(), sizeof(double) * a.size());
  if (wts) std::memcpy(wts, b.data(), sizeof(double) * b.size());
  return n;
}

void get_rfacs_(const int64_t* norder, const int64_t* iptype, double* rfac, double* rfac0) {
  (void)iptype;
  fmmhost::rfacs((int)*norder, rfac, rfac0);
}

}  // extern "C"
