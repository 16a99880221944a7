// Host-side discretization tables (see host_basis.hpp).
#include "host_basis.hpp"

#include <cmath>
#include <cstring>

#include "rules_data.h"

namespace fmmhost {

// Orthonormal Koornwinder polynomials on the simplex (0,0),(1,0),(0,1):
// K_{n,k}(u,v) = P_{n-k}^{(0,2k+1)}(1-2v) (1-v)^k P_k((2u+v-1)/(1-v)) * sqrt((2k+1)(2n+2)),
// ordered (0,0),(1,0),(1,1),(2,0),... as koornf_pols (src/tria_routs/koornexps.f90:909-1016).
void koorn_pols(double u, double v, int nmax, double* pols) {
  double leg[FMM_MAX_ORDER + 24], jac[FMM_MAX_ORDER + 24];
  const double z = 2 * u + v - 1, y = 1 - v, x = 1 - 2 * v;
  leg[0] = 1;
  leg[1] = z;
  for (int k = 1; k <= nmax - 1; k++)
    leg[k + 1] = ((2 * k + 1) / (k + 1.0)) * z * leg[k] + (-k / (k + 1.0)) * leg[k - 1] * y * y;
  for (int k = 0; k <= nmax; k++) {
    const double beta = 2 * k + 1;
    jac[0] = 1;
    if (nmax - k >= 1) jac[1] = ((2 + beta) / 2) * x + (-beta / 2);
    for (int n = 1; n <= nmax - k - 1; n++) {
      const double an = (2 * n + beta + 1) * (2 * n + beta + 2) / 2 / (n + 1.0) / (n + beta + 1);
      const double bn =
          (-beta * beta) * (2 * n + beta + 1) / 2 / (n + 1.0) / (n + beta + 1) / (2 * n + beta);
      const double cn =
          n * (n + beta) * (2 * n + beta + 2) / (n + 1.0) / (n + beta + 1) / (2 * n + beta);
      jac[n + 1] = (an * x + bn) * jac[n] - cn * jac[n - 1];
    }
    for (int m = 0; m <= nmax - k; m++) {
      const int n = m + k;
      pols[n * (n + 1) / 2 + k] = leg[k] * jac[m] * (1.0 / std::sqrt(1.0 / (2 * k + 1) / (2 * n + 2)));
    }
  }
}

int rv_nodes(int n, double* uvs, double* wts) {
  if (n < 0 || n > RULES_RV_MAXORDER) return -1;
  const int a = rules_rv_off[n], b = rules_rv_off[n + 1];
  for (int i = a; i < b; i++) {
    if (uvs) {
      uvs[2 * (i - a)] = rules_rv_u[i];
      uvs[2 * (i - a) + 1] = rules_rv_v[i];
    }
    if (wts) wts[i - a] = rules_rv_w[i];
  }
  return b - a;
}

static int gauss_jordan_inverse(int n, const std::vector<double>& a, std::vector<double>& inv) {
  std::vector<double> m((size_t)n * 2 * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++)
      m[(size_t)i * 2 * n + j] = j < n ? a[i + (size_t)j * n] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(m[(size_t)r * 2 * n + c]) > std::fabs(m[(size_t)p * 2 * n + c])) p = r;
    if (m[(size_t)p * 2 * n + c] == 0) return -1;
    if (p != c)
      for (int j = 0; j < 2 * n; j++) std::swap(m[(size_t)c * 2 * n + j], m[(size_t)p * 2 * n + j]);
    const double piv = m[(size_t)c * 2 * n + c];
    for (int j = 0; j < 2 * n; j++) m[(size_t)c * 2 * n + j] /= piv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = m[(size_t)r * 2 * n + c];
      if (f != 0)
        for (int j = 0; j < 2 * n; j++) m[(size_t)r * 2 * n + j] -= f * m[(size_t)c * 2 * n + j];
    }
  }
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inv[i + (size_t)j * n] = m[(size_t)i * 2 * n + n + j];
  return 0;
}

int rv_umat(int n, std::vector<double>& umat) {
  const int p = npols(n);
  std::vector<double> uvs(2 * p), v((size_t)p * p), pols(p);
  rv_nodes(n, uvs.data(), nullptr);
  for (int i = 0; i < p; i++) {
    koorn_pols(uvs[2 * i], uvs[2 * i + 1], n, pols.data());
    for (int j = 0; j < p; j++) v[i + (size_t)j * p] = pols[j];
  }
  return gauss_jordan_inverse(p, v, umat);
}

int xg_rule(int deg, std::vector<double>& uv, std::vector<double>& w) {
  if (deg < 1 || deg > RULES_XG_MAXDEG) return -1;
  const int a = rules_xg_off[deg], b = rules_xg_off[deg + 1];
  uv.resize(2 * (b - a));
  w.resize(b - a);
  for (int i = a; i < b; i++) {
    uv[2 * (i - a)] = rules_xg_u[i];
    uv[2 * (i - a) + 1] = rules_xg_v[i];
    w[i - a] = rules_xg_w[i];
  }
  return b - a;
}

// Sub-triangles of the simplex tree are (x0,y0) + diag(dx,dy)*(u,v); midpoint
// refinement (gettrichildren, src/tria_routs/triatreerouts.f:313-392) keeps that form.
int xg_unif_nodes(int nlev, int deg, std::vector<double>& uv, std::vector<double>& w) {
  std::vector<double> q, qw;
  const int nq = xg_rule(deg, q, qw);
  if (nq < 0) return -1;
  std::vector<double> tri = {0, 0, 1, 1}, nxt;
  for (int l = 0; l < nlev; l++) {
    nxt.clear();
    for (size_t i = 0; i < tri.size(); i += 4) {
      const double x0 = tri[i], y0 = tri[i + 1], hx = tri[i + 2] / 2, hy = tri[i + 3] / 2;
      const double c[16] = {x0, y0, hx, hy, x0 + hx, y0, hx, hy, x0, y0 + hy, hx, hy, x0 + hx, y0 + hy, -hx, -hy};
      nxt.insert(nxt.end(), c, c + 16);
    }
    tri.swap(nxt);
  }
  const int nt = (int)tri.size() / 4;
  uv.resize(2 * (size_t)nt * nq);
  w.resize((size_t)nt * nq);
  for (int i = 0; i < nt; i++)
    for (int j = 0; j < nq; j++) {
      uv[2 * (i * nq + j)] = tri[4 * i] + tri[4 * i + 2] * q[2 * j];
      uv[2 * (i * nq + j) + 1] = tri[4 * i + 1] + tri[4 * i + 3] * q[2 * j + 1];
      w[i * nq + j] = qw[j] / nt;
    }
  return nt * nq;
}

void gauss_legendre01(int n, double* xs, double* ws) {
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; k++) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double dp = n * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1, p1 = z;
    for (int k = 2; k <= n; k++) {
      const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const double dp = n * (z * p1 - p0) / (z * z - 1);
    xs[n - 1 - i] = 0.5 * (1 + z);
    ws[n - 1 - i] = 1.0 / ((1 - z * z) * dp * dp);
  }
}

void quadparams_adap(double eps, int norder, int* nqorder, double* eps_adap, int* nlev,
                     int* nqorder_f) {
  int iprec = 0;
  if (eps < 0.5e-2) iprec = 1;
  if (eps < 0.5e-3) iprec = 2;
  if (eps < 0.5e-6) iprec = 3;
  if (eps < 0.5e-9) iprec = 4;
  static const int qh[5] = {12, 12, 16, 22, 30}, fh[5] = {8, 10, 16, 22, 30};
  static const double eh[5] = {1.0e-2, 0.2e-2, 0.1e-4, 7.0e-8, 3.0e-10};
  static const int ql[5] = {8, 8, 12, 16, 24}, fl[5] = {4, 6, 12, 20, 26};
  static const double el[5] = {1.0e-2, 0.2e-2, 0.1e-4, 3.0e-8, 2.0e-10};
  *nlev = 1;
  if (norder >= 4) {
    *nqorder = qh[iprec]; *eps_adap = eh[iprec]; *nqorder_f = fh[iprec];
  } else {
    *nqorder = ql[iprec]; *eps_adap = el[iprec]; *nqorder_f = fl[iprec];
  }
}

void rfacs(int norder, double* rfac, double* rfac0) {
  *rfac = 1.25;
  if (norder <= 2) *rfac = 2.75;
  if (norder <= 6 && norder > 2) *rfac = 2.0;
  if (norder > 6) *rfac = 1.25;
  *rfac0 = 1.25;
}

void fill_koorn_const(KoornConst* kc) {
  std::memset(kc, 0, sizeof(*kc));
  for (int k = 1; k <= FMM_MAX_ORDER; k++) {
    kc->leg_a[k] = (2 * k + 1) / (k + 1.0);
    kc->leg_b[k] = -k / (k + 1.0);
  }
  for (int k = 0; k <= FMM_MAX_ORDER; k++) {
    const double beta = 2 * k + 1;
    kc->jac[k][1][0] = (2 + beta) / 2;
    kc->jac[k][1][1] = -beta / 2;
    for (int n = 1; n < FMM_MAX_ORDER; n++) {
      kc->jac[k][n + 1][0] = (2 * n + beta + 1) * (2 * n + beta + 2) / 2 / (n + 1.0) / (n + beta + 1);
      kc->jac[k][n + 1][1] =
          (-beta * beta) * (2 * n + beta + 1) / 2 / (n + 1.0) / (n + beta + 1) / (2 * n + beta);
      kc->jac[k][n + 1][2] =
          -(n * (n + beta) * (2 * n + beta + 2) / (n + 1.0) / (n + beta + 1) / (2 * n + beta));
    }
  }
  for (int n = 0; n <= FMM_MAX_ORDER; n++)
    for (int k = 0; k <= n; k++) kc->rsc[n - k][k] = 1.0 / std::sqrt(1.0 / (2 * k + 1) / (2 * n + 2));
}

void interp_matrix(int norder, int nover, std::vector<double>& imat) {
  const int p = npols(norder), po = npols(nover);
  std::vector<double> umat, uvo(2 * po), pols(p);
  rv_umat(norder, umat);
  rv_nodes(nover, uvo.data(), nullptr);
  imat.assign((size_t)po * p, 0.0);
  for (int i = 0; i < po; i++) {
    koorn_pols(uvo[2 * i], uvo[2 * i + 1], norder, pols.data());
    for (int l = 0; l < p; l++) {
      double s = 0;
      for (int k = 0; k < p; k++) s += pols[k] * umat[k + (size_t)l * p];
      imat[(size_t)i * p + l] = s;
    }
  }
}

// x_{k+1} = a x_k mod 2^48 with a = 33952834046453 held as four 12-bit limbs
// (494, 322, 2508, 2549); the seed limbs are updated in place and the result is the
// seed scaled to [0, 1), regenerated if it rounds to 1.
double dlaran(int iseed[4]) {
  static const int a[4] = {494, 322, 2508, 2549};
  const int base = 4096;
  const double r = 1.0 / base;
  for (;;) {
    // schoolbook product of the limb vectors, keeping the low 48 bits
    int p3 = iseed[3] * a[3];
    int c3 = p3 / base;
    p3 -= base * c3;
    int p2 = c3 + iseed[2] * a[3] + iseed[3] * a[2];
    int c2 = p2 / base;
    p2 -= base * c2;
    int p1 = c2 + iseed[1] * a[3] + iseed[2] * a[2] + iseed[3] * a[1];
    int c1 = p1 / base;
    p1 -= base * c1;
    int p0 = (c1 + iseed[0] * a[3] + iseed[1] * a[2] + iseed[2] * a[1] + iseed[3] * a[0]) % base;
    iseed[0] = p0;
    iseed[1] = p1;
    iseed[2] = p2;
    iseed[3] = p3;
    const double v = r * ((double)p0 + r * ((double)p1 + r * ((double)p2 + r * (double)p3)));
    if (v != 1.0) return v;
  }
}

double hkrand(int seed) {
  int s[4] = {0, 0, 0, (2 * seed + 1) % 4096};
  return dlaran(s);
}

}  // namespace fmmhost
