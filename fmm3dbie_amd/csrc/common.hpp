// Shared device-side building blocks for the gfx950 kernels: the orthonormal
// Koornwinder basis on the simplex, the combined-field Green's functions, and the
// fixed-size structures passed to kernels.  Everything here is plain HIP for CDNA4.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trig.hpp"

#define FMM_MAX_ORDER 10
#define FMM_MAX_NP 66  // (10+1)*(10+2)/2
#define OVER4PI 0.07957747154594767

// Recurrence constants of the Koornwinder basis (koornf_init,
// src/tria_routs/koornexps.f90:840-905), filled once by the host.
struct KoornConst {
  double leg_a[FMM_MAX_ORDER + 1];                       // (2k+1)/(k+1)
  double leg_b[FMM_MAX_ORDER + 1];                       // -k/(k+1)
  double jac[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1][3];   // [k][n+1] -> (a, b, -c)
  double rsc[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1];      // [n-k][k] normalisation
};
extern __constant__ KoornConst g_koorn;

// Compile-time recurrence constants of the Koornwinder basis (the same numbers as
// koornf_init, src/tria_routs/koornexps.f90:840-905).  Folded into the unrolled
// recurrence as immediates, so they occupy no registers.
constexpr double cx_sqrt(double x) {
  double r = x > 1.0 ? x : 1.0;
  for (int i = 0; i < 200; i++) {
    const double nr = 0.5 * (r + x / r);
    if (nr == r) break;
    r = nr;
  }
  return r;
}
struct KoornTab {
  double la[FMM_MAX_ORDER + 1], lb[FMM_MAX_ORDER + 1];
  double ja[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1], jb[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1],
      jc[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1];
  double rs[FMM_MAX_ORDER + 1][FMM_MAX_ORDER + 1];
};
constexpr KoornTab make_koorn_tab() {
  KoornTab t{};
  for (int k = 1; k <= FMM_MAX_ORDER; k++) {
    t.la[k] = (2 * k + 1) / (k + 1.0);
    t.lb[k] = -k / (k + 1.0);
  }
  for (int k = 0; k <= FMM_MAX_ORDER; k++) {
    const double beta = 2 * k + 1;
    t.ja[k][1] = (2 + beta) / 2;
    t.jb[k][1] = -beta / 2;
    for (int n = 1; n < FMM_MAX_ORDER; n++) {
      t.ja[k][n + 1] = (2 * n + beta + 1) * (2 * n + beta + 2) / 2 / (n + 1.0) / (n + beta + 1);
      t.jb[k][n + 1] = (-beta * beta) * (2 * n + beta + 1) / 2 / (n + 1.0) / (n + beta + 1) / (2 * n + beta);
      t.jc[k][n + 1] = -(n * (n + beta) * (2 * n + beta + 2) / (n + 1.0) / (n + beta + 1) / (2 * n + beta));
    }
  }
  for (int n = 0; n <= FMM_MAX_ORDER; n++)
    for (int k = 0; k <= n; k++) t.rs[n - k][k] = 1.0 / cx_sqrt(1.0 / (2 * k + 1) / (2 * n + 2));
  return t;
}
constexpr KoornTab kTab = make_koorn_tab();

// P[idx(n,k)] for n=0..N, k=0..n, idx = n(n+1)/2 + k  (koornf_pols, koornexps.f90:909-1016)
template <int N>
__device__ __forceinline__ void koorn_eval(double u, double v, double* P) {
  const double z = 2.0 * u + v - 1.0, y = 1.0 - v, x = 1.0 - 2.0 * v;
  const double y2 = y * y;
  double leg[N + 1];
  leg[0] = 1.0;
  if (N >= 1) leg[1] = z;
#pragma unroll
  for (int k = 1; k <= N - 1; k++) leg[k + 1] = kTab.la[k] * z * leg[k] + kTab.lb[k] * leg[k - 1] * y2;
#pragma unroll
  for (int k = 0; k <= N; k++) {
    double jm1 = 0.0, j0 = 1.0;
#pragma unroll
    for (int m = 0; m <= N - k; m++) {
      const int n = m + k;
      P[n * (n + 1) / 2 + k] = leg[k] * j0 * kTab.rs[m][k];
      if (m + 1 <= N - k) {
        double jn;
        if (m == 0) jn = kTab.ja[k][1] * x + kTab.jb[k][1];
        else jn = (kTab.ja[k][m + 1] * x + kTab.jb[k][m + 1]) * j0 + kTab.jc[k][m + 1] * jm1;
        jm1 = j0;
        j0 = jn;
      }
    }
  }
}


// ---- FP64 elementary functions for the kernel evaluations ------------------------
// sin/cos: sincos_red (trig.hpp), valid for |k r| <= 1.6e6, with no large-argument path:
// the device library's sincos evaluates a large-argument reduction beside the small one
// and selects, which made it most of a node evaluation's instructions.
// 1/sqrt(x) for x > 0: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double t = x * y;
  double e = fma(-t, y, 1.0);
  y = fma(0.5 * y, e, y);
  t = x * y;
  e = fma(-t, y, 1.0);
  return fma(0.5 * y, e, y);
}

// Kernel parameters: Helmholtz zk, alpha, beta (complex) or Laplace alpha, beta (real
// parts used).  u = alpha*S + beta*D.
struct KPar {
  double zkr, zki, ar, ai, br, bi;
};

// h3d_comb (src/kernels/helm_kernels.f90:120-163):
//   (beta*d*(1-ikr)/r^3 + alpha/r) * exp(ikr) / (4 pi),  d = (x-y).n_y
__device__ __forceinline__ void helm_comb(double dx, double dy, double dz, double nx, double ny,
                                          double nz, const KPar& kp, double& vr, double& vi) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double d = dx * nx + dy * ny + dz * nz;
  double s, c;
  sincos_red(kp.zkr * r, s, c);
  double er = c, ei = s;
  if (kp.zki != 0.0) {
    const double dec = exp(-kp.zki * r);
    er *= dec;
    ei *= dec;
  }
  // (1 - i k r) with k = zkr + i zki : (1 + zki r) - i zkr r
  const double ar1 = 1.0 + kp.zki * r, ai1 = -kp.zkr * r;
  const double d3 = d * rinv * rinv * rinv;
  // beta*d3*(1-ikr) + alpha*rinv
  const double tr = (kp.br * ar1 - kp.bi * ai1) * d3 + kp.ar * rinv;
  const double ti = (kp.br * ai1 + kp.bi * ar1) * d3 + kp.ai * rinv;
  vr = (tr * er - ti * ei) * OVER4PI;
  vi = (tr * ei + ti * er) * OVER4PI;
}

// l3d_comb (src/kernels/lap_kernels.f90:123-169): (beta*d/r^3 + alpha/r)/(4 pi)
__device__ __forceinline__ double lap_comb(double dx, double dy, double dz, double nx, double ny,
                                           double nz, const KPar& kp) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rinv = rsqrt_nr(r2);
  const double d = dx * nx + dy * ny + dz * nz;
  return (kp.br * d * rinv * rinv * rinv + kp.ar * rinv) * OVER4PI;
}

// Geometry of one point: srcvals(1:9) = srccoefs * P, normal, |x_u x x_v|.
template <int NP>
__device__ __forceinline__ double geom_eval(const double* __restrict__ sc, const double* P, double* g) {
#pragma unroll
  for (int m = 0; m < 9; m++) g[m] = 0.0;
#pragma unroll
  for (int l = 0; l < NP; l++) {
#pragma unroll
    for (int m = 0; m < 9; m++) g[m] = fma(sc[9 * l + m], P[l], g[m]);
  }
  const double nx = g[4] * g[8] - g[5] * g[7];
  const double ny = g[5] * g[6] - g[3] * g[8];
  const double nz = g[3] * g[7] - g[4] * g[6];
  const double rr = sqrt(nx * nx + ny * ny + nz * nz);
  const double ri = 1.0 / rr;
  g[9] = nx * ri;
  g[10] = ny * ri;
  g[11] = nz * ri;
  return rr;
}


// Basis into an LDS staging row, then geometry from that row:
//   row[l] = P_l(u,v) (koornf_pols ordering),  g[m] = sum_l sc[9l+m] * row[l].
// With the coefficients in global memory (scalar loads) the geometry loop is kept rolled
// so the 9*NP reads stay next to their use instead of being hoisted into registers; with
// the coefficients staged in LDS (SC_LDS) it is unrolled by 3, which overlaps the reads
// of three basis functions.
template <int N>
__device__ __forceinline__ void koorn_row(double u, double v, double* row) {
  const double z = 2.0 * u + v - 1.0, y = 1.0 - v, x = 1.0 - 2.0 * v;
  const double y2 = y * y;
  double lkm1 = 0.0, lk = 1.0;
#pragma unroll
  for (int k = 0; k <= N; k++) {
    if (k == 1) { lkm1 = 1.0; lk = z; }
    if (k >= 2) {
      const double ln = kTab.la[k - 1] * z * lk + kTab.lb[k - 1] * lkm1 * y2;
      lkm1 = lk;
      lk = ln;
    }
    double jm1 = 0.0, j0 = 1.0;
#pragma unroll
    for (int m = 0; m <= N - k; m++) {
      const int n = m + k;
      row[n * (n + 1) / 2 + k] = lk * j0 * kTab.rs[m][k];
      if (m + 1 <= N - k) {
        double jn;
        if (m == 0) jn = kTab.ja[k][1] * x + kTab.jb[k][1];
        else jn = (kTab.ja[k][m + 1] * x + kTab.jb[k][m + 1]) * j0 + kTab.jc[k][m + 1] * jm1;
        jm1 = j0;
        j0 = jn;
      }
    }
  }
}

template <int NP, bool SC_LDS>
__device__ __forceinline__ void geom_from_row(const double* __restrict__ sc, const double* __restrict__ row,
                                              double* g) {
#pragma unroll
  for (int m = 0; m < 9; m++) g[m] = 0.0;
  if (SC_LDS) {
#pragma unroll 3
    for (int l = 0; l < NP; l++) {
      const double p = row[l];
#pragma unroll
      for (int m = 0; m < 9; m++) g[m] = fma(sc[9 * l + m], p, g[m]);
    }
  } else {
#pragma unroll 1
    for (int l = 0; l < NP; l++) {
      const double p = row[l];
#pragma unroll
      for (int m = 0; m < 9; m++) g[m] = fma(sc[9 * l + m], p, g[m]);
    }
  }
}

// Basis into an LDS staging row and geometry from the same values as they are generated
// (no read-back of the row): g[m] = sum_l sc[9l+m] P_l(u,v), summed in generation order.
template <int N>
__device__ __forceinline__ void koorn_geom_fused(double u, double v, const double* __restrict__ sc, double* g,
                                                 double* row) {
  const double z = 2.0 * u + v - 1.0, y = 1.0 - v, x = 1.0 - 2.0 * v;
  const double y2 = y * y;
#pragma unroll
  for (int m = 0; m < 9; m++) g[m] = 0.0;
  double lkm1 = 0.0, lk = 1.0;
#pragma unroll
  for (int k = 0; k <= N; k++) {
    if (k == 1) { lkm1 = 1.0; lk = z; }
    if (k >= 2) {
      const double ln = kTab.la[k - 1] * z * lk + kTab.lb[k - 1] * lkm1 * y2;
      lkm1 = lk;
      lk = ln;
    }
    double jm1 = 0.0, j0 = 1.0;
#pragma unroll
    for (int m = 0; m <= N - k; m++) {
      const int n = m + k;
      const int idx = n * (n + 1) / 2 + k;
      const double val = lk * j0 * kTab.rs[m][k];
      row[idx] = val;
#pragma unroll
      for (int c = 0; c < 9; c++) g[c] = fma(sc[9 * idx + c], val, g[c]);
      if (m + 1 <= N - k) {
        double jn;
        if (m == 0) jn = kTab.ja[k][1] * x + kTab.jb[k][1];
        else jn = (kTab.ja[k][m + 1] * x + kTab.jb[k][m + 1]) * j0 + kTab.jc[k][m + 1] * jm1;
        jm1 = j0;
        j0 = jn;
      }
    }
  }
}

template <int N, bool SC_LDS = false>
__device__ __forceinline__ void koorn_geom_row(double u, double v, const double* __restrict__ sc, double* g,
                                               double* row) {
  koorn_row<N>(u, v, row);
  geom_from_row<(N + 1) * (N + 2) / 2, SC_LDS>(sc, row, g);
}

// normal and area element from g[3..8]; returns |x_u x x_v|, writes unit normal to g[9..11]
__device__ __forceinline__ double geom_finish(double* g) {
  const double nx = g[4] * g[8] - g[5] * g[7];
  const double ny = g[5] * g[6] - g[3] * g[8];
  const double nz = g[3] * g[7] - g[4] * g[6];
  const double rr = sqrt(nx * nx + ny * ny + nz * nz);
  const double ri = 1.0 / rr;
  g[9] = nx * ri;
  g[10] = ny * ri;
  g[11] = nz * ri;
  return rr;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// value of the first active lane (for values known to be equal across the wavefront)
__device__ __forceinline__ double wave_uniform(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    T o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
