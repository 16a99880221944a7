// Device GMRES for (zid*I + K) x = rhs (zgmres_guru, src/common/gmres_routs.f90:13-371):
// quadrature-weighted inner products, Givens rotations (zrotmat_gmres,
// rotmat_gmres.f:60-100), stop at |s_{it+1}|/|b| <= eps_gmres, one extra apply for the
// true residual.  Orthogonalisation is classical Gram-Schmidt applied twice (CGS2: the
// same Krylov basis as the reference's modified Gram-Schmidt, two fused kernels per
// iteration instead of 2*it dependent ones).  nrhs systems iterate in lockstep and
// share one batched apply per iteration; each stops at its own convergence.
// T = double (FP64 solve) or float (the FP32 inner solve of iterative refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "apply.hpp"
#include "common.hpp"
#include "gmres.hpp"

namespace {

// Iteration lines as the reference prints them ("iter=", it, errs(it),
// gmres_routs.f90:303), here on stderr and every g_print-th iteration (0: none).
std::atomic<int> g_print{0};

template <typename T>
struct C2 {
  T x, y;
};

// Partial weighted dots: hp[((r*ldh + k)*S + z)] = sum over rows j of slice z of
// w_j conj(V_k[j]) x_r[j]; one block per (k, r, slice z) so that a Krylov basis of a few
// vectors still spreads over the GPU.
template <typename T>
__global__ void k_dots(int64_t n, int64_t chunk, const C2<T>* __restrict__ V, int64_t ldv, int64_t ldvr,
                       const C2<T>* __restrict__ x, int64_t ldx, const double* __restrict__ w,
                       double* __restrict__ hp, int ldh) {
  const int k = blockIdx.x, r = blockIdx.y, z = blockIdx.z, S = gridDim.z;
  const C2<T>* vk = V + r * ldvr + k * ldv;
  const C2<T>* xr = x + r * ldx;
  const int64_t j0 = z * chunk, j1 = j0 + chunk < n ? j0 + chunk : n;
  double sr = 0.0, si = 0.0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const double ar = vk[j].x, ai = -(double)vk[j].y, br = xr[j].x, bi = xr[j].y;
    sr += w[j] * (ar * br - ai * bi);
    si += w[j] * (ar * bi + ai * br);
  }
  __shared__ double shr[256], shi[256];
  shr[threadIdx.x] = sr;
  shi[threadIdx.x] = si;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      shr[threadIdx.x] += shr[threadIdx.x + st];
      shi[threadIdx.x] += shi[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t o = ((int64_t)(r * ldh + k) * S + z) * 2;
    hp[o] = shr[0];
    hp[o + 1] = shi[0];
  }
}

// h[r*ldh + k] = sum over the S slices (one wavefront per (k, r))
__global__ void k_dots_sum(int S, const double* __restrict__ hp, double* __restrict__ h, int ldh) {
  const int k = blockIdx.x, r = blockIdx.y, lane = threadIdx.x;
  const int64_t o = (int64_t)(r * ldh + k) * S;
  double sr = 0.0, si = 0.0;
  for (int z = lane; z < S; z += 64) {
    sr += hp[(o + z) * 2];
    si += hp[(o + z) * 2 + 1];
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (lane == 0) {
    h[2 * (r * ldh + k)] = sr;
    h[2 * (r * ldh + k) + 1] = si;
  }
}

constexpr int64_t DOT_ROWS = 4096;  // rows per dot-product slice
constexpr int DOT_SMAX = 64;

// h[r*ldh + k] = sum_j w_j conj(V_k[j]) x_r[j]  for k < nk (hp: scratch of
// 2 * ldh * nrhs * DOT_SMAX doubles)
template <typename T>
void dots(int64_t n, int nk, const C2<T>* V, int64_t ldv, int64_t ldvr, const C2<T>* x, int64_t ldx,
          const double* w, double* h, int ldh, int nrhs, double* hp, hipStream_t s) {
  int S = (int)((n + DOT_ROWS - 1) / DOT_ROWS);
  S = S < 1 ? 1 : (S > DOT_SMAX ? DOT_SMAX : S);
  const int64_t chunk = (n + S - 1) / S;
  k_dots<T><<<dim3(nk, nrhs, S), 256, 0, s>>>(n, chunk, V, ldv, ldvr, x, ldx, w, hp, ldh);
  k_dots_sum<<<dim3(nk, nrhs), 64, 0, s>>>(S, hp, h, ldh);
}

// x_r -= sum_k h[r,k] V_k
template <typename T>
__global__ void k_sub(int64_t n, int nk, const C2<T>* __restrict__ V, int64_t ldv, int64_t ldvr,
                      C2<T>* __restrict__ x, int64_t ldx, const double* __restrict__ h, int ldh) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= n) return;
  double xr = x[r * ldx + j].x, xi = x[r * ldx + j].y;
  for (int k = 0; k < nk; k++) {
    const double hr = h[2 * (r * ldh + k)], hi = h[2 * (r * ldh + k) + 1];
    const C2<T> v = V[r * ldvr + k * ldv + j];
    xr -= hr * v.x - hi * v.y;
    xi -= hr * v.y + hi * v.x;
  }
  x[r * ldx + j] = {(T)xr, (T)xi};
}

// y_r = x_r * scale_r (complex scale stored as double pairs)
template <typename T>
__global__ void k_scale(int64_t n, const C2<T>* __restrict__ x, int64_t ldx, C2<T>* __restrict__ y, int64_t ldy,
                        const double* __restrict__ sc) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= n) return;
  const double a = sc[2 * r], b = sc[2 * r + 1];
  const C2<T> v = x[r * ldx + j];
  y[r * ldy + j] = {(T)(a * v.x - b * v.y), (T)(a * v.y + b * v.x)};
}

// out_r = sum_k y[r,k] V_k
template <typename T, typename TO>
__global__ void k_combine(int64_t n, int nk, const C2<T>* __restrict__ V, int64_t ldv, int64_t ldvr,
                          const double* __restrict__ y, int ldy, C2<TO>* __restrict__ out, int64_t ldo, int add) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= n) return;
  double sr = 0.0, si = 0.0;
  if (add) {
    sr = out[r * ldo + j].x;
    si = out[r * ldo + j].y;
  }
  for (int k = 0; k < nk; k++) {
    const double yr = y[2 * (r * ldy + k)], yi = y[2 * (r * ldy + k) + 1];
    const C2<T> v = V[r * ldvr + k * ldv + j];
    sr += yr * v.x - yi * v.y;
    si += yr * v.y + yi * v.x;
  }
  out[r * ldo + j] = {(TO)sr, (TO)si};
}

// res_r = rhs_r - (zid*x_r + Kx_r); also returns nothing (norm via k_dots)
template <typename T>
__global__ void k_resid(int64_t n, const C2<T>* __restrict__ rhs, const C2<T>* __restrict__ x,
                        const C2<T>* __restrict__ kx, double zr, double zi, C2<T>* __restrict__ res, int64_t ld) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= n) return;
  const C2<T> a = x[r * ld + j], b = kx[r * ld + j], c = rhs[r * ld + j];
  res[r * ld + j] = {(T)((double)c.x - (zr * a.x - zi * a.y) - b.x), (T)((double)c.y - (zr * a.y + zi * a.x) - b.y)};
}

template <typename A, typename B>
__global__ void k_cvt(int64_t n, const C2<A>* __restrict__ a, C2<B>* __restrict__ b) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) b[j] = {(B)a[j].x, (B)a[j].y};
}

template <typename T>
__global__ void k_add(int64_t n, C2<double>* __restrict__ x, const C2<T>* __restrict__ c) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) {
    x[j].x += (double)c[j].x;
    x[j].y += (double)c[j].y;
  }
}

// x[r] += c[r] for the systems with upd[r] != 0 (grid y = system)
template <typename T>
__global__ void k_add_masked(int64_t n, C2<double>* __restrict__ x, const C2<T>* __restrict__ c,
                             const int* __restrict__ upd) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j < n && upd[r]) {
    x[r * n + j].x += (double)c[r * n + j].x;
    x[r * n + j].y += (double)c[r * n + j].y;
  }
}

// y_r = V_r column k (the Krylov vector of each system, gathered for the batched apply)
template <typename T>
__global__ void k_gather(int64_t n, const C2<T>* __restrict__ V, int64_t ldvr, C2<T>* __restrict__ y) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j < n) y[r * n + j] = V[r * ldvr + j];
}

inline dim3 g2(int64_t n, int r) { return dim3((unsigned)((n + 255) / 256), (unsigned)r); }

// Device buffers of one solve, freed on every return path (early error returns included).
struct DevBufs {
  std::vector<void*> p;
  template <typename T>
  hipError_t alloc(T** out, size_t bytes) {
    *out = nullptr;
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
    if (e == hipSuccess) {
      p.push_back(q);
      *out = (T*)q;
    }
    return e;
  }
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
};

#define GCHK(x)                                                                    \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "fmm3dbie_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

template <typename T>
int gmres_impl(ApplyPlan* P, int mode, const double* zid, const void* rhs_v, void* soln_v, const double* wts,
               int nrhs, int numit, double eps_gmres, int* niter, double* errs, double* rres, hipStream_t s,
               GmresStats* gs, const double* eps_rhs = nullptr) {
  using cd = std::complex<double>;
  const int64_t N = P->npts;
  const int prec = sizeof(T) == 8 ? 64 : 32;
  const C2<T>* rhs = (const C2<T>*)rhs_v;
  C2<T>* soln = (C2<T>*)soln_v;
  const int64_t ldv = N, ldvr = N * (int64_t)(numit + 1);
  C2<T>*V = nullptr, *W = nullptr, *KX = nullptr;
  double* dh = nullptr;
  DevBufs bufs;
  GCHK(bufs.alloc(&V, sizeof(C2<T>) * ldvr * nrhs));
  GCHK(bufs.alloc(&W, sizeof(C2<T>) * N * nrhs));
  GCHK(bufs.alloc(&KX, sizeof(C2<T>) * N * nrhs));
  // dot products of the two Gram-Schmidt passes and the norm, each 2 (numit + 2) nrhs
  // doubles, read back together once per iteration; then the scale factors
  const int64_t LH = 2 * (int64_t)(numit + 2) * nrhs;
  GCHK(bufs.alloc(&dh, sizeof(double) * (3 * LH + 2 * nrhs + LH * DOT_SMAX)));
  double* dsc = dh + 3 * LH;
  double* dhp = dsc + 2 * nrhs;  // partial dot products of the row slices
  std::vector<double> hh(LH), hh2(LH), hh3(3 * LH);
  // rb = ||rhs||_w
  dots<T>(N, 1, rhs, N, N, rhs, N, wts, dh, numit + 2, nrhs, dhp, s);
  GCHK(hipMemcpyAsync(hh.data(), dh, sizeof(double) * 2 * (numit + 2) * nrhs, hipMemcpyDeviceToHost, s));
  GCHK(hipStreamSynchronize(s));
  std::vector<double> rb(nrhs);
  std::vector<double> scv(2 * nrhs);
  for (int r = 0; r < nrhs; r++) {
    rb[r] = std::sqrt(hh[2 * (r * (numit + 2))]);
    scv[2 * r] = 1.0 / rb[r];
    scv[2 * r + 1] = 0.0;
  }
  GCHK(hipMemcpyAsync(dsc, scv.data(), sizeof(double) * 2 * nrhs, hipMemcpyHostToDevice, s));
  k_scale<T><<<g2(N, nrhs), 256, 0, s>>>(N, rhs, N, V, ldvr, dsc);
  std::vector<cd> H((size_t)numit * numit * nrhs), cs(numit * nrhs), sn(numit * nrhs), sv((numit + 1) * nrhs);
  for (int r = 0; r < nrhs; r++) sv[r * (numit + 1)] = rb[r];
  std::vector<int> done(nrhs, 0), its(nrhs, numit);
  const cd z(zid[0], zid[1]);
  int napply = 0;
  auto Hm = [&](int r, int i, int j) -> cd& { return H[(size_t)r * numit * numit + i + (size_t)j * numit]; };
  int it = 0;
  for (it = 1; it <= numit; it++) {
    // W = K V_{it-1}  (all systems in one batched apply; V_{it-1} of system r at r*ldvr)
    const C2<T>* vin = V + (it - 1) * ldv;
    if (nrhs > 1) {
      k_gather<T><<<g2(N, nrhs), 256, 0, s>>>(N, vin, ldvr, KX);
      vin = KX;
    }
    int rc = apply_plan_run(P, prec, prec == 32 ? APPLY_MASKED : mode, vin, W, nrhs, s, nullptr);
    if (rc) return rc;
    napply++;
    // CGS2 (two passes, each on the device with its own dot products), then the norm;
    // one read-back per iteration
    for (int pass = 0; pass < 2; pass++) {
      dots<T>(N, it, V, ldv, ldvr, W, N, wts, dh + pass * LH, numit + 2, nrhs, dhp, s);
      k_sub<T><<<g2(N, nrhs), 256, 0, s>>>(N, it, V, ldv, ldvr, W, N, dh + pass * LH, numit + 2);
    }
    dots<T>(N, 1, W, N, N, W, N, wts, dh + 2 * LH, numit + 2, nrhs, dhp, s);
    GCHK(hipMemcpyAsync(hh3.data(), dh, sizeof(double) * 3 * LH, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    std::vector<double> htot(LH, 0.0);
    for (int r = 0; r < nrhs; r++)
      for (int k = 0; k < it; k++)
        for (int c = 0; c < 2; c++) {
          const int64_t o = 2 * (r * (numit + 2) + k) + c;
          htot[o] = hh3[o] + hh3[LH + o];
        }
    std::copy(hh3.begin() + 2 * LH, hh3.end(), hh2.begin());
    for (int r = 0; r < nrhs; r++) {
      const double wn = std::sqrt(hh2[2 * (r * (numit + 2))]);
      scv[2 * r] = wn > 0 ? 1.0 / wn : 0.0;
      scv[2 * r + 1] = 0.0;
      if (done[r]) continue;
      for (int k = 0; k < it; k++) Hm(r, k, it - 1) = cd(htot[2 * (r * (numit + 2) + k)], htot[2 * (r * (numit + 2) + k) + 1]);
      Hm(r, it - 1, it - 1) += z;
      for (int k = 1; k <= it - 1; k++) {
        const cd t2 = cs[r * numit + k - 1] * Hm(r, k - 1, it - 1) + std::conj(sn[r * numit + k - 1]) * Hm(r, k, it - 1);
        Hm(r, k, it - 1) = -sn[r * numit + k - 1] * Hm(r, k - 1, it - 1) + cs[r * numit + k - 1] * Hm(r, k, it - 1);
        Hm(r, k - 1, it - 1) = t2;
      }
      const cd a = Hm(r, it - 1, it - 1), b = wn;
      cd c, sg;
      if (a == cd(0, 0)) {
        sg = 1;
        c = 0;
      } else {
        const cd temp = b / a;
        c = 1.0 / std::sqrt(1.0 + std::norm(temp));
        sg = temp * c;
      }
      cs[r * numit + it - 1] = c;
      sn[r * numit + it - 1] = sg;
      Hm(r, it - 1, it - 1) = c * Hm(r, it - 1, it - 1) + std::conj(sg) * wn;
      sv[r * (numit + 1) + it] = -sg * sv[r * (numit + 1) + it - 1];
      sv[r * (numit + 1) + it - 1] = c * sv[r * (numit + 1) + it - 1];
      const double err = std::abs(sv[r * (numit + 1) + it]) / rb[r];
      errs[r * numit + it - 1] = err;
      if (err <= (eps_rhs ? eps_rhs[r] : eps_gmres) || it == numit) {
        done[r] = 1;
        its[r] = it;
      }
    }
    const int every = g_print.load();
    if (every > 0 && it % every == 0) {
      double mx = 0.0;
      for (int r = 0; r < nrhs; r++) mx = std::max(mx, errs[r * numit + it - 1]);
      fprintf(stderr, "gmres (%s, %d systems) iter= %d %.6e\n", sizeof(T) == 4 ? "fp32" : "fp64", nrhs, it, mx);
    }
    GCHK(hipMemcpyAsync(dsc, scv.data(), sizeof(double) * 2 * nrhs, hipMemcpyHostToDevice, s));
    if (it < numit) k_scale<T><<<g2(N, nrhs), 256, 0, s>>>(N, W, N, V + it * ldv, ldvr, dsc);
    bool all = true;
    for (int r = 0; r < nrhs; r++) all = all && done[r];
    if (all) break;
  }
  // back substitution per system and solution
  std::vector<double> yv(2 * (numit + 2) * nrhs, 0.0);
  for (int r = 0; r < nrhs; r++) {
    const int m = its[r];
    std::vector<cd> y(m);
    for (int j = 1; j <= m; j++) {
      const int ii = m - j + 1;
      cd acc = sv[r * (numit + 1) + ii - 1];
      for (int l = ii + 1; l <= m; l++) acc -= Hm(r, ii - 1, l - 1) * y[l - 1];
      y[ii - 1] = acc / Hm(r, ii - 1, ii - 1);
    }
    for (int k = 0; k < m; k++) {
      yv[2 * (r * (numit + 2) + k)] = y[k].real();
      yv[2 * (r * (numit + 2) + k) + 1] = y[k].imag();
    }
    niter[r] = m;
  }
  GCHK(hipMemcpyAsync(dh, yv.data(), sizeof(double) * 2 * (numit + 2) * nrhs, hipMemcpyHostToDevice, s));
  int mmax = 0;
  for (int r = 0; r < nrhs; r++) mmax = its[r] > mmax ? its[r] : mmax;
  k_combine<T, T><<<g2(N, nrhs), 256, 0, s>>>(N, mmax, V, ldv, ldvr, dh, numit + 2, soln, N, 0);
  // true residual
  if (rres) {
    int rc = apply_plan_run(P, prec, prec == 32 ? APPLY_MASKED : mode, soln, KX, nrhs, s, nullptr);
    if (rc) return rc;
    napply++;
    k_resid<T><<<g2(N, nrhs), 256, 0, s>>>(N, rhs, soln, KX, zid[0], zid[1], W, N);
    dots<T>(N, 1, W, N, N, W, N, wts, dh, numit + 2, nrhs, dhp, s);
    GCHK(hipMemcpyAsync(hh.data(), dh, sizeof(double) * 2 * (numit + 2) * nrhs, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    for (int r = 0; r < nrhs; r++) rres[r] = std::sqrt(hh[2 * (r * (numit + 2))]) / rb[r];
  }
  GCHK(hipStreamSynchronize(s));
  if (gs) gs->napply += napply;
  return 0;
}

}  // namespace

void gmres_set_print(int every) { g_print.store(every > 0 ? every : 0); }

int gmres_device(ApplyPlan* P, int prec, int mode, const double* zid, const void* rhs, void* soln,
                 const double* wts, int nrhs, int numit, double eps_gmres, int* niter, double* errs, double* rres,
                 hipStream_t s, GmresStats* gs) {
  if (P->ntarg != P->npts) {
    fprintf(stderr, "fmm3dbie_amd: GMRES needs the on-surface operator (plan targets = discretization points)\n");
    return -6;
  }
  if (prec != 32 && prec != 64) {
    fprintf(stderr, "fmm3dbie_amd: GMRES precision %d (must be 32 or 64)\n", prec);
    return -6;
  }
  if (nrhs < 1 || nrhs > 64 || numit < 1) return -6;
  if (prec == 64)
    return gmres_impl<double>(P, mode, zid, rhs, soln, wts, nrhs, numit, eps_gmres, niter, errs, rres, s, gs);
  return gmres_impl<float>(P, mode, zid, rhs, soln, wts, nrhs, numit, eps_gmres, niter, errs, rres, s, gs);
}

// Mixed-precision iterative refinement: inner FP32/complex64 GMRES on the masked FP32
// operator, outer residual r = rhs - (zid + K64) x in FP64, until ||r||_w/||rhs||_w <= tol.
int gmres_ir_device(ApplyPlan* P, int mode, const double* zid, const void* rhs_v, void* soln_v,
                    const double* wts, int nrhs, int numit, double eps_inner, double tol, int max_outer,
                    int* outer_out, int* inner_total, double* rres, hipStream_t s, GmresStats* gs) {
  if (P->ntarg != P->npts) {
    fprintf(stderr, "fmm3dbie_amd: GMRES needs the on-surface operator (plan targets = discretization points)\n");
    return -6;
  }
  if (nrhs < 1 || nrhs > 64 || numit < 1 || max_outer < 1) return -6;
  const int64_t N = P->npts;
  const C2<double>* rhs = (const C2<double>*)rhs_v;
  C2<double>* x = (C2<double>*)soln_v;
  C2<double>*r64 = nullptr, *kx = nullptr;
  C2<float>*r32 = nullptr, *c32 = nullptr;
  double* dh = nullptr;
  DevBufs bufs;
  GCHK(bufs.alloc(&r64, sizeof(C2<double>) * N * nrhs));
  GCHK(bufs.alloc(&kx, sizeof(C2<double>) * N * nrhs));
  GCHK(bufs.alloc(&r32, sizeof(C2<float>) * N * nrhs));
  GCHK(bufs.alloc(&c32, sizeof(C2<float>) * N * nrhs));
  GCHK(bufs.alloc(&dh, sizeof(double) * 4 * nrhs * (1 + DOT_SMAX)));
  double* dhp = dh + 4 * nrhs;  // partial dot products of the row slices
  GCHK(hipMemsetAsync(x, 0, sizeof(C2<double>) * N * nrhs, s));
  GCHK(hipMemcpyAsync(r64, rhs, sizeof(C2<double>) * N * nrhs, hipMemcpyDeviceToDevice, s));
  std::vector<double> hb(2 * nrhs * 2), hr(2 * nrhs * 2);
  dots<double>(N, 1, rhs, N, N, rhs, N, wts, dh, 2, nrhs, dhp, s);
  GCHK(hipMemcpyAsync(hb.data(), dh, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, s));
  GCHK(hipStreamSynchronize(s));
  std::vector<int> nit(nrhs), upd(nrhs, 1);
  std::vector<double> errs((size_t)numit * nrhs), eps_r(nrhs, eps_inner), cur(nrhs, 1.0);
  int* dupd = nullptr;
  GCHK(bufs.alloc(&dupd, sizeof(int) * nrhs));
  int outer = 0, inner = 0;
  double worst = 1.0;
  for (outer = 1; outer <= max_outer; outer++) {
    // inner tolerance per system: eps_inner for the first solve, afterwards just the
    // reduction still needed (with a factor 0.3 of margin) -- the inner residual is
    // relative to the current FP64 residual; systems already at tol are left unchanged
    for (int r = 0; r < nrhs; r++) {
      upd[r] = cur[r] > tol ? 1 : 0;
      const double need = 0.3 * tol / cur[r];
      eps_r[r] = outer == 1 ? eps_inner : (upd[r] ? std::min(0.5, std::max(eps_inner, need)) : 0.5);
    }
    GCHK(hipMemcpyAsync(dupd, upd.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
    k_cvt<double, float><<<(unsigned)((N * nrhs + 255) / 256), 256, 0, s>>>(N * nrhs, r64, r32);
    int rc = gmres_impl<float>(P, mode, zid, r32, c32, wts, nrhs, numit, eps_inner, nit.data(), errs.data(),
                               nullptr, s, gs, eps_r.data());
    if (rc) return rc;
    int mx = 0;
    for (int r = 0; r < nrhs; r++) mx = nit[r] > mx ? nit[r] : mx;
    inner += mx;
    k_add_masked<float><<<g2(N, nrhs), 256, 0, s>>>(N, x, c32, dupd);
    rc = apply_plan_run(P, 64, mode, x, kx, nrhs, s, nullptr);
    if (rc) return rc;
    if (gs) gs->napply64++;
    k_resid<double><<<g2(N, nrhs), 256, 0, s>>>(N, rhs, x, kx, zid[0], zid[1], r64, N);
    dots<double>(N, 1, r64, N, N, r64, N, wts, dh, 2, nrhs, dhp, s);
    GCHK(hipMemcpyAsync(hr.data(), dh, sizeof(double) * 4 * nrhs, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    worst = 0.0;
    for (int r = 0; r < nrhs; r++) {
      const double e = std::sqrt(hr[2 * (r * 2)]) / std::sqrt(hb[2 * (r * 2)]);
      if (rres) rres[r] = e;
      cur[r] = e;
      worst = e > worst ? e : worst;
    }
    if (g_print.load() > 0)
      fprintf(stderr, "gmres refinement %d: %d inner iterations, fp64 residual %.3e\n", outer, mx, worst);
    if (worst <= tol) break;
  }
  if (outer > max_outer) outer = max_outer;
  if (outer_out) *outer_out = outer;
  if (inner_total) *inner_total = inner;
  // the buffers are freed on return; the stream's work on them is complete (the last
  // read-back above synchronised it)
  return 0;
}
