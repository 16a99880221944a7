// Near-field quadrature generation on gfx950 (getnearquad_{helm,lap}_comb_dir*).
//
// Pipeline (all on one stream, inputs resident in HBM):
//   1. k_patch_prep   : per patch centroid/radius of the vertex images
//                       (get_centroid_rads, surf_routs.f90:195-302) and the adaptive
//                       tolerance epsp = eps_adap*rsc^(1/4) (ggq-quads.f:329-337).
//   2. k_classify     : per CSR entry -> self / adaptive (|x-c| <= rfac0*rad) / fixed
//                       (ggq-quads.f:341-365, get_near_far_split_pt surf_routs.f90:1653).
//   3. k_fixed_mfma   : one workgroup per source patch; the patch's oversampled rule
//      (k_fixed)        geometry is built once in LDS and every far-near target of the
//                       patch is integrated against it, 16 targets per wavefront with
//                       the moments on FP64 MFMA (ctriaints_wnodes, triaintrouts.f:67-271
//                       + umat^T, ggq-quads.f:412).
//   4. k_tree_cache_mfma : node geometry of the refinement tree's depths 0..3 or 0..4
//                       per patch, 16 patches at a time on FP64 MFMA.
//   5. k_adaptive     : persistent wavefronts pull (target, patch) pairs from a queue
//                       and run the 1->4 refinement of triaadap_main
//                       (triaintrouts.f:1269-1464) depth-first with a per-wave stack in
//                       HBM; the nodes of a step's four children are evaluated up
//                       to 64 at a time (55 at order 6 and above, where a rule of 28
//                       to 55 nodes takes one chunk per child).  With learned order selection each pair runs
//                       the XG rule of its predicted class.
//   6. k_self         : one wavefront per on-surface target; a polar/sinh product rule
//                       on the three triangles meeting at the target replaces the GGQ
//                       rule of zget_ggq_self_quad_pt (ggq-quads.f:933-1122).
// Moments are formed by staging kernel values f_q and basis rows P_j(q) in LDS and
// letting lane j sum column j, then weights = umat^T * moments.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "host_basis.hpp"
#include "learned.hpp"
#include "quad.hpp"

__constant__ KoornConst g_koorn;
extern __shared__ double fmm_dyn_lds[];


namespace {

constexpr int WAVE = 64;
constexpr int STACK_CAP = 72;
constexpr int NLMAX = 20;
// The second stream is used on surfaces of up to SIDE_MAX_PATCHES patches: measured 7-8%
// faster on cube spheres of 768-1728 patches of orders 4-6, equal on a 1536-patch order-6
// torus, and 6-8% slower on a 12288-patch order-6 surface, where the adaptive kernel alone
// keeps the GPU busy and the kernels beside it only take its slots.
constexpr int64_t SIDE_MAX_PATCHES = 4096;

// Gauss-Legendre rules on [0, 1] of the self rule: radial (r) and sinh-parameter (t).
struct SelfGL {
  double xr[16], wr[16];
  double xt[16], wt[16];
};
__constant__ SelfGL g_self_gl;

template <bool HELM>
__device__ __forceinline__ void kern_eval(const double* tg, const double* s, const KPar& kp, double& vr,
                                          double& vi) {
  if (HELM) {
    helm_comb(tg[0] - s[0], tg[1] - s[1], tg[2] - s[2], s[9], s[10], s[11], kp, vr, vi);
  } else {
    vr = lap_comb(tg[0] - s[0], tg[1] - s[1], tg[2] - s[2], s[9], s[10], s[11], kp);
    vi = 0.0;
  }
}

// ------------------------------------------------------------------ 1. patch prep
template <int N>
__global__ void k_patch_prep(int64_t npatches, const int64_t* __restrict__ ixyzs,
                             const double* __restrict__ srccoefs, const double* __restrict__ srcvals,
                             const double* __restrict__ rvw, double eps_adap, double* __restrict__ cms,
                             double* __restrict__ rads, double* __restrict__ epsp, double* __restrict__ area) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (ip >= npatches) return;
  const int64_t i0 = ixyzs[ip] - 1;
  const double* sc = srccoefs + 9 * i0;
  double P[NP], xyz[3][3];
  const double vu[3] = {0.0, 1.0, 0.0}, vv[3] = {0.0, 0.0, 1.0};
  for (int k = 0; k < 3; k++) {
    koorn_eval<N>(vu[k], vv[k], P);
    for (int m = 0; m < 3; m++) {
      double s = 0.0;
      for (int l = 0; l < NP; l++) s += sc[9 * l + m] * P[l];
      xyz[k][m] = s;
    }
  }
  double c[3];
  for (int m = 0; m < 3; m++) c[m] = (xyz[0][m] + xyz[1][m] + xyz[2][m]) / 3;
  double r = 0.0;
  for (int k = 0; k < 3; k++) {
    const double d = (c[0] - xyz[k][0]) * (c[0] - xyz[k][0]) + (c[1] - xyz[k][1]) * (c[1] - xyz[k][1]) +
                     (c[2] - xyz[k][2]) * (c[2] - xyz[k][2]);
    if (d >= r) r = d;
  }
  cms[3 * ip] = c[0];
  cms[3 * ip + 1] = c[1];
  cms[3 * ip + 2] = c[2];
  rads[ip] = sqrt(r);
  double rsc = 0.0, ar = 0.0;
  for (int i = 0; i < NP; i++) {
    const double* s = srcvals + 12 * (i0 + i);
    const double nx = s[4] * s[8] - s[5] * s[7], ny = s[5] * s[6] - s[3] * s[8],
                 nz = s[3] * s[7] - s[4] * s[6];
    const double rr = (nx * nx + ny * ny + nz * nz) * rvw[i] * rvw[i];
    if (i == 0 || rr < rsc) rsc = rr;
    ar += sqrt(nx * nx + ny * ny + nz * nz) * rvw[i];
  }
  epsp[ip] = eps_adap * pow(rsc, 0.25);
  area[ip] = ar;
}

// ------------------------------------------------------------------ 2. classify
__global__ void k_classify(int64_t ntarg, int64_t ndtarg, const double* __restrict__ targs,
                           const int64_t* __restrict__ ipatch_id, const int64_t* __restrict__ row_ptr,
                           const int64_t* __restrict__ col_ind, const double* __restrict__ cms,
                           const double* __restrict__ rads, double rfac0, int* __restrict__ fixed_cnt,
                           int2* __restrict__ adap_list, int* __restrict__ adap_n,
                           int2* __restrict__ self_list, int* __restrict__ self_n,
                           int* __restrict__ pair_class) {
  // Pass 1 classifies each pair of the target's row (0 self, 1 adaptive, 2 fixed rule)
  // into pair_class and counts them; the lanes' counts are scanned within the wavefront,
  // so that one atomic per wavefront and list reserves the wavefront's slots (the list
  // counters are single addresses that every wavefront updates); pass 2 writes the
  // entries.  blockDim must be a multiple of 64.
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int64_t j0 = 0, j1 = 0, pid = -1;
  double tx = 0.0, ty = 0.0, tz = 0.0;
  if (t < ntarg) {
    tx = targs[ndtarg * t];
    ty = targs[ndtarg * t + 1];
    tz = targs[ndtarg * t + 2];
    pid = ipatch_id ? ipatch_id[t] : -1;
    j0 = row_ptr[t] - 1;
    j1 = row_ptr[t + 1] - 1;
  }
  int ns = 0, na = 0;
  for (int64_t j = j0; j < j1; j++) {
    const int64_t ip = col_ind[j] - 1;
    int cls;
    if (pid == ip + 1) {
      cls = 0;
      ns++;
    } else {
      const double dx = tx - cms[3 * ip], dy = ty - cms[3 * ip + 1], dz = tz - cms[3 * ip + 2];
      const double rr = rfac0 * rads[ip];
      if (dx * dx + dy * dy + dz * dz <= rr * rr) {
        cls = 1;
        na++;
      } else {
        cls = 2;
        atomicAdd(fixed_cnt + ip, 1);
      }
    }
    pair_class[j] = cls;
  }
  int ps = ns, pa = na;  // inclusive scans over the lanes
  for (int d = 1; d < 64; d <<= 1) {
    const int us = __shfl_up(ps, d, 64), ua = __shfl_up(pa, d, 64);
    if (lane >= d) {
      ps += us;
      pa += ua;
    }
  }
  int bs = 0, ba = 0;
  if (lane == 63) {
    if (ps > 0) bs = atomicAdd(self_n, ps);
    if (pa > 0) ba = atomicAdd(adap_n, pa);
  }
  int ks = __shfl(bs, 63, 64) + ps - ns, ka = __shfl(ba, 63, 64) + pa - na;
  for (int64_t j = j0; j < j1; j++) {
    const int cls = pair_class[j];
    if (cls == 0) self_list[ks++] = make_int2((int)j, (int)t);
    else if (cls == 1) adap_list[ka++] = make_int2((int)j, (int)t);
  }
}

__global__ void __launch_bounds__(1024) k_scan(int64_t n, const int* __restrict__ cnt, int* __restrict__ off) {
  // Single-workgroup exclusive scan, off[n] = total: each thread sums a contiguous chunk,
  // the chunk sums are scanned within each wavefront (shuffles) and across the
  // wavefronts (LDS), then each thread writes its chunk's offsets.
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nw = (int)(blockDim.x >> 6);
  const int64_t chunk = (n + blockDim.x - 1) / blockDim.x;
  const int64_t a = tid * chunk, b = a + chunk < n ? a + chunk : n;
  int s = 0;
  for (int64_t i = a; i < b; i++) s += cnt[i];
  int v = s;  // inclusive scan over the lanes
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (w == 0) {
    int x = lane < nw ? wsum[lane] : 0;
    for (int d = 1; d < 16; d <<= 1) {
      const int u = __shfl_up(x, d, 64);
      if (lane >= d) x += u;
    }
    if (lane < nw) wsum[lane] = x;
  }
  __syncthreads();
  int acc = v - s + (w > 0 ? wsum[w - 1] : 0);
  if (tid == (int)blockDim.x - 1) off[n] = acc + s;
  for (int64_t i = a; i < b; i++) {
    off[i] = acc;
    acc += cnt[i];
  }
}

__global__ void k_scatter_fixed(int64_t ntarg, const int64_t* __restrict__ row_ptr,
                                const int64_t* __restrict__ col_ind, const int* __restrict__ pair_class,
                                const int* __restrict__ off, int* __restrict__ fill,
                                int2* __restrict__ fixed_list) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntarg) return;
  for (int64_t j = row_ptr[t] - 1; j < row_ptr[t + 1] - 1; j++) {
    if (pair_class[j] != 2) continue;
    const int64_t ip = col_ind[j] - 1;
    const int k = atomicAdd(fill + ip, 1);
    fixed_list[off[ip] + k] = make_int2((int)j, (int)t);
  }
}

// ------------------------------------------------------------------ 3. fixed rule
// Bq[q*NP + j] = w_q * sum_l P_l(q) umat(l,j); Pq[q*NP + l] = P_l(q).
template <int N, bool HELM>
__global__ void __launch_bounds__(128) k_fixed(int nf, const double* __restrict__ Pq, const double* __restrict__ Bq,
                                               const int64_t* __restrict__ ixyzs,
                                               const double* __restrict__ srccoefs, int64_t ndtarg,
                                               const double* __restrict__ targs,
                                               const int64_t* __restrict__ iquad, const int* __restrict__ off,
                                               const int2* __restrict__ fixed_list, KPar kp,
                                               double* __restrict__ wnear) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  double* lds = fmm_dyn_lds;
  const int ip = blockIdx.x;
  const int a = off[ip], b = off[ip + 1];
  if (a == b) return;
  double* sc = lds;              // 9*NP
  double* node = lds + 9 * NP;   // nf * 7 : xyz, n, |J|
  const double* gsc = srccoefs + 9 * (ixyzs[ip] - 1);
  for (int i = threadIdx.x; i < 9 * NP; i += blockDim.x) sc[i] = gsc[i];
  __syncthreads();
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    double P[NP], g[12];
#pragma unroll
    for (int l = 0; l < NP; l++) P[l] = Pq[q * NP + l];
    const double jac = geom_eval<NP>(sc, P, g);
    double* o = node + 7 * q;
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    o[3] = g[9]; o[4] = g[10]; o[5] = g[11];
    o[6] = jac;
  }
  __syncthreads();
  for (int k = a + threadIdx.x; k < b; k += blockDim.x) {
    const int2 jt = fixed_list[k];
    const double* tg = targs + ndtarg * (int64_t)jt.y;
    const double tx = tg[0], ty = tg[1], tz = tg[2];
    double ar[NP], ai[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) { ar[j] = 0.0; ai[j] = 0.0; }
#pragma unroll 1
    for (int q = 0; q < nf; q++) {
      const double* o = node + 7 * q;
      double vr, vi;
      if (HELM) {
        helm_comb(tx - o[0], ty - o[1], tz - o[2], o[3], o[4], o[5], kp, vr, vi);
      } else {
        vr = lap_comb(tx - o[0], ty - o[1], tz - o[2], o[3], o[4], o[5], kp);
        vi = 0.0;
      }
      vr *= o[6];
      vi *= o[6];
      const double* bq = Bq + q * NP;
#pragma unroll
      for (int j = 0; j < NP; j++) {
        const double bb = bq[j];
        ar[j] = fma(vr, bb, ar[j]);
        if (HELM) ai[j] = fma(vi, bb, ai[j]);
      }
    }
    const int64_t q0 = iquad[jt.x] - 1;
    if (HELM) {
#pragma unroll
      for (int j = 0; j < NP; j++) {
        wnear[2 * (q0 + j)] = ar[j];
        wnear[2 * (q0 + j) + 1] = ai[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < NP; j++) wnear[q0 + j] = ar[j];
    }
  }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Fixed rule on the matrix cores: the weights of a patch's far-near targets are the
// product W = F B of the kernel matrix F (targets x nf nodes, F_tq = K(x_t, y_q) |J_q|)
// and the patch-independent table B (nf x NP, B_qj = w_q sum_l P_l(q) umat(l, j)), i.e.
// ctriaints_wnodes (triaintrouts.f:67-271) followed by umat^T (ggq-quads.f:412).  A
// workgroup takes one source patch: its rule geometry is built once into LDS, then each
// wavefront takes 16 targets at a time.  In a step of 4 nodes, lane l evaluates F for
// target l % 16 at node l / 16 (64 independent kernel evaluations) and the FP64 MFMA
// 16x16x4 multiplies that 16x4 block of F (real and imaginary parts) by the 4x16 block
// of B (basis functions 16 ct .. 16 ct + 15); accumulator element r of lane l holds
// target l / 16 + 4 r, basis function l % 16 + 16 ct.
template <int N, bool HELM>
__global__ void __launch_bounds__(256) k_fixed_mfma(int nf, const double* __restrict__ Pq,
                                                    const double* __restrict__ Bq, const int64_t* __restrict__ ixyzs,
                                                    const double* __restrict__ srccoefs, int64_t ndtarg,
                                                    const double* __restrict__ targs,
                                                    const int64_t* __restrict__ iquad, const int* __restrict__ off,
                                                    const int2* __restrict__ fixed_list, KPar kp,
                                                    double* __restrict__ wnear) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  constexpr int CT = (NP + 15) / 16;
  constexpr int NPART = HELM ? 2 : 1;
  double* lds = fmm_dyn_lds;
  const int ip = blockIdx.x;
  const int a = off[ip], b = off[ip + 1];
  if (a == b) return;
  double* sc = lds;             // 9*NP
  double* node = lds + 9 * NP;  // nf * 8: xyz, n, |J|, pad
  const double* gsc = srccoefs + 9 * (ixyzs[ip] - 1);
  for (int i = threadIdx.x; i < 9 * NP; i += blockDim.x) sc[i] = gsc[i];
  __syncthreads();
  for (int q = threadIdx.x; q < nf; q += blockDim.x) {
    double P[NP], g[12];
#pragma unroll
    for (int l = 0; l < NP; l++) P[l] = Pq[q * NP + l];
    const double jac = geom_eval<NP>(sc, P, g);
    double* o = node + 8 * q;
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    o[3] = g[9]; o[4] = g[10]; o[5] = g[11];
    o[6] = jac;
    o[7] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int ia = lane & 15, ka = lane >> 4;
  const int ntile = (b - a + 15) / 16;
  for (int tile = wv; tile < ntile; tile += nw) {
    const int k0 = a + 16 * tile;
    const bool row_ok = k0 + ia < b;
    const int2 jt = fixed_list[row_ok ? k0 + ia : k0];
    const double* tp = targs + ndtarg * (int64_t)jt.y;
    const double tx = tp[0], ty = tp[1], tz = tp[2];
    dbl4 acc[NPART][CT];
#pragma unroll
    for (int pa = 0; pa < NPART; pa++)
#pragma unroll
      for (int ct = 0; ct < CT; ct++) acc[pa][ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int q0 = 0; q0 < nf; q0 += 4) {
      const int q = q0 + ka;  // nf is a multiple of 4 (4^nlev XG triangles)
      const double* o = node + 8 * q;
      double vr, vi;
      if (HELM) {
        helm_comb(tx - o[0], ty - o[1], tz - o[2], o[3], o[4], o[5], kp, vr, vi);
      } else {
        vr = lap_comb(tx - o[0], ty - o[1], tz - o[2], o[3], o[4], o[5], kp);
        vi = 0.0;
      }
      const double fr = row_ok ? vr * o[6] : 0.0, fi = row_ok ? vi * o[6] : 0.0;
#pragma unroll
      for (int ct = 0; ct < CT; ct++) {
        const int j = ia + 16 * ct;
        const double bb = j < NP ? Bq[q * NP + j] : 0.0;
        acc[0][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr, bb, acc[0][ct], 0, 0, 0);
        if (HELM) acc[NPART - 1][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(fi, bb, acc[NPART - 1][ct], 0, 0, 0);
      }
    }
    // accumulator element r: target k0 + ka + 4 r, basis function ia + 16 ct
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = k0 + ka + 4 * r;
      if (k >= b) continue;
      const int64_t q0 = iquad[fixed_list[k].x] - 1;
#pragma unroll
      for (int ct = 0; ct < CT; ct++) {
        const int j = ia + 16 * ct;
        if (j >= NP) continue;
        if (HELM) {
          wnear[2 * (q0 + j)] = acc[0][ct][r];
          wnear[2 * (q0 + j) + 1] = acc[NPART - 1][ct][r];
        } else {
          wnear[q0 + j] = acc[0][ct][r];
        }
      }
    }
  }
}

// LDS staging row of one node: f_re, f_im, P_0..P_{NP-1}, padded to an even number of
// doubles so that every row starts on a 16-byte boundary and (f_re, f_im) is one 128-bit read.
__host__ __device__ constexpr int row_w(int np) { return (np + 3) & ~1; }

// One complex moment component per lane: lane j < NP holds M_j (im = 0 for Laplace).
struct Mom {
  double re, im;
  __device__ Mom& operator+=(const Mom& o) {
    re += o.re;
    im += o.im;
    return *this;
  }
};

// Apply umat^T to the moments (lane j < NP holds M_j) and store the weights of one pair;
// output c is weight c/2, part c%2 (complex) or weight c (real), so the stores coalesce.
template <int NP, bool HELM>
__device__ __forceinline__ void store_weights(double* buf, Mom mom, const double* __restrict__ umat,
                                              double* __restrict__ wnear, int64_t q0) {
  const int lane = threadIdx.x;
  constexpr int NC = HELM ? 2 * NP : NP;
  if (lane < NP) {
    buf[lane] = mom.re;
    if (HELM) buf[NP + lane] = mom.im;
  }
  __syncthreads();
  for (int c = lane; c < NC; c += WAVE) {
    const int j = HELM ? c >> 1 : c, part = HELM ? c & 1 : 0;
    double s = 0.0;
    for (int l = 0; l < NP; l++) s = fma(umat[l + j * NP], buf[part * NP + l], s);
    wnear[(HELM ? 2 : 1) * q0 + c] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ 4. adaptive
// The first depths of the refinement tree (1 + 4 + 16 + 64 + 256 triangles: root 0,
// children of the root 1..4, children of t in 1..4 at 5 + 4(t-1) + c, of t in 5..20 at
// 21 + 4(t-5) + c, of t in 21..84 at 85 + 4(t-21) + c) are the same for every target of
// a patch, as in the reference's per-patch tree (triaintrouts.f:939-1033).  Their node
// geometry is computed once per patch into HBM (one 64-byte record per node: x, y, z, nx,
// ny, nz, |J|*da*w, pad, so a node is read from one base address) for depths 0..3 or
// 0..4, and the basis values at those nodes once per order.
constexpr int NCACHE_TRI = 341;
constexpr int CACHE_STRIDE = 8;
constexpr int NCACHE_PARENT = 85;  // triangles whose children can be cached (depths 0..3)

// Row q of the cached-tree basis table: two slots for node q's kernel value (zero in the
// table; the adaptive kernel writes the value there in its LDS copy of the row), then
// P_0..P_{NP-1}.  Rows are padded to an even length, so every row starts on a 16-byte
// boundary and is read with 16-byte loads, and to a length that is not a multiple of 8
// doubles: the adaptive kernel's lanes write their rows of the same layout in LDS, and a
// row stride of a multiple of 64 bytes would put the lanes' writes into few banks (order
// 4, 15 basis functions: rows of 18 instead of 16 doubles, 5% faster).
__host__ __device__ constexpr int ptab_w(int np) {
  return row_w(np) % 8 == 0 ? row_w(np) + 2 : row_w(np);
}

// Nodes per chunk of a refinement step in the adaptive kernel, and the LDS of a chunk:
// whole 1024-byte blocks (one load-to-LDS instruction each) covering chunk_rows rows.
// At order 6, 55 rows of 30 doubles take 13 blocks (13 KB): 12 wavefronts per CU, where
// 64 rows (15 KB) allow 10; the order-6 plain rule of 55 nodes then fills one chunk per
// child.
__host__ __device__ constexpr int chunk_rows(int n) { return n >= 6 ? 55 : WAVE; }
__host__ __device__ constexpr int chunk_blocks(int n) {
  return (chunk_rows(n) * ptab_w((n + 1) * (n + 2) / 2) * 8 + 1023) / 1024;
}

__host__ __device__ __forceinline__ int cached_child(int t, int c) {
  return t < 0 ? -1
               : (t == 0 ? 1 + c
                         : (t <= 4 ? 5 + 4 * (t - 1) + c
                                   : (t <= 20 ? 21 + 4 * (t - 5) + c : (t <= 84 ? 85 + 4 * (t - 21) + c : -1))));
}

// The cache of a rule may hold fewer depths (ntri = 1 + 4 + ... triangles of the first
// depths, in tree order): a rule class whose pairs stop at depth 1 or 2 needs no more.
__host__ __device__ constexpr int cache_tris(int npar) { return 4 * npar + 1; }
__host__ __device__ constexpr int cache_parents(int depth) {
  return depth <= 1 ? 1 : (depth == 2 ? 5 : (depth == 3 ? 21 : NCACHE_PARENT));
}

// The cache on the matrix cores.  For one patch the geometry of every cached node is
// the product of the basis table [nodes x NP] with the patch's coefficients [NP x 9]; with
// 16 patches side by side each of the 9 components is a product [nodes x NP] x [NP x 16]:
// per 16-node tile, ceil(NP / 4) v_mfma_f64_16x16x4f64 steps per component, A = the
// tile's basis values (4 functions per step, lane l: node l % 16, function l / 16),
// B = 4 coefficients of the 16 patches (lane l: function l / 16, patch l % 16), read from
// the workgroup's LDS copy of the 16 patches' coefficients.  Accumulator element r of lane
// l is node i0 + l / 16 + 4 r of patch p0 + l % 16, whose 9 components the lane then holds
// together and turns into the 64-byte node record.  Grid: x = groups of 16 patches, y =
// slices of the node tiles (4 wavefronts per workgroup, one tile at a time each).  Three
// wavefronts per SIMD (the compiler's register choice at the default bound leaves two);
// the matrix-core version takes 4.3 ms on 12288 order-6 patches of 341 triangles of 55
// nodes, where one thread per node with vector FMAs took 7.5 ms.
template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_tree_cache_mfma(int64_t npatches, int ncn, const double* __restrict__ twq, const double* __restrict__ ptab,
                  const int64_t* __restrict__ ixyzs, const double* __restrict__ srccoefs,
                  double* __restrict__ cache) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  constexpr int KS = (NP + 3) / 4;
  constexpr int PW = ptab_w(NP);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ia = lane & 15, ka = lane >> 4;
  // B fragments of the 16 patches in LDS: step s, component m, lane l at ((s 9 + m) 64 + l)
  __shared__ double bs[KS * 9 * 64];
  __shared__ __align__(16) double rec[4 * 512];  // per wavefront: 64 records of 8 doubles
  for (int i = threadIdx.x; i < KS * 9 * 64; i += blockDim.x) {
    const int li = i & 63, sm = i >> 6;
    const int s = sm / 9, m = sm - 9 * s;
    const int l = 4 * s + (li >> 4);
    const int64_t pi = (int64_t)blockIdx.x * 16 + (li & 15);
    bs[i] = (pi < npatches && l < NP) ? srccoefs[9 * (ixyzs[pi] - 1) + 9 * l + m] : 0.0;
  }
  __syncthreads();
  const int ntile = (ncn + 15) / 16;
  const int tstep = 4 * gridDim.y;
  // A fragments of the next tile are loaded while the current one is multiplied
  double an[KS];
  auto load_a = [&](int tt) {
    // rows 16 tt .. 16 tt + 15 of the table (padded by 64 rows past its last node)
    const double* arow = ptab + (size_t)(16 * tt + ia) * PW + 2;
#pragma unroll
    for (int s = 0; s < KS; s++) {
      const int l = 4 * s + ka;
      an[s] = l < NP ? arow[l] : 0.0;
    }
  };
  int t = blockIdx.y * 4 + wv;
  if (t < ntile) load_a(t);
  for (; t < ntile; t += tstep) {
    const int i0 = 16 * t;
    double a[KS];
#pragma unroll
    for (int s = 0; s < KS; s++) a[s] = an[s];
    if (t + tstep < ntile) load_a(t + tstep);
    // read the B fragments from LDS in every tile rather than keep them in 126 registers
    asm volatile("" ::: "memory");
    dbl4 acc[9];
#pragma unroll
    for (int m = 0; m < 9; m++) acc[m] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; s++) {
#pragma unroll
      for (int m = 0; m < 9; m++)
        acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bs[(s * 9 + m) * 64 + lane], acc[m], 0, 0, 0);
    }
    // records through LDS, 64 at a time (node i0 + ka + 4 r of patch p0 + ia at slot
    // 4 ia + ka), then stored as 16-byte pieces by consecutive lanes: each store
    // instruction writes 4 patches' 4 consecutive records (256 contiguous bytes each)
    double* st = rec + wv * 512;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int n = i0 + ka + 4 * r;
      double g[12];
#pragma unroll
      for (int m = 0; m < 9; m++) g[m] = acc[m][r];
      const double jac = geom_finish(g);
      double* o = st + 8 * (4 * ia + ka);
      *reinterpret_cast<dbl4*>(o) = dbl4{g[0], g[1], g[2], g[9]};
      *reinterpret_cast<dbl4*>(o + 4) = dbl4{g[10], g[11], jac * (n < ncn ? twq[n] : 0.0), 0.0};
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the wavefront's own LDS writes
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int q = 64 * k + lane, rec_i = q >> 2, part = q & 3;
        const int64_t pp = (int64_t)blockIdx.x * 16 + (rec_i >> 2);
        const int nn = i0 + (rec_i & 3) + 4 * r;
        if (pp < npatches && nn < ncn)
          *reinterpret_cast<double2*>(cache + ((size_t)pp * ncn + nn) * CACHE_STRIDE + 2 * part) =
              *reinterpret_cast<const double2*>(st + 2 * q);
      }
      __builtin_amdgcn_s_waitcnt(0xC07F);
      __builtin_amdgcn_wave_barrier();
    }
  }
}

template <int N>
void launch_tree_cache(int64_t npat, int ncn, const double* twq, const double* ptab, const int64_t* ixyzs,
                       const double* srccoefs, double* cache, hipStream_t s) {
  const int64_t gx = (npat + 15) / 16;
  const int ntile = (ncn + 15) / 16;
  // about 4096 workgroups, each wavefront taking every (4 gy)-th tile of its slice
  int64_t gy = (4096 + gx - 1) / gx;
  if (gy > (ntile + 3) / 4) gy = (ntile + 3) / 4;
  if (gy < 1) gy = 1;
  k_tree_cache_mfma<N><<<dim3((unsigned)gx, (unsigned)gy), 256, 0, s>>>(npat, ncn, twq, ptab, ixyzs, srccoefs,
                                                                         cache);
}

// Rules of the adaptive integrator, one per class: the XG rule (kq nodes), the node
// geometry of the cached tree for that rule (per patch) and its basis table, and the
// deepest level a pair of the class may reach.  The work list is sorted by class;
// seg[c] .. seg[c + 1] - 1 are the pairs of class c.  A plain generation has one class:
// the reference's rule and level cap (NLMAX).
constexpr int MAX_RULES = 6;
struct AdapRules {
  int nrules;
  int kq[MAX_RULES];
  int lmax[MAX_RULES];
  int npar[MAX_RULES];  // cached triangles whose children are cached too (1, 5 or 21)
  int seg[MAX_RULES + 1];
  const double* xg[MAX_RULES];
  const double* cache[MAX_RULES];
  const double* ptab[MAX_RULES];
};

// Column sums of staged rows r0 <= r < r1 (no barriers: the caller brackets the chunk):
// row q at pb + q * PW holds the kernel value f_q of node q (re, im) and its basis values
// P_0..P_{NP-1}.
template <int NP, bool HELM>
__device__ __forceinline__ Mom column_sum_rows(const double* pb, int r0, int r1) {
  constexpr int PW = ptab_w(NP);
  constexpr int G = WAVE / NP;
  const int lane = threadIdx.x;
  const int grp = lane / NP, j = lane - grp * NP;
  double sr = 0.0, si = 0.0;
  if (grp < G) {
#pragma unroll 4
    for (int q = r0 + grp; q < r1; q += G) {
      const double p = pb[q * PW + 2 + j];
      const double2 f = *reinterpret_cast<const double2*>(pb + q * PW);
      sr = fma(f.x, p, sr);
      if (HELM) si = fma(f.y, p, si);
    }
  }
  const double pr = sr, pi = si;
#pragma unroll
  for (int k = 1; k < G; k++) {
    sr += __shfl(pr, (lane + k * NP) & 63, 64);
    if (HELM) si += __shfl(pi, (lane + k * NP) & 63, 64);
  }
  return lane < NP ? Mom{sr, HELM ? si : 0.0} : Mom{0.0, 0.0};
}

// Moments of the triangles of one refinement step (root: the patch triangle, m[0];
// otherwise the four children of the triangle (x0, y0, dx, dy) at tree id tid, m[0..3]):
// their ntri * kq nodes are spread over the lanes in chunks of at most chunk_rows(N)
// nodes, of equal size, so a small rule still fills the wavefront and the kernel
// evaluations of all children are in flight together.  Children of a cached triangle
// have consecutive tree ids, so a chunk's rows are one contiguous block of the basis
// table: it is copied to LDS by the memory unit (64 x 16 bytes per load-to-LDS
// instruction, no registers) while the lanes read the nodes' geometry from the tree
// cache and evaluate the kernel; each lane then puts its node's kernel value into the
// row's first two slots.  Below the cached depths
// each lane writes its node's basis row and sums the geometry from it and the patch
// coefficients sc in LDS.  CACHED selects the path at compile time (the caller tests the
// step's tree id), so each path is scheduled without the other's registers.
template <int N, bool HELM, bool CACHED>
__device__ __forceinline__ void step_moments(double* pb, const double* sc, bool root, int tid, double x0, double y0, double dx, double dy, int kq,
                                             const double* __restrict__ xg, const double* __restrict__ cpatch,
                                             const double* __restrict__ ptab, const double* tg, const KPar& kp,
                                             Mom (&m)[4]) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  constexpr int PW = ptab_w(NP);
  const int lane = threadIdx.x;
  const int ntri = root ? 1 : 4;
  const int total = ntri * kq;
  // cached: the root itself, or children of a cached triangle of depth <= 2
  constexpr bool cached = CACHED;
  const int t0 = root ? 0 : cached_child(tid, 0);  // first cached tree id of the step
  const double hx = 0.5 * dx, hy = 0.5 * dy, da = hx * hx;
#pragma unroll
  for (int c = 0; c < 4; c++) m[c] = Mom{0.0, 0.0};
  constexpr int CHM = chunk_rows(N);
  const int nch = (total + CHM - 1) / CHM;
  const int ch = (total + nch - 1) / nch;  // rows per chunk (the last one possibly fewer)
  for (int base = 0; base < total; base += ch) {
    const int n = base + lane;
    const int nrows = total - base < ch ? total - base : ch;
    if (cached) {
      // chunk_blocks(N) KB of the table from row t0 kq + base: the chunk's rows and a
      // part of the next ones (the table is padded by 64 rows past its last triangle)
      const double* blk = ptab + ((size_t)t0 * kq + base) * PW;
#pragma unroll
      for (int k = 0; k < chunk_blocks(N); k++)
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(blk + 2 * (k * WAVE + lane)),
            (__attribute__((address_space(3))) void*)(pb + 2 * k * WAVE), 16, 0, 0);
    }
    double2 fv = make_double2(0.0, 0.0);
    if (lane < nrows) {
      const int c = ntri == 1 ? 0 : (n >= kq) + (n >= 2 * kq) + (n >= 3 * kq);
      const int q = n - c * kq;
      double g[12], w;
      if (cached) {
        const double* cn = cpatch + (size_t)CACHE_STRIDE * ((size_t)t0 * kq + n);
        g[0] = cn[0];
        g[1] = cn[1];
        g[2] = cn[2];
        g[9] = cn[3];
        g[10] = cn[4];
        g[11] = cn[5];
        w = cn[6];
      } else {
        // child c: corner (x0, y0) + ((c & 1) hx, (c >> 1) hy), legs (hx, hy), the fourth reversed
        const double cx = (c & 1) ? x0 + hx : x0, cy = (c & 2) ? y0 + hy : y0;
        const double cdx = c == 3 ? -hx : hx, cdy = c == 3 ? -hy : hy;
        const double u = cx + cdx * xg[3 * q], v = cy + cdy * xg[3 * q + 1];
        koorn_geom_row<N>(u, v, sc, g, pb + lane * PW + 2);
        w = geom_finish(g) * da * xg[3 * q + 2];
      }
      double vr, vi;
      kern_eval<HELM>(tg, g, kp, vr, vi);
      fv = make_double2(vr * w, vi * w);
    }
    // the loads to LDS (counted by vmcnt) also fill the rows' kernel-value slots: wait
    // for them before the lanes write their values there
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (lane < nrows) *reinterpret_cast<double2*>(pb + lane * PW) = fv;
    __syncthreads();
    // the rows of this chunk belong to triangles c0..c1 (wave-uniform)
    const int c0 = ntri == 1 ? 0 : base / kq, c1 = ntri == 1 ? 0 : (base + nrows - 1) / kq;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (c < c0 || c > c1) continue;
      const int r0 = c * kq - base > 0 ? c * kq - base : 0;
      const int r1 = (c + 1) * kq - base < nrows ? (c + 1) * kq - base : nrows;
      m[c] += column_sum_rows<NP, HELM>(pb, r0, r1);
    }
    __syncthreads();
  }
}

// Moments of staged rows 0 <= q < nr, M_j = sum_q f_q P_j(q) (row q at pb + q PW holds
// f_q (re, im), then P_0..P_{NP-1}; lane j < NP returns M_j): the G = 64 / NP lane groups
// take rows G i + grp for i < nr / G (a wave-uniform trip count), then groups
// grp < nr mod G one more row each, one basis read serving the real and imaginary parts;
// the groups' partial sums are added with shuffles.
template <int NP, bool HELM, int PW = ptab_w(NP)>
__device__ __forceinline__ Mom column_sum_tri(const double* pb, int nr) {
  constexpr int G = WAVE / NP;
  const int lane = threadIdx.x;
  const int grp = lane / NP, j = lane - grp * NP;
  double sr = 0.0, si = 0.0;
  if (grp < G) {
    const double* b = pb + grp * PW + 2 + j;
    const int nfull = nr / G;
#pragma unroll 8
    for (int i = 0; i < nfull; i++) {
      const double p = b[G * i * PW];
      const double2 f = *reinterpret_cast<const double2*>(b - 2 - j + G * i * PW);
      sr = fma(f.x, p, sr);
      if (HELM) si = fma(f.y, p, si);
    }
    if (grp < nr - G * nfull) {
      const double p = b[G * nfull * PW];
      const double2 f = *reinterpret_cast<const double2*>(b - 2 - j + G * nfull * PW);
      sr = fma(f.x, p, sr);
      if (HELM) si = fma(f.y, p, si);
    }
  }
  const double pr = sr, pi = si;
#pragma unroll
  for (int k = 1; k < G; k++) {
    sr += __shfl(pr, (lane + k * NP) & 63, 64);
    if (HELM) si += __shfl(pi, (lane + k * NP) & 63, 64);
  }
  return lane < NP ? Mom{sr, HELM ? si : 0.0} : Mom{0.0, 0.0};
}

// step_moments for rules of at most chunk_rows(N) nodes: one chunk per triangle, so the
// triangle index is wave-uniform and the column sums run a wave-uniform trip count (on the
// deformed sphere of order 6 with the 55-node rule the adaptive kernel took 75.8 ms through
// step_moments, with the same chunks but per-lane child indices and per-lane row bounds in
// the column sums, and 71.0 ms here).
template <int N, bool HELM, bool CACHED>
__device__ __forceinline__ void step_tris(double* pb, const double* sc, bool root, int tid, double x0, double y0,
                                          double dx, double dy, int kq, const double* __restrict__ xg,
                                          const double* __restrict__ cpatch, const double* __restrict__ ptab,
                                          const double* tg, const KPar& kp, Mom (&m)[4]) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  constexpr int PW = ptab_w(NP);
  const int lane = threadIdx.x;
  const int ntri = root ? 1 : 4;
  const int t0 = root ? 0 : cached_child(tid, 0);
  const double hx = 0.5 * dx, hy = 0.5 * dy, da = hx * hx;
#pragma unroll
  for (int c = 0; c < 4; c++) m[c] = Mom{0.0, 0.0};
  for (int c = 0; c < ntri; c++) {
    if (CACHED) {
      const double* blk = ptab + ((size_t)(t0 + c) * kq) * PW;
#pragma unroll
      for (int k = 0; k < chunk_blocks(N); k++)
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(blk + 2 * (k * WAVE + lane)),
            (__attribute__((address_space(3))) void*)(pb + 2 * k * WAVE), 16, 0, 0);
    }
    double2 fv = make_double2(0.0, 0.0);
    if (lane < kq) {
      double g[12], w;
      if (CACHED) {
        const double* cn = cpatch + (size_t)CACHE_STRIDE * ((size_t)(t0 + c) * kq + lane);
        g[0] = cn[0];
        g[1] = cn[1];
        g[2] = cn[2];
        g[9] = cn[3];
        g[10] = cn[4];
        g[11] = cn[5];
        w = cn[6];
      } else {
        const double cx = (c & 1) ? x0 + hx : x0, cy = (c & 2) ? y0 + hy : y0;
        const double cdx = c == 3 ? -hx : hx, cdy = c == 3 ? -hy : hy;
        const double u = cx + cdx * xg[3 * lane], v = cy + cdy * xg[3 * lane + 1];
        koorn_geom_row<N>(u, v, sc, g, pb + lane * PW + 2);
        w = geom_finish(g) * da * xg[3 * lane + 2];
      }
      double vr, vi;
      kern_eval<HELM>(tg, g, kp, vr, vi);
      fv = make_double2(vr * w, vi * w);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (lane < kq) *reinterpret_cast<double2*>(pb + lane * PW) = fv;
    __syncthreads();
    m[c] = column_sum_tri<NP, HELM>(pb, kq);
    __syncthreads();
  }
}

// Adaptive integration (triaadap_main, triaintrouts.f:1269-1464): persistent wavefronts
// pull (target, patch) pairs from a queue; each pair's refinement tree is walked depth
// first from a per-wavefront stack (the leaf set is that of the reference's level-by-level
// walk: every decision depends only on its own triangle).  A step evaluates the four
// children of the popped triangle (step_tris: one chunk per child; step_moments: the
// children's nodes in equal chunks) and refines iff max_j |sum_c M_c,j - M_j| >
// epsp (triaintrouts.f:1435-1448); the children are pushed only when refined.  A pair of
// a routed class whose tree needs leaves below the class's level cap (lmax < NLMAX) is
// appended to fb_out
// (the plain integrator takes it afterwards) and gets no weights here.  The list is read
// from nlist or, when nlist_dev is set, from the device count (a second launch over the
// pairs a first one sent back).
template <int N, bool HELM>
__global__ void __launch_bounds__(64) k_adaptive(int nlist, const int* __restrict__ nlist_dev,
                                                 const int2* __restrict__ list, AdapRules R,
                                                 int2* __restrict__ fb_out, int* __restrict__ fb_cnt,
                                                 int* __restrict__ qctr, const int64_t* __restrict__ col_ind,
                                                 const int64_t* __restrict__ ixyzs,
                                                 const double* __restrict__ srccoefs, int64_t ndtarg,
                                                 const double* __restrict__ targs,
                                                 const int64_t* __restrict__ iquad,
                                                 const double* __restrict__ epsp_patch,
                                                 const double* __restrict__ umat, KPar kp,
                                                 double* __restrict__ stack_all, int* __restrict__ nref_out,
                                                 int* __restrict__ lev_out, int* __restrict__ unc_out,
                                                 int* __restrict__ overflow, double* __restrict__ wnear) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  constexpr int ES = 6 + 2 * NP;  // stack entry: x0,y0,dx,dy,level,tree id, moments
  __shared__ __align__(16) double pb[chunk_blocks(N) * 128];
  const int lane = threadIdx.x;
  double* stack = stack_all + (size_t)blockIdx.x * STACK_CAP * ES;
  const int nl = nlist_dev ? *nlist_dev : nlist;
  // Work queue: the list (sorted by class, then patch) in 8 contiguous segments, one per
  // XCD (workgroups are dispatched to the XCDs round-robin, so workgroup b starts on
  // segment b % 8 and the patch geometry it reads stays in its XCD's L2), then the other
  // segments once its own is drained.
  const int home = blockIdx.x & 7;
  int tried = 0;  // segments found empty (lane 0)
  for (;;) {
    // lane 0 claims the next pair; the whole wavefront is active here, so lane 0 is the
    // first active lane and readfirstlane broadcasts its claim
    int got = -1;
    if (lane == 0) {
      while (tried < 8) {
        const int sg = (home + tried) & 7;
        const int a = (int)((int64_t)nl * sg / 8), b = (int)((int64_t)nl * (sg + 1) / 8);
        const int k = atomicAdd(qctr + sg, 1);
        if (a + k < b) {
          got = a + k;
          break;
        }
        tried++;
      }
    }
    const int idx = __builtin_amdgcn_readfirstlane(got);
    if (idx < 0) break;
    int cls = 0;
    while (cls + 1 < R.nrules && idx >= R.seg[cls + 1]) cls++;
    cls = __builtin_amdgcn_readfirstlane(cls);
    const int kq = R.kq[cls], lmax = R.lmax[cls];
    const double* xg = R.xg[cls];
    const double* ptab = R.ptab[cls];
    const int2 jt = list[idx];
    const int64_t ip = col_ind[jt.x] - 1;
    // patch coefficients for the geometry below the cached tree, read with wave-uniform
    // (scalar) loads: in LDS they would cost the LDS of a wavefront per CU at order 6
    const double* lsc = srccoefs + 9 * (ixyzs[ip] - 1);
    const double* tp = targs + ndtarg * (int64_t)jt.y;
    const double tg[3] = {tp[0], tp[1], tp[2]};
    const double eps = epsp_patch[ip];
    const int npar = R.npar[cls];
    const double* cpatch = R.cache[cls] + (size_t)ip * CACHE_STRIDE * cache_tris(npar) * kq;
    Mom mc[4];
    // (at order 6 and above: at lower orders the second path costs a wavefront per SIMD)
    const bool per_tri = N >= 6 && kq <= chunk_rows(N) && 2 * kq > chunk_rows(N);
    if (per_tri)
      step_tris<N, HELM, true>(pb, lsc, true, 0, 0.0, 0.0, 1.0, 1.0, kq, xg, cpatch, ptab, tg, kp, mc);
    else
      step_moments<N, HELM, true>(pb, lsc, true, 0, 0.0, 0.0, 1.0, 1.0, kq, xg, cpatch, ptab, tg, kp, mc);
    // The triangle being tested is held in registers and the stack holds the others:
    // a refinement pushes children 0..2 and goes on with child 3 (the one the stack
    // order takes next).  At order 6 and above the entry on top of the stack is loaded at
    // the start of each test, so a leaf's successor is in registers when the test ends
    // (at lower orders those registers would cost a wavefront per SIMD).
    constexpr bool PREFETCH = N >= 6;
    double cx0 = 0.0, cy0 = 0.0, cdx = 1.0, cdy = 1.0;
    int clevel = 0, ctid = 0;
    Mom cpar = lane < NP ? mc[0] : Mom{0.0, 0.0};
    int sp = 0, nref = 0, deep = 0, nunc = 0;
    bool aborted = false, have = true;
    Mom result{0.0, 0.0};
    while (have) {
      double nx0 = 0.0, ny0 = 0.0, ndx = 0.0, ndy = 0.0, nlev = 0.0, ntid = 0.0;
      Mom tpar{0.0, 0.0};
      if (PREFETCH && sp > 0) {
        const double* e = stack + (size_t)(sp - 1) * ES;
        nx0 = e[0];
        ny0 = e[1];
        ndx = e[2];
        ndy = e[3];
        nlev = e[4];
        ntid = e[5];
        if (lane < NP) tpar = Mom{e[6 + lane], e[6 + NP + lane]};
      }
      const double x0 = cx0, y0 = cy0, dx = cdx, dy = cdy;
      const int level = clevel, tid = ctid;  // cached tree id, -1 below the cached depths
      const Mom mpar = cpar;
      if (per_tri) {
        if (tid >= 0 && tid < npar)
          step_tris<N, HELM, true>(pb, lsc, false, tid, x0, y0, dx, dy, kq, xg, cpatch, ptab, tg, kp, mc);
        else
          step_tris<N, HELM, false>(pb, lsc, false, tid, x0, y0, dx, dy, kq, xg, cpatch, ptab, tg, kp, mc);
      } else if (tid >= 0 && tid < npar) {
        step_moments<N, HELM, true>(pb, lsc, false, tid, x0, y0, dx, dy, kq, xg, cpatch, ptab, tg, kp, mc);
      } else {
        step_moments<N, HELM, false>(pb, lsc, false, tid, x0, y0, dx, dy, kq, xg, cpatch, ptab, tg, kp, mc);
      }
      nref++;
      nunc += (tid < 0 || tid >= npar) ? 1 : 0;  // children below the cached tree
      Mom ssum = mc[0];
      ssum += mc[1];
      ssum += mc[2];
      ssum += mc[3];
      const double dre = ssum.re - mpar.re, dim = ssum.im - mpar.im;
      double err = HELM ? sqrt(dre * dre + dim * dim) : fabs(dre);
      err = wave_max(lane < NP ? err : 0.0);
      const bool refine = err > eps;
      if (refine && lmax < NLMAX && level + 2 > lmax) {
        // refining would put leaves past the class's level cap: the prediction was wrong
        aborted = true;
        break;
      }
      // (sp + 4: the capacity rule of a stack that holds all four children)
      if (refine && level + 1 <= NLMAX && sp + 4 <= STACK_CAP) {
        const double hx = 0.5 * dx, hy = 0.5 * dy;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          double* ne = stack + (size_t)(sp + c) * ES;
          if (lane == 0) {
            const int ct = cached_child(tid, c);
            ne[0] = (c & 1) ? x0 + hx : x0;
            ne[1] = (c & 2) ? y0 + hy : y0;
            ne[2] = hx;
            ne[3] = hy;
            ne[4] = (double)(level + 1);
            ne[5] = (double)(ct >= 0 && ct < NCACHE_PARENT ? ct : -1);
          }
          if (lane < NP) {
            ne[6 + lane] = mc[c].re;
            ne[6 + NP + lane] = mc[c].im;
          }
        }
        sp += 3;
        const int ct3 = cached_child(tid, 3);
        cx0 = x0 + hx;
        cy0 = y0 + hy;
        cdx = -hx;
        cdy = -hy;
        clevel = level + 1;
        ctid = ct3 >= 0 && ct3 < NCACHE_PARENT ? ct3 : -1;
        cpar = lane < NP ? mc[3] : Mom{0.0, 0.0};
      } else {
        if (refine && lane == 0) atomicAdd(overflow, 1);  // level 20 or stack depth reached
        result += ssum;
        deep = level + 1 > deep ? level + 1 : deep;
        if (sp > 0) {
          sp--;
          if (!PREFETCH) {
            const double* e = stack + (size_t)sp * ES;
            nx0 = e[0];
            ny0 = e[1];
            ndx = e[2];
            ndy = e[3];
            nlev = e[4];
            ntid = e[5];
            if (lane < NP) tpar = Mom{e[6 + lane], e[6 + NP + lane]};
          }
          cx0 = wave_uniform(nx0);
          cy0 = wave_uniform(ny0);
          cdx = wave_uniform(ndx);
          cdy = wave_uniform(ndy);
          clevel = __builtin_amdgcn_readfirstlane((int)nlev);
          ctid = __builtin_amdgcn_readfirstlane((int)ntid);
          cpar = tpar;
        } else {
          have = false;
        }
      }
      __syncthreads();
    }
    __syncthreads();
    if (aborted) {
      if (lane == 0 && fb_out) fb_out[atomicAdd(fb_cnt, 1)] = jt;
      continue;
    }
    if (lane == 0) {
      if (nref_out) nref_out[jt.x] = nref;
      if (lev_out) lev_out[jt.x] = deep;
      if (unc_out) unc_out[jt.x] = nunc;
    }
    store_weights<NP, HELM>(pb, result, umat, wnear, iquad[jt.x] - 1);
  }
}

// ------------------------------------------------------------------ 4b. learned path
// Selector features of an adaptive pair (target x, source patch with centroid c, radius
// rad, area A, nodes y_i with unit normals n_i at RV (u_i, v_i)):
//   |x-c|/rad, |(x-c).n_mean|/rad, min_i |x-y_i|/rad, |(x-y_*).n_*|/rad,
//   min(u_*, v_*, 1-u_*-v_*), log10(epsp), |k| rad, rad^2/A        (* = nearest node)
template <int N>
__global__ void k_features(int nadap, const int2* __restrict__ list, const int64_t* __restrict__ col_ind,
                           const int64_t* __restrict__ ixyzs, const double* __restrict__ srcvals, int64_t ndtarg,
                           const double* __restrict__ targs, const double* __restrict__ cms,
                           const double* __restrict__ rads, const double* __restrict__ area,
                           const double* __restrict__ epsp, const double* __restrict__ rvuv, double kabs,
                           float* __restrict__ X, float* __restrict__ feat_full) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nadap) return;
  const int2 jt = list[idx];
  const int64_t ip = col_ind[jt.x] - 1, i0 = ixyzs[ip] - 1;
  const double* tp = targs + ndtarg * (int64_t)jt.y;
  const double x = tp[0], y = tp[1], z = tp[2];
  const double rad = rads[ip];
  const double cx = x - cms[3 * ip], cy = y - cms[3 * ip + 1], cz = z - cms[3 * ip + 2];
  double mx = 0.0, my = 0.0, mz = 0.0, dmin = 0.0, hn = 0.0;
  int imin = 0;
  for (int i = 0; i < NP; i++) {
    const double* sv = srcvals + 12 * (i0 + i);
    mx += sv[9];
    my += sv[10];
    mz += sv[11];
    const double dx = x - sv[0], dy = y - sv[1], dz = z - sv[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (i == 0 || d2 < dmin) {
      dmin = d2;
      imin = i;
      hn = fabs(dx * sv[9] + dy * sv[10] + dz * sv[11]);
    }
  }
  const double mn = sqrt(mx * mx + my * my + mz * mz);
  const double h = mn > 0.0 ? fabs(cx * mx + cy * my + cz * mz) / mn : 0.0;
  const double u = rvuv[2 * imin], v = rvuv[2 * imin + 1];
  const double edge = fmin(fmin(u, v), 1.0 - u - v);
  float f[8];
  f[0] = (float)(sqrt(cx * cx + cy * cy + cz * cz) / rad);
  f[1] = (float)(h / rad);
  f[2] = (float)(sqrt(dmin) / rad);
  f[3] = (float)(hn / rad);
  f[4] = (float)edge;
  f[5] = (float)log10(epsp[ip]);
  f[6] = (float)(kabs * rad);
  f[7] = (float)(rad * rad / area[ip]);
  for (int k = 0; k < 8; k++) X[8 * (int64_t)idx + k] = f[k];
  if (feat_full)
    for (int k = 0; k < 8; k++) feat_full[8 * (int64_t)jt.x + k] = f[k];
}

// Bucket key of an adaptive pair (MLP_OUT * npatches keys): rule class, then patch.
// Without a prediction every pair is in class 0 (the plain rule), sorted by patch.
__device__ __forceinline__ int64_t bucket_key(int c, int64_t patch, int64_t npat) { return c * npat + patch; }

__global__ void k_bucket_count(int n, const int2* __restrict__ list, const int* __restrict__ pred,
                               const int64_t* __restrict__ col_ind, int64_t npat, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(cnt + bucket_key(pred ? pred[i] : 0, col_ind[list[i].x] - 1, npat), 1);
}

__global__ void k_bucket_scatter(int n, const int2* __restrict__ list, const int* __restrict__ pred,
                                 const int64_t* __restrict__ col_ind, int64_t npat, const int* __restrict__ off,
                                 int* __restrict__ fill, int2* __restrict__ out, int* __restrict__ pred_full) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 jt = list[i];
  const int64_t key = bucket_key(pred ? pred[i] : 0, col_ind[jt.x] - 1, npat);
  out[off[key] + atomicAdd(fill + key, 1)] = jt;
  if (pred_full) pred_full[jt.x] = pred[i];
}

// ------------------------------------------------------------------ 5. self
template <int N, bool HELM>
__global__ void __launch_bounds__(64) k_self(int nself, const int2* __restrict__ list,
                                             const int64_t* __restrict__ col_ind,
                                             const int64_t* __restrict__ ixyzs,
                                             const double* __restrict__ srccoefs, int64_t ndtarg,
                                             const double* __restrict__ targs,
                                             const double* __restrict__ uvs_targ,
                                             const int64_t* __restrict__ iquad,
                                             const double* __restrict__ umat, KPar kp, int nr, int nt,
                                             double tau_panel, double* __restrict__ wnear) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  __shared__ __align__(16) double buf[WAVE * row_w(NP)];
  const int lane = threadIdx.x;
  const int idx = blockIdx.x;
  if (idx >= nself) return;
  const int2 jt = list[idx];
  const int64_t ip = col_ind[jt.x] - 1;
  const double* sc = srccoefs + 9 * (ixyzs[ip] - 1);
  const double* tp = targs + ndtarg * (int64_t)jt.y;
  const double tg[3] = {tp[0], tp[1], tp[2]};
  const double u0 = uvs_targ[2 * (int64_t)jt.y], v0 = uvs_targ[2 * (int64_t)jt.y + 1];
  double P[NP], g0[12];
  koorn_eval<N>(u0, v0, P);
  geom_eval<NP>(sc, P, g0);
  const double g11 = g0[3] * g0[3] + g0[4] * g0[4] + g0[5] * g0[5];
  const double g12 = g0[3] * g0[6] + g0[4] * g0[7] + g0[5] * g0[8];
  const double g22 = g0[6] * g0[6] + g0[7] * g0[7] + g0[8] * g0[8];
  const double l11 = sqrt(g11), l12 = g12 / l11, l22 = sqrt(g22 - l12 * l12);
  const double detl = l11 * l22;
  const double vx[3] = {0.0, 1.0, 0.0}, vy[3] = {0.0, 0.0, 1.0};
  Mom acc{0.0, 0.0};
  const int nn = nr * nt;
  // Principal value over a disk about the target in the conformal frame (the reference's
  // GGQ self rule, ggq-selfquad-routs.f90:16-214: frame from the QR factor of (x_u, x_v),
  // disk radius rcut = half the distance to the nearest edge line).  On each ray of the
  // product rule the radial Gauss rule gives G(0) C_GL for the 1/rho part of the
  // integrand (C_GL = sum_i w_i / rho_i), which the disk convention integrates from
  // rcut/R to 1: every ray adds G(0) (ln(R / rcut) - C_GL).  G(0) is the double layer's
  // 1/r^2 coefficient, nonzero when the derivative data are not the derivatives of the
  // xyz expansion: -beta (D.n0) / (4 pi |D|^3) P_j(uv0) |J0| (ray Jacobian), D = grad X
  // along the ray.  grad X comes from an 11-point central difference (exact for the
  // expansion's degree <= 10); lanes 0..19 take one stencil point each.
  __shared__ double fd[20][3];
  __shared__ double pj[NP];
  const bool has_dl = kp.br != 0.0 || kp.bi != 0.0;
  double gx[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  if (has_dl) {
    if (lane < 20) {
      const int dir = lane / 10, k = 1 + (lane % 10) / 2, sgn = (lane & 1) ? 1 : -1;
      constexpr double hfd = 0.1;
      const double cfd = k == 1 ? 5.0 / 6 : (k == 2 ? -5.0 / 21 : (k == 3 ? 5.0 / 84 : (k == 4 ? -5.0 / 504 : 1.0 / 1260)));
      double Pf[NP], gf[12];
      koorn_eval<N>(u0 + (dir == 0 ? sgn * k * hfd : 0.0), v0 + (dir == 1 ? sgn * k * hfd : 0.0), Pf);
      geom_eval<NP>(sc, Pf, gf);
      for (int m = 0; m < 3; m++) fd[lane][m] = sgn * cfd * gf[m] / hfd;
    }
    if (lane < NP) {
      double Pa[NP];
      koorn_eval<N>(u0, v0, Pa);
#pragma unroll
      for (int j = 0; j < NP; j++)
        if (j == lane) pj[j] = Pa[j];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 10; q++)
#pragma unroll
      for (int m = 0; m < 3; m++) {
        gx[0][m] += fd[q][m];
        gx[1][m] += fd[10 + q][m];
      }
  }
  double cgl = 0.0;
  for (int i = 0; i < nr; i++) cgl += g_self_gl.wr[i] / g_self_gl.xr[i];
  double pvs = 0.0;  // this lane's rays: sum of wtau (ray Jacobian) kappa (ln(R / rcut) - C_GL)
  const float inv_nr = 1.0f / (float)nr;
  // The three triangles at the target, each split at the foot of the perpendicular: up to
  // six segments of sinh-graded panels.  Their parameters go to LDS, and the nodes of all
  // panels of all segments are taken 64 at a time (one pass per 64 nodes rather than per
  // panel: a 100-node panel would leave 28 lanes of its second pass idle).
  __shared__ double segd[6][8];  // fx, fy, ex, ey, d, ta, tb, npan
  __shared__ int segc[7];        // first node of each segment; segc[nseg] = node count
  __shared__ int segt[7];        // first ray (tau node) of each segment
  __shared__ double sh_rcut;
  // Lane 2e + k < 6 takes piece k of edge e: the edge parameter from its start to the
  // foot (k = 0; to the end when the foot is outside the edge) or from the foot to the
  // end (k = 1, when the foot splits the edge).  The pieces keep the order edge by edge.
  bool has_piece = false;
  if (lane < 6) {
    const int e = lane >> 1, k = lane & 1;
    const double Ax = vx[e] - u0, Ay = vy[e] - v0;
    const double Bx = vx[(e + 1) % 3] - u0, By = vy[(e + 1) % 3] - v0;
    const double ax = l11 * Ax + l12 * Ay, ay = l22 * Ay;
    const double bx = l11 * Bx + l12 * By, by = l22 * By;
    const double len = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
    const double ex = (bx - ax) / len, ey = (by - ay) / len;
    const double ts = -(ax * ex + ay * ey);
    const double fx = ax + ts * ex, fy = ay + ts * ey;
    const double d = sqrt(fx * fx + fy * fy);
    const double s0 = -ts, s1 = len - ts;
    const bool split = s0 < 0.0 && s1 > 0.0;
    has_piece = k == 0 || split;
    if (has_piece) {
      const double sa = k == 0 ? s0 : 0.0, sb = (k == 0 && split) ? 0.0 : s1;
      const double ta = asinh(sa / d), tb = asinh(sb / d);
      const double span = fabs(tb - ta) / tau_panel;
      int npan = span < 64.0 ? (int)ceil(span) : 64;  // also maps NaN to 1 below
      if (!(npan >= 1)) npan = 1;
      const uint64_t below = __ballot(1) & ((1ull << lane) - 1);  // active lanes: the pieces
      double* sd = segd[__popcll(below)];
      sd[0] = fx; sd[1] = fy; sd[2] = ex; sd[3] = ey; sd[4] = d; sd[5] = ta; sd[6] = tb; sd[7] = npan;
    }
  }
  const int nseg = __popcll(__ballot(has_piece));
  __syncthreads();
  if (lane == 0) {
    int cnt = 0, cntt = 0;
    double hmin = 1e300;  // every edge has a piece k = 0
    for (int s = 0; s < nseg; s++) {
      const int npan = (int)segd[s][7];
      hmin = segd[s][4] < hmin ? segd[s][4] : hmin;
      segt[s] = cntt;
      segc[s] = cnt;
      cnt += npan * nn;
      cntt += npan * nt;
    }
    for (int k = nseg; k < 7; k++) {
      segc[k] = cnt;
      segt[k] = cntt;
    }
    sh_rcut = 0.5 * hmin;
  }
  __syncthreads();
  const int total = segc[6];
  const double rcut = sh_rcut;
  for (int base = 0; base < total; base += WAVE) {
    const int k = base + lane;
    if (k < total) {
      int sgi = 0;
      while (sgi < 5 && k >= segc[sgi + 1]) sgi++;
      const double* sd = segd[sgi];
      const double fx = sd[0], fy = sd[1], ex = sd[2], ey = sd[3], d = sd[4], ta = sd[5], tb = sd[6];
      const int npan = (int)sd[7];
      const int loc = k - segc[sgi];
      const int ipan = loc / nn, kk = loc - ipan * nn;
      const double t0 = ta + (tb - ta) * ipan / npan, t1 = ta + (tb - ta) * (ipan + 1) / npan;
      double* row = buf + lane * row_w(NP);
      // kk < 256: a float reciprocal gives the exact quotient
      const int it = (int)(((float)kk + 0.5f) * inv_nr), ir = kk - it * nr;
      const double tau = t0 + (t1 - t0) * g_self_gl.xt[it];
      // sinh and cosh from one exponential (only d * sinh enters, so the absolute error
      // of the difference near tau = 0 is what matters)
      const double et = exp(tau), eti = 1.0 / et;
      const double sh = 0.5 * (et - eti), ch = 0.5 * (et + eti);
      const double wtau = g_self_gl.wt[it] * (t1 - t0);
      const double px = fx + d * sh * ex, py = fy + d * sh * ey;
      const double rho = g_self_gl.xr[ir];
      const double xi0 = rho * px, xi1 = rho * py;
      const double dv = xi1 / l22, du = (xi0 - l12 * dv) / l11;
      double g[12];
      koorn_geom_row<N>(u0 + du, v0 + dv, sc, g, row + 2);
      const double jac = geom_finish(g);
      const double w = g_self_gl.wr[ir] * wtau * rho * d * d * ch / detl * jac;
      double vr, vi;
      kern_eval<HELM>(tg, g, kp, vr, vi);
      row[0] = vr * w;
      row[1] = vi * w;
    }
    const int nrows = total - base < WAVE ? total - base : WAVE;
    __syncthreads();
    acc += column_sum_tri<NP, HELM, row_w(NP)>(buf, nrows);
    __syncthreads();
  }
  if (has_dl) {
    // the disk correction of every ray (tau node): rays spread over the lanes
    const int ntau = segt[6];
    for (int k = lane; k < ntau; k += WAVE) {
      int sgi = 0;
      while (sgi < 5 && k >= segt[sgi + 1]) sgi++;
      const double* sd = segd[sgi];
      const double fx = sd[0], fy = sd[1], ex = sd[2], ey = sd[3], d = sd[4], ta = sd[5], tb = sd[6];
      const int npan = (int)sd[7];
      const int loc = k - segt[sgi];
      const int ipan = loc / nt, it = loc - ipan * nt;
      const double t0 = ta + (tb - ta) * ipan / npan, t1 = ta + (tb - ta) * (ipan + 1) / npan;
      const double tau = t0 + (t1 - t0) * g_self_gl.xt[it];
      const double et = exp(tau), eti = 1.0 / et;
      const double sh = 0.5 * (et - eti), ch = 0.5 * (et + eti);
      const double wtau = g_self_gl.wt[it] * (t1 - t0);
      const double px = fx + d * sh * ex, py = fy + d * sh * ey;
      const double dv0 = py / l22, du0 = (px - l12 * dv0) / l11;
      const double Dx = gx[0][0] * du0 + gx[1][0] * dv0, Dy = gx[0][1] * du0 + gx[1][1] * dv0,
                   Dz = gx[0][2] * du0 + gx[1][2] * dv0;
      const double dn = Dx * g0[9] + Dy * g0[10] + Dz * g0[11], d2 = Dx * Dx + Dy * Dy + Dz * Dz;
      const double kap = -dn / (4.0 * M_PI * d2 * sqrt(d2));
      pvs += wtau * d * d * ch / detl * kap * (log(d * ch / rcut) - cgl);
    }
    // beta * (sum over rays) * |J0| * P_j(uv0) joins moment j (lane j)
    const double pv = wave_sum(pvs);
    const double jac0 = sqrt((g0[4] * g0[8] - g0[5] * g0[7]) * (g0[4] * g0[8] - g0[5] * g0[7]) +
                             (g0[5] * g0[6] - g0[3] * g0[8]) * (g0[5] * g0[6] - g0[3] * g0[8]) +
                             (g0[3] * g0[7] - g0[4] * g0[6]) * (g0[3] * g0[7] - g0[4] * g0[6]));
    if (lane < NP) {
      const double c = pv * jac0 * pj[lane];
      acc.re += kp.br * c;
      if (HELM) acc.im += kp.bi * c;
    }
  }
  store_weights<NP, HELM>(buf, acc, umat, wnear, iquad[jt.x] - 1);
}

template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
  return (T*)p;
}

// Device workspace reused across calls: one grow-only buffer per slot and device.  Calls
// are serialised (g_quad_mutex) and every call synchronises its stream before returning,
// so a buffer is never resized while a kernel of an earlier call uses it.
enum WsSlot {
  W_CMS, W_RADS, W_EPSP, W_AREA, W_CNT, W_OFF, W_FILL, W_CLS, W_COUNTS, W_ADAP, W_SELF, W_FIXED, W_NREF,
  W_CACHE, W_STACK, W_X, W_PRED, W_BCNT, W_BOFF, W_SORTED, W_FB, W_PFULL, W_QCTR, W_LEV, W_SEG, W_UNC,
  W_RCACHE0,  // tree caches of the routed rule classes: W_RCACHE0 + class
  W_NSLOTS = W_RCACHE0 + MAX_RULES
};
struct WsBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
std::mutex g_quad_mutex;
// per device: a second stream for the fixed-rule and self kernels, which run beside the
// tree cache and the adaptive kernel (their pairs and weight slices are disjoint)
std::map<int, hipStream_t> g_side;
int g_use_side = 1;
std::map<int, std::vector<WsBuf>> g_ws;

template <typename T>
T* wsalloc(int slot, size_t n) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto& v = g_ws[dev];
  if (v.size() < W_NSLOTS) v.resize(W_NSLOTS);
  const size_t b = (n ? n : 1) * sizeof(T);
  if (v[slot].bytes < b) {
    if (v[slot].p) (void)hipFree(v[slot].p);
    v[slot].p = nullptr;
    v[slot].bytes = 0;
    if (hipMalloc(&v[slot].p, b) != hipSuccess) return nullptr;
    v[slot].bytes = b;
  }
  return (T*)v[slot].p;
}

// Patch-independent tables of one (order, rule) combination, built once per device.
struct OrderTables {
  int kq = 0, nf = 0;
  double *umat, *rvw, *rvuv, *xg, *Pq, *Bq, *twq, *ptab;
};
std::map<std::tuple<int, int, int, int, int>, OrderTables> g_tables;

// The adaptive integrator's XG rule of degree q with the node tables of the cached tree
// (depths 0..4) for patches of order N; built once per device.
struct RuleTables {
  int kq = 0;
  double *xg, *twq, *ptab;
};
std::map<std::tuple<int, int, int>, RuleTables> g_rules;

#define HCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "fmm3dbie_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return -1;                                                                  \
    }                                                                             \
  } while (0)

// Constant-memory tables are per device: uploaded once per device (Koornwinder
// constants) and when the self rule size changes (its Gauss-Legendre rule).
std::map<int, int> g_const_dev;  // device -> self rule sizes last uploaded, 16 nr + nt (0: none)

int ensure_constants() {
  int dev = 0;
  HCHK(hipGetDevice(&dev));
  if (g_const_dev.count(dev)) return 0;
  KoornConst kc;
  fmmhost::fill_koorn_const(&kc);
  HCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_koorn), &kc, sizeof(kc)));
  g_const_dev[dev] = 0;
  return 0;
}

int set_self_gl(int nr, int nt) {
  int dev = 0;
  HCHK(hipGetDevice(&dev));
  const int key = 16 * nr + nt;
  auto it = g_const_dev.find(dev);
  if (it != g_const_dev.end() && it->second == key) return 0;
  SelfGL gl = {};
  fmmhost::gauss_legendre01(nr, gl.xr, gl.wr);
  fmmhost::gauss_legendre01(nt, gl.xt, gl.wt);
  HCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_self_gl), &gl, sizeof(gl)));
  g_const_dev[dev] = key;
  return 0;
}

// Timing events of the generation, created once per device and reused: every call
// records each of them before reading any, and synchronises before it returns.
constexpr int NQ_EVENTS = 15;
std::map<int, std::array<hipEvent_t, NQ_EVENTS>> g_events;

hipEvent_t* quad_events() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto it = g_events.find(dev);
  if (it == g_events.end()) {
    std::array<hipEvent_t, NQ_EVENTS> e{};
    for (auto& x : e) {
      if (hipEventCreate(&x) != hipSuccess) {
        for (auto& y : e)
          if (y) (void)hipEventDestroy(y);
        return nullptr;
      }
    }
    it = g_events.emplace(dev, e).first;
  }
  return it->second.data();
}

template <typename T>
T* upload(const std::vector<T>& h, hipStream_t s) {
  T* d = dalloc<T>(h.size());
  if (d && hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess) return nullptr;
  return d;
}

// Node tables of the cached tree (depths 0..4: 341 triangles) for the rule (xq, xw):
// (u, v) and area-scaled weight of every node and the basis values there (patch
// independent; rows of ptab_w(NP) doubles, the basis from slot 2).
template <int N>
void cached_tree_tables(const std::vector<double>& xq, const std::vector<double>& xw, std::vector<double>& tuv,
                        std::vector<double>& twq, std::vector<double>& ptab) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  const int kq = (int)xw.size();
  const int ncn = NCACHE_TRI * kq;
  tuv.assign(2 * ncn, 0.0);
  twq.assign(ncn, 0.0);
  // 64 rows of padding: the adaptive kernel copies whole 1024-byte blocks of the table
  // from the first row of a chunk, up to chunk_blocks KB
  ptab.assign((size_t)(ncn + WAVE) * ptab_w(NP), 0.0);
  std::vector<double> pols(NP);
  struct Tri { double x0, y0, dx, dy; };
  std::vector<Tri> tris(NCACHE_TRI);
  tris[0] = {0.0, 0.0, 1.0, 1.0};
  auto kids = [](Tri t, int c) -> Tri {
    const double hx = t.dx / 2, hy = t.dy / 2;
    const Tri k[4] = {{t.x0, t.y0, hx, hy}, {t.x0 + hx, t.y0, hx, hy}, {t.x0, t.y0 + hy, hx, hy},
                      {t.x0 + hx, t.y0 + hy, -hx, -hy}};
    return k[c];
  };
  for (int t = 0; t < NCACHE_PARENT; t++)
    for (int c = 0; c < 4; c++) tris[cached_child(t, c)] = kids(tris[t], c);
  for (int t = 0; t < NCACHE_TRI; t++) {
    const double da = t == 0 ? 1.0 : (t <= 4 ? 0.25 : (t <= 20 ? 0.0625 : (t <= 84 ? 0.015625 : 0.00390625)));
    for (int q = 0; q < kq; q++) {
      const int n = t * kq + q;
      tuv[2 * n] = tris[t].x0 + tris[t].dx * xq[2 * q];
      tuv[2 * n + 1] = tris[t].y0 + tris[t].dy * xq[2 * q + 1];
      twq[n] = da * xw[q];
      fmmhost::koorn_pols(tuv[2 * n], tuv[2 * n + 1], N, pols.data());
      for (int l = 0; l < NP; l++) ptab[(size_t)n * ptab_w(NP) + 2 + l] = pols[l];
    }
  }
}

// umat, RV nodes/weights, the XG rule of the adaptive integrator, the fixed rule with its
// basis (Pq) and weight-folded (Bq) tables, and the node tables of the cached tree
template <int N>
int build_tables(int nqorder, int nlev, int nqorder_f, OrderTables& T, hipStream_t s) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  std::vector<double> umat, rvuv(2 * NP), rvw(NP);
  fmmhost::rv_umat(N, umat);
  fmmhost::rv_nodes(N, rvuv.data(), rvw.data());
  std::vector<double> xq, xw, fuv, fw;
  const int kq = fmmhost::xg_rule(nqorder, xq, xw);
  const int nf = fmmhost::xg_unif_nodes(nlev, nqorder_f, fuv, fw);
  std::vector<double> xg(3 * kq), Pq((size_t)nf * NP), Bq((size_t)nf * NP), pols(NP);
  for (int q = 0; q < kq; q++) { xg[3 * q] = xq[2 * q]; xg[3 * q + 1] = xq[2 * q + 1]; xg[3 * q + 2] = xw[q]; }
  for (int q = 0; q < nf; q++) {
    fmmhost::koorn_pols(fuv[2 * q], fuv[2 * q + 1], N, pols.data());
    for (int l = 0; l < NP; l++) Pq[(size_t)q * NP + l] = pols[l];
    for (int j = 0; j < NP; j++) {
      double sacc = 0.0;
      for (int l = 0; l < NP; l++) sacc += pols[l] * umat[l + (size_t)j * NP];
      Bq[(size_t)q * NP + j] = fw[q] * sacc;
    }
  }
  std::vector<double> tuv, twq, ptab;
  cached_tree_tables<N>(xq, xw, tuv, twq, ptab);
  T.kq = kq;
  T.nf = nf;
  T.umat = upload(umat, s);
  T.rvw = upload(rvw, s);
  T.rvuv = upload(rvuv, s);
  T.xg = upload(xg, s);
  T.Pq = upload(Pq, s);
  T.Bq = upload(Bq, s);
  T.twq = upload(twq, s);
  T.ptab = upload(ptab, s);
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  if (!T.umat || !T.rvw || !T.rvuv || !T.xg || !T.Pq || !T.Bq || !T.twq || !T.ptab) return -2;
  return 0;
}

// tables of the adaptive rule of degree q for order N (routed classes)
template <int N>
int rule_tables(int q, const RuleTables** out, hipStream_t s) {
  int dev = 0;
  HCHK(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, N, q);
  auto it = g_rules.find(key);
  if (it == g_rules.end()) {
    std::vector<double> xq, xw;
    const int kq = fmmhost::xg_rule(q, xq, xw);
    std::vector<double> xg(3 * kq), tuv, twq, ptab;
    for (int k = 0; k < kq; k++) { xg[3 * k] = xq[2 * k]; xg[3 * k + 1] = xq[2 * k + 1]; xg[3 * k + 2] = xw[k]; }
    cached_tree_tables<N>(xq, xw, tuv, twq, ptab);
    RuleTables T;
    T.kq = kq;
    T.xg = upload(xg, s);
    T.twq = upload(twq, s);
    T.ptab = upload(ptab, s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (!T.xg || !T.twq || !T.ptab) return -2;
    it = g_rules.emplace(key, T).first;
  }
  *out = &it->second;
  return 0;
}

// XG degree of rule class c around the plain degree q0 (class 2 is the plain rule); no
// class goes below degree 8, where the children-against-parent test stops being a
// reliable error estimate, or above 30, the largest rule in the tables
inline int class_degree(int q0, int c) {
  static const int off[MAX_RULES] = {-4, -2, 0, 4, 8, 12};
  const int q = q0 + off[c];
  return q < 8 ? (q0 < 8 ? q0 : 8) : (q > 30 ? 30 : q);
}

__global__ void k_seg_bounds(int nrules, int64_t npat, const int* __restrict__ boff, int* __restrict__ seg) {
  const int c = threadIdx.x;
  if (c <= nrules) seg[c] = boff[c * npat];
}

// Persistent wavefronts of the adaptive kernel: one per resident slot on this device
// (CUs x workgroups per CU at the kernel's register and LDS use); more would only wait
// for a slot and find the queue drained, fewer leave slots to the side-stream kernels
// but slow the kernel alone.
template <int N, bool HELM>
int adaptive_slots() {
  static std::map<int, int> slots;  // per device
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  auto it = slots.find(dev);
  if (it != slots.end()) return it->second;
  int ncu = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adaptive<N, HELM>, 64, 0) != hipSuccess)
    return 0;
  int n = ncu * per_cu;
  if (n < 64) n = 64;
  if (n > 8192) n = 8192;
  slots[dev] = n;
  return n;
}

template <int N, bool HELM>
void launch_adaptive(int g, hipStream_t s, int nlist, const int* nlist_dev, const int2* list, const AdapRules& R,
                     int2* fb_out, int* fb_cnt, int* qctr, const NearQuadArgs& A, const double* epsp,
                     const double* umat, const KPar& kp, double* stack, int* nref, int* lev, int* unc, int* over) {
  k_adaptive<N, HELM><<<g, 64, 0, s>>>(nlist, nlist_dev, list, R, fb_out, fb_cnt, qctr, A.col_ind, A.ixyzs,
                                       A.srccoefs, A.ndtarg, A.targs, A.iquad, epsp, umat, kp, stack, nref, lev, unc,
                                       over, A.wnear);
}

template <int N, bool HELM>
int run_order(const NearQuadArgs& A, NearQuadStats* st, hipStream_t s) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  const int64_t npat = A.npatches;
  // tables (built and uploaded on the first call for this order and rule)
  int nqorder, nlev, nqorder_f;
  double eps_adap;
  fmmhost::quadparams_adap(A.eps, A.norder_branch, &nqorder, &eps_adap, &nlev, &nqorder_f);
  int dev = 0;
  HCHK(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, N, nqorder, nlev, nqorder_f);
  auto tit = g_tables.find(key);
  if (tit == g_tables.end()) {
    OrderTables T;
    if (build_tables<N>(nqorder, nlev, nqorder_f, T, s) != 0) return -2;
    tit = g_tables.emplace(key, T).first;
  }
  const OrderTables& TB = tit->second;
  const int kq = TB.kq, nf = TB.nf;
  double *d_umat = TB.umat, *d_rvw = TB.rvw, *d_rvuv = TB.rvuv, *d_xg = TB.xg, *d_Pq = TB.Pq, *d_Bq = TB.Bq;
  double *d_cms = wsalloc<double>(W_CMS, 3 * npat), *d_rads = wsalloc<double>(W_RADS, npat),
         *d_epsp = wsalloc<double>(W_EPSP, npat), *d_area = wsalloc<double>(W_AREA, npat);
  int *d_cnt = wsalloc<int>(W_CNT, npat), *d_off = wsalloc<int>(W_OFF, npat + 1),
      *d_fill = wsalloc<int>(W_FILL, npat);
  int *d_cls = wsalloc<int>(W_CLS, A.nnz), *d_counts = wsalloc<int>(W_COUNTS, 8);
  int2 *d_adap = wsalloc<int2>(W_ADAP, A.nnz), *d_self = wsalloc<int2>(W_SELF, A.nnz),
       *d_fixed = wsalloc<int2>(W_FIXED, A.nnz);
  int* d_nref = A.want_nref ? wsalloc<int>(W_NREF, A.nnz) : nullptr;
  if (!d_cms || !d_rads || !d_epsp || !d_cnt || !d_off || !d_fill || !d_cls || !d_counts || !d_adap || !d_self ||
      !d_fixed || !d_area || (A.want_nref && !d_nref)) {
    fprintf(stderr, "fmm3dbie_amd: device allocation failed in getnearquad\n");
    return -2;
  }
  HCHK(hipMemsetAsync(d_cnt, 0, sizeof(int) * npat, s));
  HCHK(hipMemsetAsync(d_fill, 0, sizeof(int) * npat, s));
  HCHK(hipMemsetAsync(d_counts, 0, sizeof(int) * 8, s));
  KPar kp;
  kp.zkr = A.zk[0]; kp.zki = A.zk[1];
  kp.ar = A.alpha[0]; kp.ai = A.alpha[1];
  kp.br = A.beta[0]; kp.bi = A.beta[1];

  hipEvent_t* ev = quad_events();
  if (!ev) {
    fprintf(stderr, "fmm3dbie_amd: event creation failed in getnearquad\n");
    return -1;
  }
  HCHK(hipEventRecord(ev[0], s));
  // side stream: the tree cache (on-surface targets, where adaptive pairs are all but
  // certain: it then runs beside the list building and its host read-back), then the
  // fixed-rule and self kernels
  hipStream_t s2 = s;
  if (g_use_side && npat <= SIDE_MAX_PATCHES) {
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    auto it = g_side.find(dev);
    if (it == g_side.end()) {
      hipStream_t ns;
      HCHK(hipStreamCreateWithFlags(&ns, hipStreamNonBlocking));
      it = g_side.emplace(dev, ns).first;
    }
    s2 = it->second;
  }
  // Depth of the plain rule's tree cache: 0..4 when the surface has many near pairs per
  // patch, else 0..3.  Depth 4 costs one more pass over 256 triangles per patch (measured
  // 4.5e-4 ms per patch) and saves each refinement test of a depth-3 triangle its
  // geometry (about 8e-6 ms); on the benchmark surfaces a third of the near pairs are
  // adaptive and 8-22% of those reach depth 4, which pays from about 300 near pairs per
  // patch (sphere 12x12 order 4: 235, depth 3; torus 48x16 order 6: 510 and deformed
  // sphere 32x32 order 6: 436, depth 4, 6-12% less adaptive time including the cache).
  const int npar_plain = cache_parents(A.nnz >= 300 * npat ? 4 : 3);
  const int ntri_plain = cache_tris(npar_plain);
  const int ncn = ntri_plain * kq;
  double *d_twq = TB.twq, *d_ptab = TB.ptab;
  const bool early_tree = s2 != s && A.ipatch_id != nullptr && npat > 0;
  double* d_cache = nullptr;
  const int tb = 256;
  k_patch_prep<N><<<(npat + tb - 1) / tb, tb, 0, s>>>(npat, A.ixyzs, A.srccoefs, A.srcvals, d_rvw, eps_adap,
                                                      d_cms, d_rads, d_epsp, d_area);
  k_classify<<<(A.ntarg + tb - 1) / tb, tb, 0, s>>>(A.ntarg, A.ndtarg, A.targs, A.ipatch_id, A.row_ptr,
                                                    A.col_ind, d_cms, d_rads, A.rfac0, d_cnt, d_adap,
                                                    d_counts + 0, d_self, d_counts + 1, d_cls);
  // The tree cache starts after the two small kernels above (it fills every CU and would
  // hold them back) and runs beside the list building and its host read-back.
  if (early_tree) {
    d_cache = wsalloc<double>(W_CACHE, (size_t)npat * CACHE_STRIDE * ncn);
    if (!d_cache) return -2;
    HCHK(hipEventRecord(ev[14], s));
    HCHK(hipStreamWaitEvent(s2, ev[14], 0));
    HCHK(hipEventRecord(ev[9], s2));
    launch_tree_cache<N>(npat, ntri_plain * kq, d_twq, d_ptab, A.ixyzs, A.srccoefs, d_cache, s2);
    HCHK(hipGetLastError());
    HCHK(hipEventRecord(ev[8], s2));
  }
  k_scan<<<1, 1024, 0, s>>>(npat, d_cnt, d_off);
  k_scatter_fixed<<<(A.ntarg + tb - 1) / tb, tb, 0, s>>>(A.ntarg, A.row_ptr, A.col_ind, d_cls, d_off, d_fill,
                                                         d_fixed);
  int counts[4] = {0, 0, 0, 0};
  HCHK(hipMemcpyAsync(counts, d_counts, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  HCHK(hipStreamSynchronize(s));
  const int nadap = counts[0], nself = counts[1];
  HCHK(hipEventRecord(ev[1], s));
  // fixed-rule and self pairs on the side stream, after the classification
  if (nself > 0 && !A.uvs_targ) {
    fprintf(stderr, "fmm3dbie_amd: self pairs present but no uvs_targ\n");
    return -3;
  }
  if (s2 != s) HCHK(hipStreamWaitEvent(s2, ev[1], 0));
  if (nf % 4 == 0) {
    const size_t lds_fixed = sizeof(double) * (9 * NP + 8 * (size_t)nf);
    k_fixed_mfma<N, HELM><<<npat, 256, lds_fixed, s2>>>(nf, d_Pq, d_Bq, A.ixyzs, A.srccoefs, A.ndtarg, A.targs,
                                                        A.iquad, d_off, d_fixed, kp, A.wnear);
  } else {
    const size_t lds_fixed = sizeof(double) * (9 * NP + 7 * (size_t)nf);
    k_fixed<N, HELM><<<npat, 128, lds_fixed, s2>>>(nf, d_Pq, d_Bq, A.ixyzs, A.srccoefs, A.ndtarg, A.targs,
                                                   A.iquad, d_off, d_fixed, kp, A.wnear);
  }
  HCHK(hipGetLastError());
  HCHK(hipEventRecord(ev[2], s2));
  const bool self_side = s2 != s;
  auto launch_self = [&](hipStream_t ss) -> int {
    HCHK(hipEventRecord(ev[12], ss));
    if (nself > 0) {
      k_self<N, HELM><<<nself, 64, 0, ss>>>(nself, d_self, A.col_ind, A.ixyzs, A.srccoefs, A.ndtarg, A.targs,
                                            A.uvs_targ, A.iquad, d_umat, kp, A.self_nr, A.self_nt,
                                            A.self_tau_panel, A.wnear);
      HCHK(hipGetLastError());
    }
    HCHK(hipEventRecord(ev[13], ss));
    return 0;
  };
  if (self_side && launch_self(s2) != 0) return -1;
  HCHK(hipEventRecord(ev[4], s2));
  HCHK(hipEventRecord(ev[10], s));
  if (!early_tree && nadap > 0) {
    d_cache = wsalloc<double>(W_CACHE, (size_t)npat * CACHE_STRIDE * ncn);
    if (!d_cache) return -2;
  }
  // adaptive
  const int nwaves = A.adaptive_waves > 0 ? A.adaptive_waves : adaptive_slots<N, HELM>();
  if (nwaves <= 0) {
    fprintf(stderr, "fmm3dbie_amd: occupancy query for the adaptive kernel failed\n");
    return -1;
  }
  constexpr int ES = 6 + 2 * NP;
  double* d_stack = wsalloc<double>(W_STACK, (size_t)nwaves * STACK_CAP * ES);
  if (!d_stack) return -2;
  int* d_over = d_counts + 2;
  // Adaptive pairs are bucketed by patch (and, with a selector, by predicted rule class
  // first) so that consecutive work items share a patch's cached geometry.
  const bool learned = A.mlp_params != nullptr && nadap > 0;
  const bool want_feat = (learned || A.feat_out) && nadap > 0;
  float* d_X = want_feat ? wsalloc<float>(W_X, (size_t)nadap * 8) : nullptr;
  int *d_pred = nullptr, *d_boff = nullptr, *d_bcnt = nullptr, *d_pfull = nullptr, *d_qctr = nullptr;
  int *d_lev = nullptr, *d_seg = nullptr, *d_unc = nullptr;
  int2 *d_sorted = nullptr, *d_fb = nullptr;
  if (want_feat && !d_X) return -2;
  if (nadap > 0) {
    d_bcnt = wsalloc<int>(W_BCNT, MLP_OUT * npat);
    d_boff = wsalloc<int>(W_BOFF, MLP_OUT * npat + 1);
    d_sorted = wsalloc<int2>(W_SORTED, nadap);
    d_qctr = wsalloc<int>(W_QCTR, 16);
    if (A.lev_host) d_lev = wsalloc<int>(W_LEV, A.nnz);
    if (A.unc_host) d_unc = wsalloc<int>(W_UNC, A.nnz);
    if (!d_bcnt || !d_boff || !d_sorted || !d_qctr || (A.lev_host && !d_lev) || (A.unc_host && !d_unc)) return -2;
  }
  if (learned) {
    d_pred = wsalloc<int>(W_PRED, nadap);
    d_fb = wsalloc<int2>(W_FB, nadap);
    d_seg = wsalloc<int>(W_SEG, MAX_RULES + 1);
    if (A.pred_host) d_pfull = wsalloc<int>(W_PFULL, A.nnz);
    if (!d_pred || !d_fb || !d_seg || (A.pred_host && !d_pfull)) return -2;
  }
  HCHK(hipEventRecord(ev[5], s));
  if (want_feat) {
    const double kabs = HELM ? sqrt(A.zk[0] * A.zk[0] + A.zk[1] * A.zk[1]) : 0.0;
    k_features<N><<<(nadap + tb - 1) / tb, tb, 0, s>>>(nadap, d_adap, A.col_ind, A.ixyzs, A.srcvals, A.ndtarg,
                                                       A.targs, d_cms, d_rads, d_area, d_epsp, d_rvuv, kabs, d_X,
                                                       A.feat_out);
    HCHK(hipGetLastError());
  }
  if (learned) {
    if (mlp_predict_device(A.mlp_params, d_X, nadap, d_pred, nullptr, s) != 0) return -1;
    if (d_pfull) HCHK(hipMemsetAsync(d_pfull, 0xff, sizeof(int) * A.nnz, s));
  }
  if (d_lev) HCHK(hipMemsetAsync(d_lev, 0, sizeof(int) * A.nnz, s));
  if (d_unc) HCHK(hipMemsetAsync(d_unc, 0, sizeof(int) * A.nnz, s));
  if (nadap > 0) {
    HCHK(hipMemsetAsync(d_bcnt, 0, sizeof(int) * MLP_OUT * npat, s));
    HCHK(hipMemsetAsync(d_qctr, 0, sizeof(int) * 16, s));
    k_bucket_count<<<(nadap + tb - 1) / tb, tb, 0, s>>>(nadap, d_adap, d_pred, A.col_ind, npat, d_bcnt);
    k_scan<<<1, 1024, 0, s>>>(MLP_OUT * npat, d_bcnt, d_boff);
    HCHK(hipMemsetAsync(d_bcnt, 0, sizeof(int) * MLP_OUT * npat, s));
    k_bucket_scatter<<<(nadap + tb - 1) / tb, tb, 0, s>>>(nadap, d_adap, d_pred, A.col_ind, npat, d_boff, d_bcnt,
                                                          d_sorted, d_pfull);
    HCHK(hipGetLastError());
  }
  // rule classes: the plain rule alone, or with a selector every class (class c: XG
  // degree class_degree(nqorder, c), level cap from route_lmax) whose segment of the
  // sorted list is not empty
  AdapRules R = {};
  int seg[MAX_RULES + 1] = {0};
  int nroute = 0;  // routed classes with pairs
  if (learned) {
    k_seg_bounds<<<1, 64, 0, s>>>(MAX_RULES, npat, d_boff, d_seg);
    HCHK(hipMemcpyAsync(seg, d_seg, sizeof(int) * (MAX_RULES + 1), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    R.nrules = MAX_RULES;
    for (int c = 0; c <= MAX_RULES; c++) R.seg[c] = seg[c];
  } else {
    R.nrules = 1;
    R.seg[0] = 0;
    R.seg[1] = nadap;
    R.kq[0] = kq;
    R.lmax[0] = NLMAX;
    R.npar[0] = npar_plain;
    R.xg[0] = d_xg;
    R.ptab[0] = d_ptab;
  }
  HCHK(hipEventRecord(ev[6], s));
  if (nadap > 0) {
    if (early_tree) {
      HCHK(hipStreamWaitEvent(s, ev[8], 0));
    } else {
      HCHK(hipEventRecord(ev[9], s));
      launch_tree_cache<N>(npat, ntri_plain * kq, d_twq, d_ptab, A.ixyzs, A.srccoefs, d_cache, s);
      HCHK(hipEventRecord(ev[8], s));
    }
    HCHK(hipEventRecord(ev[11], s));
    if (learned) {
      // tree caches of the routed classes (the plain class shares the plain cache)
      for (int c = 0; c < MAX_RULES; c++) {
        const int q = class_degree(nqorder, c);
        const RuleTables* RT = nullptr;
        if (rule_tables<N>(q, &RT, s) != 0) return -2;
        R.kq[c] = RT->kq;
        R.xg[c] = RT->xg;
        R.ptab[c] = RT->ptab;
        R.lmax[c] = A.route_lmax ? A.route_lmax[c] : NLMAX;
        if (R.lmax[c] < 1 || R.lmax[c] > NLMAX) R.lmax[c] = NLMAX;
        R.cache[c] = nullptr;
        if (seg[c + 1] == seg[c]) continue;
        nroute++;
        if (q == nqorder) {
          R.npar[c] = npar_plain;
          R.cache[c] = d_cache;
          continue;
        }
        // the class's pairs evaluate no triangle below depth lmax (deeper ones go back to
        // the plain integrator); its cache holds depths 0 .. min(2, lmax): the pairs routed
        // to a rule of another degree mostly end at levels 1 and 2 (on the wavy torus, 99%
        // of them), and a deeper cache would cost every generation 4 times as much for
        // the few refinement tests below it
        R.npar[c] = cache_parents(R.lmax[c] < 2 ? R.lmax[c] : 2);
        const int ntri = cache_tris(R.npar[c]);
        double* cc = wsalloc<double>(W_RCACHE0 + c, (size_t)npat * CACHE_STRIDE * ntri * RT->kq);
        if (!cc) return -2;
        launch_tree_cache<N>(npat, ntri * RT->kq, RT->twq, RT->ptab, A.ixyzs, A.srccoefs, cc, s);
        R.cache[c] = cc;
      }
      HCHK(hipGetLastError());
    } else {
      R.cache[0] = d_cache;
    }
    HCHK(hipEventRecord(ev[7], s));
    const int g = nadap < nwaves ? nadap : nwaves;
    launch_adaptive<N, HELM>(g, s, nadap, nullptr, d_sorted, R, d_fb, d_counts + 4, d_qctr, A, d_epsp, d_umat,
                             kp, d_stack, d_nref, d_lev, d_unc, d_over);
    HCHK(hipGetLastError());
    if (learned) {
      // pairs whose tree went past their class's level: the plain rule and level cap
      AdapRules P0 = {};
      P0.nrules = 1;
      P0.seg[0] = 0;
      P0.seg[1] = 0x7fffffff;
      P0.kq[0] = kq;
      P0.lmax[0] = NLMAX;
      P0.npar[0] = npar_plain;
      P0.xg[0] = d_xg;
      P0.ptab[0] = d_ptab;
      P0.cache[0] = d_cache;
      launch_adaptive<N, HELM>(g, s, 0, d_counts + 4, d_fb, P0, nullptr, nullptr, d_qctr + 8, A, d_epsp,
                               d_umat, kp, d_stack, d_nref, d_lev, d_unc, d_over);
      HCHK(hipGetLastError());
    }
  } else {
    if (!early_tree) {
      HCHK(hipEventRecord(ev[9], s));
      HCHK(hipEventRecord(ev[8], s));
    }
    HCHK(hipEventRecord(ev[11], s));
    HCHK(hipEventRecord(ev[7], s));
  }
  HCHK(hipEventRecord(ev[3], s));
  if (s2 != s) HCHK(hipStreamWaitEvent(s, ev[4], 0));  // the side stream's kernels
  if (!self_side && launch_self(s) != 0) return -1;
  int counts8[8];
  HCHK(hipMemcpyAsync(counts8, d_counts, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
  HCHK(hipStreamSynchronize(s));
  for (int i = 0; i < 4; i++) counts[i] = counts8[i];
  if (d_pfull) HCHK(hipMemcpy(A.pred_host, d_pfull, sizeof(int) * A.nnz, hipMemcpyDeviceToHost));
  if (d_lev) HCHK(hipMemcpy(A.lev_host, d_lev, sizeof(int) * A.nnz, hipMemcpyDeviceToHost));
  if (d_unc) HCHK(hipMemcpy(A.unc_host, d_unc, sizeof(int) * A.nnz, hipMemcpyDeviceToHost));
  if (st) {
    st->nself = nself;
    st->nadap = nadap;
    st->nfixed = (int)(A.nnz - nself - nadap);
    st->overflow = counts[2];
    st->nf = nf;
    st->kq = kq;
    auto elapsed = [&](int a, int b) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
      return (double)ms;
    };
    st->ms_prep = elapsed(0, 1);
    st->ms_fixed = elapsed(1, 2);
    st->ms_tables = elapsed(10, 5);
    st->ms_select = elapsed(5, 6);
    st->ms_tree = elapsed(9, 8);
    st->ms_shallow = elapsed(11, 7);
    st->ms_adap = elapsed(7, 3);
    st->ms_self = elapsed(12, 13);
    st->nroute = nroute;
    st->cache_tris = ntri_plain;
    st->nfallback = counts8[4];
    for (int c = 0; c < MAX_RULES; c++) st->ncls[c] = learned ? seg[c + 1] - seg[c] : (c == 0 ? nadap : 0);
    st->nref_total = 0;
    if (d_nref && A.nref_host) HCHK(hipMemcpy(A.nref_host, d_nref, sizeof(int) * A.nnz, hipMemcpyDeviceToHost));
    if (A.class_host) HCHK(hipMemcpy(A.class_host, d_cls, sizeof(int) * A.nnz, hipMemcpyDeviceToHost));
  }
  (void)d_umat;
  return 0;
}

template <bool HELM>
int dispatch(const NearQuadArgs& A, NearQuadStats* st, hipStream_t s) {
  switch (A.norder) {
    case 1: return run_order<1, HELM>(A, st, s);
    case 2: return run_order<2, HELM>(A, st, s);
    case 3: return run_order<3, HELM>(A, st, s);
    case 4: return run_order<4, HELM>(A, st, s);
    case 5: return run_order<5, HELM>(A, st, s);
    case 6: return run_order<6, HELM>(A, st, s);
    case 7: return run_order<7, HELM>(A, st, s);
    case 8: return run_order<8, HELM>(A, st, s);
    case 9: return run_order<9, HELM>(A, st, s);
    default:
      // one moment component per lane of a wavefront: (N + 1)(N + 2) / 2 <= 64
      fprintf(stderr, "fmm3dbie_amd: patch order %d not supported by the near quadrature (1..9)\n", A.norder);
      return -4;
  }
}

}  // namespace

int fmm_nearquad_device(const NearQuadArgs& A, NearQuadStats* st, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_quad_mutex);
  if (ensure_constants() != 0) return -1;
  if (A.self_nr < 1 || A.self_nr > 16 || A.self_nt < 1 || A.self_nt > 16) {
    fprintf(stderr, "fmm3dbie_amd: self rule size %d x %d out of range (1..16)\n", A.self_nr, A.self_nt);
    return -5;
  }
  if (set_self_gl(A.self_nr, A.self_nt) != 0) return -1;
  return A.helm ? dispatch<true>(A, st, s) : dispatch<false>(A, st, s);
}

int fmm_koorn_constants() {
  std::lock_guard<std::mutex> lock(g_quad_mutex);
  return ensure_constants();
}

void fmm_nearquad_set_side_stream(int enable) {
  std::lock_guard<std::mutex> lock(g_quad_mutex);
  g_use_side = enable ? 1 : 0;
}

void fmm_nearquad_release() {
  std::lock_guard<std::mutex> lock(g_quad_mutex);
  for (auto& kv : g_side) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first);
    (void)hipStreamSynchronize(kv.second);
    (void)hipStreamDestroy(kv.second);
    (void)hipSetDevice(cur);
  }
  g_side.clear();
  for (auto& kv : g_events) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first);
    for (auto& e : kv.second) (void)hipEventDestroy(e);
    (void)hipSetDevice(cur);
  }
  g_events.clear();
  for (auto& kv : g_ws)
    for (auto& b : kv.second)
      if (b.p) (void)hipFree(b.p);
  g_ws.clear();
}
