// Layer-potential apply on gfx950 (lpcomp_{helm,lap}_comb_dir_addsub,
// helm_comb_dir.f:856-1208; lap_comb_dir.f90:806-1188):
//
//   pot = sum_{all oversampled j} K(x_i, y_j) s_j      (FMM3D call in the reference)
//       + sum_{near pairs} wnear . sigma                 (CSR correction, :1138-1151)
//       - sum_{near pairs} sum_{j in patch} K(x_i,y_j) s_j (subtraction, :1156-1193)
// with s_j = sigma_over_j * whtsover_j and K the combined kernel (alpha S + beta D).
//
// The FMM is replaced by a tiled all-pairs sum: 256 targets per workgroup, source tiles
// staged in LDS.  Two modes:
//   ADDSUB : all sources, then an explicit subtraction kernel (the reference's algebra);
//   MASKED : sources of the target's near patches are skipped inside the tiled sum
//            (algebraically identical, no cancellation; used for FP32).
// Sources are sorted by patch, so the near-patch test is a moving pointer into the
// target's sorted near list.  Sources with |x-y| <= thresh are skipped in every sum
// (get_fmm_thresh, src/common/get_fmm_thresh.f:81).
// Laplace runs through the same code with k = 0 and real densities (the kernel then
// reduces exactly to (beta d/r^3 + alpha/r)/(4 pi)).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "apply.hpp"
#include "common.hpp"
#include "host_basis.hpp"

namespace {

constexpr int TPB = 256;

template <typename T>
struct V2 {
  T x, y;
};

// sigma_over = I * sigma per patch; strengths q = alpha*s, m = beta*s, s = sigma_over*w.
template <typename T>
__global__ void k_oversample(int64_t nptso, const int* __restrict__ opatch, const int64_t* __restrict__ ixyzs,
                             const int64_t* __restrict__ ixyzso, const int64_t* __restrict__ ioff,
                             const int* __restrict__ np_of, const double* __restrict__ imat,
                             const V2<T>* __restrict__ sigma, int nrhs, int64_t ldsig,
                             const T* __restrict__ wover, double ar, double ai, double br, double bi,
                             V2<T>* __restrict__ qstr, V2<T>* __restrict__ mstr, int64_t ldstr) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nptso) return;
  const int ip = opatch[i];
  const int p = np_of[ip];
  const int64_t li = i - (ixyzso[ip] - 1);
  const double* row = imat + ioff[ip] + li * p;
  const int64_t s0 = ixyzs[ip] - 1;
  const double w = (double)wover[i];
  for (int r = 0; r < nrhs; r++) {
    const V2<T>* sg = sigma + r * ldsig + s0;
    double sr = 0.0, si = 0.0;
    for (int l = 0; l < p; l++) {
      sr = fma(row[l], (double)sg[l].x, sr);
      si = fma(row[l], (double)sg[l].y, si);
    }
    sr *= w;
    si *= w;
    qstr[r * ldstr + i] = {(T)(ar * sr - ai * si), (T)(ar * si + ai * sr)};
    mstr[r * ldstr + i] = {(T)(br * sr - bi * si), (T)(br * si + bi * sr)};
  }
}

// 16-byte LDS word of T values (4 floats or 2 doubles)
template <typename T>
struct Word16;
template <>
struct Word16<float> {
  using type = float4;
  static __device__ __forceinline__ float4 pack(const float* v) { return make_float4(v[0], v[1], v[2], v[3]); }
  static __device__ __forceinline__ void unpack(const float4 w, float* v) {
    v[0] = w.x; v[1] = w.y; v[2] = w.z; v[3] = w.w;
  }
};
template <>
struct Word16<double> {
  using type = double2;
  static __device__ __forceinline__ double2 pack(const double* v) { return make_double2(v[0], v[1]); }
  static __device__ __forceinline__ void unpack(const double2 w, double* v) {
    v[0] = w.x; v[1] = w.y;
  }
};

// e^{ikr}/r and the dipole factor; T = double or float.  GEN: complex wavenumber, or (FP32)
// phases beyond the hardware sine's range, reduced to one revolution first; otherwise the
// FP32 phase k r / (2 pi) = r * krev goes to the hardware sine and cosine directly
// (valid for |k r| / (2 pi) < 256, which the plan checks against its bounding box).
template <typename T, bool GEN>
__device__ __forceinline__ void pair_kernel(T dx, T dy, T dz, T nx, T ny, T nz, T kr, T krev, T ki, T& er, T& ei,
                                            T& tdr, T& tdi, T& r2out) {
  const T r2 = dx * dx + dy * dy + dz * dz;
  r2out = r2;
  T rinv, r;
  if constexpr (sizeof(T) == 4) {
    rinv = __frsqrt_rn(r2);
    r = r2 * rinv;
  } else {
    rinv = rsqrt_nr(r2);
    r = r2 * rinv;
  }
  const T d = dx * nx + dy * ny + dz * nz;
  T s, c;
  if constexpr (sizeof(T) == 4) {
    T rev = r * krev;
    if constexpr (GEN) rev -= rintf(rev);
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
  } else {
    (void)krev;
    sincos_red(kr * r, s, c);
  }
  T mag = rinv;
  if constexpr (GEN) {
    if constexpr (sizeof(T) == 4) mag *= __expf(-ki * r);
    else mag *= exp(-ki * r);
  }
  er = c * mag;
  ei = s * mag;
  // dipole factor t (1 - i k r), t = d / r^2, k = kr + i ki: (1 + ki r) - i kr r, and
  // t kr r = kr d / r
  const T u = d * rinv;
  const T t = u * rinv;
  tdr = GEN ? t * ((T)1 + ki * r) : t;
  tdi = -kr * u;
}

// Far-sum pair factors with 1/r left to the caller: (c, s) = e^{i kr r} (times e^{-ki r}
// when GEN), the dipole factor t (1 - i k r) = (tdr, tdi) with t = d / r^2, from r^2, d =
// (x - y).n and rinv = 1/r.  A pair to be skipped passes rinv = 0: r is then 0 and every
// factor stays finite, so its contribution (factors x rinv) is an exact zero.
template <typename T, bool GEN>
__device__ __forceinline__ void pair_factors(T r2, T rinv, T d, T kr, T krev, T ki, T& c, T& s, T& tdr, T& tdi) {
  const T r = r2 * rinv;
  if constexpr (sizeof(T) == 4) {
    T rev = r * krev;
    if constexpr (GEN) rev -= rintf(rev);
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
  } else {
    (void)krev;
    sincos_red(kr * r, s, c);
  }
  if constexpr (GEN) {
    T mag;
    if constexpr (sizeof(T) == 4) mag = __expf(-ki * r);
    else mag = exp(-ki * r);
    c *= mag;
    s *= mag;
  }
  const T u = d * rinv;
  const T t = u * rinv;
  tdr = GEN ? t * ((T)1 + ki * r) : t;
  tdi = -kr * u;
}

// Tiled far sum.  pot[i] (+)= sum_j [q_j + m_j * t_ij (1-ikr)] e^{ikr}/r / (4 pi)
// MASK: skip sources whose patch is in the target's sorted near list.  Sources are sorted
// by patch, so each thread keeps the source interval [ea, eb) of the next near patch
// (patch p owns sources ixyzso[p]-1 .. ixyzso[p+1]-2) and moves it forward when the tile
// passes eb; the pair loop itself has no branches (a skipped pair adds an exact zero).
template <typename T, bool MASK, bool GEN, int NRHS>
__global__ void __launch_bounds__(TPB) k_farsum(int64_t ntarg, const T* __restrict__ tx, const T* __restrict__ ty,
                                                const T* __restrict__ tz, int64_t nsrc,
                                                const T* __restrict__ sx, const T* __restrict__ sy,
                                                const T* __restrict__ sz, const T* __restrict__ snx,
                                                const T* __restrict__ sny, const T* __restrict__ snz,
                                                const int64_t* __restrict__ ixyzso, const V2<T>* __restrict__ qstr,
                                                const V2<T>* __restrict__ mstr, int64_t ldstr,
                                                const int64_t* __restrict__ row_ptr,
                                                const int* __restrict__ near_sorted, T kr, T krev, T ki, T thresh2,
                                                int64_t chunk, V2<T>* __restrict__ part, int64_t ldpart) {
  // One LDS record per source of the tile: x, y, z, nx, ny, nz, then (q, m) of each
  // right-hand side, padded to whole 16-byte words, so that a pair's operands are a few
  // 16-byte reads at immediate offsets from one address (all lanes read the same record).
  using W16 = typename Word16<T>::type;
  constexpr int PER = 16 / (int)sizeof(T);            // values per 16-byte word
  constexpr int NV = 6 + 4 * NRHS;                    // values per source
  constexpr int REC = (NV + PER - 1) / PER;           // 16-byte words per source
  __shared__ W16 tile[TPB * REC];
  // blockIdx.z selects a group of NRHS right-hand sides: the geometry and e^{ikr} of a
  // pair are computed once per group
  const int64_t rg = blockIdx.z;
  qstr += rg * NRHS * ldstr;
  mstr += rg * NRHS * ldstr;
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  const bool act = i < ntarg;
  const T x = act ? tx[i] : (T)0, y = act ? ty[i] : (T)0, z = act ? tz[i] : (T)0;
  T accr[NRHS], acci[NRHS];
#pragma unroll
  for (int r = 0; r < NRHS; r++) { accr[r] = 0; acci[r] = 0; }
  const int64_t jbeg = blockIdx.y * chunk;
  const int64_t jend = jbeg + chunk < nsrc ? jbeg + chunk : nsrc;
  constexpr int64_t NONE = INT64_MAX;
  int64_t nptr = 0, nend = 0, ea = NONE, eb = NONE;
  // first near patch of the target whose sources end after j
  auto advance = [&](int64_t j) {
    ea = NONE;
    eb = NONE;
    while (nptr < nend) {
      const int p = near_sorted[nptr];
      const int64_t e = ixyzso[p + 1] - 1;
      if (e > j) {
        ea = ixyzso[p] - 1;
        eb = e;
        break;
      }
      nptr++;
    }
  };
  if (MASK && act && jbeg < jend) {
    nptr = row_ptr[i] - 1;
    nend = row_ptr[i + 1] - 1;
    advance(jbeg);
  }
  for (int64_t base = jbeg; base < jend; base += TPB) {
    const int64_t j = base + threadIdx.x;
    if (j < jend) {
      T v[REC * PER];
      v[0] = sx[j]; v[1] = sy[j]; v[2] = sz[j];
      v[3] = snx[j]; v[4] = sny[j]; v[5] = snz[j];
#pragma unroll
      for (int r = 0; r < NRHS; r++) {
        const V2<T> q = qstr[r * ldstr + j], m = mstr[r * ldstr + j];
        v[6 + 4 * r] = q.x; v[7 + 4 * r] = q.y; v[8 + 4 * r] = m.x; v[9 + 4 * r] = m.y;
      }
#pragma unroll
      for (int e = NV; e < REC * PER; e++) v[e] = (T)0;
#pragma unroll
      for (int w = 0; w < REC; w++) tile[threadIdx.x * REC + w] = Word16<T>::pack(v + w * PER);
    }
    __syncthreads();
    const int cnt = jend - base < TPB ? (int)(jend - base) : TPB;
    // FULL: no source of the tile belongs to a near patch of any target of the wavefront
    // (the next near interval of every lane starts at or after the tile's end), so every
    // pair is summed and the per-pair test and selects drop out
    auto pairs = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll 2
      for (int k = 0; k < cnt; k++) {
        const int64_t jk = base + k;
        if (MASK && !FULL && jk >= eb) advance(jk);
        T v[REC * PER];
#pragma unroll
        for (int w = 0; w < REC; w++) Word16<T>::unpack(tile[k * REC + w], v + w * PER);
        const T dx = x - v[0], dy = y - v[1], dz = z - v[2];
        const T r2 = dx * dx + dy * dy + dz * dz;
        T rinv;
        if constexpr (sizeof(T) == 4) rinv = __frsqrt_rn(r2);
        else rinv = rsqrt_nr(r2);
        // with the mask, sources within thresh of the target belong to its near patches
        if (!FULL) rinv = (MASK ? jk < ea : r2 > thresh2) ? rinv : (T)0;
        const T d = dx * v[3] + dy * v[4] + dz * v[5];
        T c, sn, tdr, tdi;
        pair_factors<T, GEN>(r2, rinv, d, kr, krev, ki, c, sn, tdr, tdi);
#pragma unroll
        for (int r = 0; r < NRHS; r++) {
          const V2<T> q = {v[6 + 4 * r], v[7 + 4 * r]}, m = {v[8 + 4 * r], v[9 + 4 * r]};
          // inner = q + m*(tdr + i tdi); contribution = inner * (c + i sn) / r
          const T ir = q.x + m.x * tdr - m.y * tdi;
          const T ii = q.y + m.x * tdi + m.y * tdr;
          const T zr = ir * c - ii * sn, zi = ir * sn + ii * c;
          accr[r] += zr * rinv;
          acci[r] += zi * rinv;
        }
      }
    };
    // (FP64 with several right-hand sides keeps the one loop: the second copy would cost
    // it occupancy)
    constexpr bool TRY_FULL = MASK && (sizeof(T) == 4 || NRHS == 1);
    bool full = false;
    if (TRY_FULL) {
      if (base >= eb) advance(base);  // the state the first pair of the tile would reach
      full = __all(ea >= base + cnt) != 0;  // wave-uniform
    }
    if (full)
      pairs(std::true_type{});
    else
      pairs(std::false_type{});
    __syncthreads();
  }
  if (act) {
#pragma unroll
    for (int r = 0; r < NRHS; r++) part[(int64_t)blockIdx.y * ldpart + (rg * NRHS + r) * ntarg + i] = {accr[r], acci[r]};
  }
}

// Far sum of 16 right-hand sides on the matrix cores (FP32, near patches masked): with
// G = e^{ikr}/r and H = t (1 - ikr) e^{ikr}/r for a pair (target, source), the group's
// potentials are the complex products pot = G q + H m of the target x source kernel
// matrices with the source x 16 strength matrices.  A wavefront takes 16 targets; in a
// step of 4 sources lane l evaluates the pair (target l % 16, source l / 16) once, and
// v_mfma_f32_16x16x4f32 multiplies 16 x 4 kernel blocks by 4 x 16 strength blocks (lane
// l supplies source l / 16, right-hand side l % 16).  Each complex product takes three
// real ones (Gauss): a1 = Re G Re q + Re H Re m, a2 = Im G Im q + Im H Im m,
// a3 = (Re G + Im G)(Re q + Im q) + (Re H + Im H)(Re m + Im m), so Re pot = a1 - a2 and
// Im pot = a3 - a1 - a2: six matrix products per step instead of eight.  Accumulator
// element r of lane l is target 4 (l / 16) + r, right-hand side l % 16.
// The workgroup stages FT sources at a time in LDS: geometry (x, y, z, nx, ny, nz) and
// the (q, m) of the 16 right-hand sides of each source.
constexpr int FT = 64;         // sources per LDS tile (18 KB of LDS per workgroup with FW = 4)
constexpr int FW = 4;          // wavefronts per workgroup (16 targets each)
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <bool GEN>
__global__ void __launch_bounds__(64 * FW) k_farsum_mfma(int64_t ntarg, const float* __restrict__ tx,
                                                        const float* __restrict__ ty, const float* __restrict__ tz,
                                                        int64_t nsrc, const float* __restrict__ sx,
                                                        const float* __restrict__ sy, const float* __restrict__ sz,
                                                        const float* __restrict__ snx, const float* __restrict__ sny,
                                                        const float* __restrict__ snz,
                                                        const int64_t* __restrict__ ixyzso,
                                                        const V2<float>* __restrict__ qstr,
                                                        const V2<float>* __restrict__ mstr, int64_t ldstr,
                                                        const int64_t* __restrict__ row_ptr,
                                                        const int* __restrict__ near_sorted, float kr, float krev,
                                                        float ki, int64_t chunk, V2<float>* __restrict__ part,
                                                        int64_t ldpart) {
  __shared__ float4 geo[FT * 2];   // x y z nx | ny nz - -
  __shared__ float4 str[FT * 16];  // (q re, q im, m re, m im) of source j, right-hand side r at j * 16 + r
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rg = blockIdx.z;
  qstr += (int64_t)rg * 16 * ldstr;
  mstr += (int64_t)rg * 16 * ldstr;
  const int64_t t0 = (blockIdx.x * (int64_t)FW + wv) * 16;
  const int64_t i = t0 + (lane & 15);  // this lane's target
  const bool act = i < ntarg;
  const float x = act ? tx[i] : 0.0f, y = act ? ty[i] : 0.0f, z = act ? tz[i] : 0.0f;
  const int ks = lane >> 4;  // this lane's source within a step of 4
  f32x4 a1 = {0.0f, 0.0f, 0.0f, 0.0f}, a2 = {0.0f, 0.0f, 0.0f, 0.0f}, a3 = {0.0f, 0.0f, 0.0f, 0.0f};
  const int64_t jbeg = blockIdx.y * chunk;
  const int64_t jend = jbeg + chunk < nsrc ? jbeg + chunk : nsrc;
  constexpr int64_t NONE = INT64_MAX;
  int64_t nptr = 0, nend = 0, ea = NONE, eb = NONE;
  auto advance = [&](int64_t j) {
    ea = NONE;
    eb = NONE;
    while (nptr < nend) {
      const int p = near_sorted[nptr];
      const int64_t e = ixyzso[p + 1] - 1;
      if (e > j) {
        ea = ixyzso[p] - 1;
        eb = e;
        break;
      }
      nptr++;
    }
  };
  if (act && jbeg < jend) {
    nptr = row_ptr[i] - 1;
    nend = row_ptr[i + 1] - 1;
    advance(jbeg);
  }
  for (int64_t base = jbeg; base < jend; base += FT) {
    const int cnt = jend - base < FT ? (int)(jend - base) : FT;
    __syncthreads();
    for (int e = threadIdx.x; e < FT; e += 64 * FW) {
      const int64_t j = base + e;
      const bool ok = e < cnt;
      geo[2 * e] = ok ? make_float4(sx[j], sy[j], sz[j], snx[j]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      geo[2 * e + 1] = ok ? make_float4(sny[j], snz[j], 0.0f, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // strengths: consecutive threads read consecutive sources of one right-hand side
    for (int e = threadIdx.x; e < FT * 16; e += 64 * FW) {
      const int jl = e % FT, r = e / FT;
      float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (jl < cnt) {
        const V2<float> q = qstr[r * ldstr + base + jl], m = mstr[r * ldstr + base + jl];
        v = make_float4(q.x, q.y, m.x, m.y);
      }
      str[jl * 16 + r] = v;
    }
    __syncthreads();
    // sources past the tile's count carry zero strengths (and are masked below)
    for (int k = 0; k < cnt; k += 4) {
      const int jl = k + ks;
      const int64_t jk = base + jl;
      if (jk >= eb) advance(jk);
      const float4 g0 = geo[2 * jl], g1 = geo[2 * jl + 1];
      const float dx = x - g0.x, dy = y - g0.y, dz = z - g0.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      float rinv = __frsqrt_rn(r2);
      rinv = (act && jl < cnt && jk < ea) ? rinv : 0.0f;
      const float d = dx * g0.w + dy * g1.x + dz * g1.y;
      float c, sn, tdr, tdi;
      pair_factors<float, GEN>(r2, rinv, d, kr, krev, ki, c, sn, tdr, tdi);
      const float gr = c * rinv, gi = sn * rinv;
      const float hr = tdr * gr - tdi * gi, hi = tdr * gi + tdi * gr;
      const float4 sv = str[(k + ks) * 16 + (lane & 15)];  // source k + l / 16, right-hand side l % 16
      a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(gr, sv.x, a1, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(hr, sv.z, a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f32_16x16x4f32(gi, sv.y, a2, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f32_16x16x4f32(hi, sv.w, a2, 0, 0, 0);
      a3 = __builtin_amdgcn_mfma_f32_16x16x4f32(gr + gi, sv.x + sv.y, a3, 0, 0, 0);
      a3 = __builtin_amdgcn_mfma_f32_16x16x4f32(hr + hi, sv.z + sv.w, a3, 0, 0, 0);
    }
  }
  // accumulator element r of lane l: target 4 (l / 16) + r, right-hand side l % 16
  const int rhs = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int64_t it = t0 + 4 * (lane >> 4) + r;
    if (it < ntarg)
      part[(int64_t)blockIdx.y * ldpart + (int64_t)(rg * 16 + rhs) * ntarg + it] = {a1[r] - a2[r], a3[r] - a1[r] - a2[r]};
  }
}

// pot[r][i] = (sum over source splits of part) / (4 pi)
template <typename T>
__global__ void k_farsum_reduce(int64_t ntarg, int nrhs, int nsplit, const V2<T>* __restrict__ part,
                                int64_t ldpart, V2<T>* __restrict__ pot, int64_t ldpot) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= ntarg || r >= nrhs) return;
  T sr = 0, si = 0;
  for (int sp = 0; sp < nsplit; sp++) {
    const V2<T> v = part[(int64_t)sp * ldpart + r * ntarg + i];
    sr += v.x;
    si += v.y;
  }
  pot[r * ldpot + i] = {(T)(sr * (T)OVER4PI), (T)(si * (T)OVER4PI)};
}

// Near subtraction (ADDSUB mode): one wavefront per target, lanes over the oversampled
// points of each near patch.
__global__ void k_near_sub(int64_t ntarg, const double* __restrict__ tx, const double* __restrict__ ty,
                           const double* __restrict__ tz, const double* __restrict__ sx,
                           const double* __restrict__ sy, const double* __restrict__ sz,
                           const double* __restrict__ snx, const double* __restrict__ sny,
                           const double* __restrict__ snz, const V2<double>* __restrict__ qstr,
                           const V2<double>* __restrict__ mstr, int64_t ldstr, int nrhs,
                           const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ind,
                           const int64_t* __restrict__ ixyzso, double kr, double ki, double thresh2,
                           V2<double>* __restrict__ pot, int64_t ldpot) {
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= ntarg) return;
  const double x = tx[i], y = ty[i], z = tz[i];
  for (int r = 0; r < nrhs; r++) {
    double sr = 0.0, si = 0.0;
    for (int64_t jj = row_ptr[i] - 1; jj < row_ptr[i + 1] - 1; jj++) {
      const int64_t ip = col_ind[jj] - 1;
      for (int64_t j = ixyzso[ip] - 1 + lane; j < ixyzso[ip + 1] - 1; j += 64) {
        double er, ei, tdr, tdi, r2;
        pair_kernel<double, true>(x - sx[j], y - sy[j], z - sz[j], snx[j], sny[j], snz[j], kr, 0.0, ki, er, ei, tdr,
                                  tdi, r2);
        if (r2 <= thresh2) continue;
        const V2<double> q = qstr[r * ldstr + j], m = mstr[r * ldstr + j];
        const double ir = q.x + m.x * tdr - m.y * tdi;
        const double ii = q.y + m.x * tdi + m.y * tdr;
        sr += ir * er - ii * ei;
        si += ir * ei + ii * er;
      }
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
      pot[r * ldpot + i].x -= sr * OVER4PI;
      pot[r * ldpot + i].y -= si * OVER4PI;
    }
  }
}

// General CSR correction (any patch sizes): one thread per target row.
template <typename T, typename TW>
__global__ void k_csr_gen(int64_t ntarg, const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ind,
                          const int64_t* __restrict__ iquad, const int64_t* __restrict__ ixyzs,
                          const V2<TW>* __restrict__ wnear, const V2<T>* __restrict__ sigma, int64_t ldsig,
                          int nrhs, V2<T>* __restrict__ pot, int64_t ldpot) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ntarg) return;
  for (int r = 0; r < nrhs; r++) {
    double sr = 0.0, si = 0.0;
    for (int64_t jj = row_ptr[i] - 1; jj < row_ptr[i + 1] - 1; jj++) {
      const int64_t ip = col_ind[jj] - 1;
      const int64_t s0 = ixyzs[ip] - 1, p = ixyzs[ip + 1] - ixyzs[ip], q0 = iquad[jj] - 1;
      for (int64_t l = 0; l < p; l++) {
        const V2<TW> w = wnear[q0 + l];
        const V2<T> s = sigma[r * ldsig + s0 + l];
        sr += (double)w.x * (double)s.x - (double)w.y * (double)s.y;
        si += (double)w.x * (double)s.y + (double)w.y * (double)s.x;
      }
    }
    pot[r * ldpot + i].x += (T)sr;
    pot[r * ldpot + i].y += (T)si;
  }
}

// Far-sum reduction fused with the CSR near correction for all right-hand sides (uniform
// patch size p): one wavefront per target row, lanes over the row's contiguous wnear
// segment; every wnear element is loaded once per group of RB right-hand sides.
//   pot[r][i] = (sum over source splits of part[r][i]) / (4 pi) + sum_e wnear_e sigma_r(src_e)
// PC > 0: the patch size as a compile-time constant (the orders 1..10), else p_rt.
template <typename T, typename TW, int RB, int PC>
__global__ void k_reduce_csr(int64_t ntarg, int nrhs, int nsplit, const V2<T>* __restrict__ part, int64_t ldpart,
                             const int64_t* __restrict__ row_ptr, const int* __restrict__ csr_src,
                             const int64_t* __restrict__ iquad, int p_rt,
                             const V2<TW>* __restrict__ wnear, const V2<T>* __restrict__ sigma, int64_t ldsig,
                             V2<T>* __restrict__ pot, int64_t ldpot) {
  const int p = PC > 0 ? PC : p_rt;
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= ntarg) return;
  const int64_t j0 = row_ptr[i] - 1, j1 = row_ptr[i + 1] - 1;
  const int64_t q0 = j0 < j1 ? iquad[j0] - 1 : 0;
  const int64_t ne = (j1 - j0) * (int64_t)p;
  for (int rb = 0; rb < nrhs; rb += RB) {
    double ar[RB], ai[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) { ar[r] = 0.0; ai[r] = 0.0; }
    for (int e = lane; e < (int)ne; e += 64) {
      const int k = e / p;  // patch entry of the row (compile-time divisor for the orders 1..10)
      const int64_t src = csr_src[j0 + k] + (e - k * p);
      const V2<TW> w = wnear[q0 + e];
#pragma unroll
      for (int r = 0; r < RB; r++) {
        if (rb + r < nrhs) {
          const V2<T> sg = sigma[(rb + r) * ldsig + src];
          ar[r] += (double)w.x * (double)sg.x - (double)w.y * (double)sg.y;
          ai[r] += (double)w.x * (double)sg.y + (double)w.y * (double)sg.x;
        }
      }
    }
    double mr = 0.0, mi = 0.0;
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const double sr = wave_sum(ar[r]), si = wave_sum(ai[r]);
      if (lane == r) { mr = sr; mi = si; }
    }
    const int r = rb + lane;
    if (lane < RB && r < nrhs) {
      T fr = 0, fi = 0;
      for (int sp = 0; sp < nsplit; sp++) {
        const V2<T> v = part[(int64_t)sp * ldpart + r * ntarg + i];
        fr += v.x;
        fi += v.y;
      }
      pot[r * ldpot + i] = {(T)(fr * (T)OVER4PI + (T)mr), (T)(fi * (T)OVER4PI + (T)mi)};
    }
  }
}

// Sum of the far-sum partials over the source splits, in place into split 0: a workgroup
// of 256 threads takes 64 targets, thread (t, g) adds splits g, g + 4, ... of target t
// (loads coalesced over t) and the four groups meet in LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_split_sum(int64_t ntarg, int nsplit, V2<T>* __restrict__ part,
                                                   int64_t ldpart) {
  __shared__ V2<T> sh[4][64];
  const int t = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t i = blockIdx.x * (int64_t)64 + t;
  const int64_t r = blockIdx.y;
  T sr = 0, si = 0;
  if (i < ntarg) {
    const V2<T>* pp = part + r * ntarg + i;
#pragma unroll 4
    for (int sp = g; sp < nsplit; sp += 4) {
      const V2<T> v = pp[(int64_t)sp * ldpart];
      sr += v.x;
      si += v.y;
    }
  }
  sh[g][t] = {sr, si};
  __syncthreads();
  if (g == 0 && i < ntarg) {
    const V2<T> a = sh[1][t], b = sh[2][t], c = sh[3][t];
    part[r * ntarg + i] = {(sr + a.x) + (b.x + c.x), (si + a.y) + (b.y + c.y)};
  }
}

// constant fill (tuning densities)
template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ a, T v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_split_src(int64_t n, const double* __restrict__ s12, double* __restrict__ x, double* __restrict__ y,
                            double* __restrict__ z, double* __restrict__ nx, double* __restrict__ ny,
                            double* __restrict__ nz, int64_t stride) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* s = s12 + stride * i;
  x[i] = s[0]; y[i] = s[1]; z[i] = s[2];
  if (nx) { nx[i] = s[9]; ny[i] = s[10]; nz[i] = s[11]; }
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A* __restrict__ a, B* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

__global__ void k_real_to_cplx(int64_t n, const double* __restrict__ a, V2<double>* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = {a[i], 0.0};
}

__global__ void k_cplx_to_real(int64_t n, const V2<double>* __restrict__ a, double* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i].x;
}

template <typename T>
T* dal(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
  return (T*)p;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

#define ACHK(x)                                                                    \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "fmm3dbie_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

ApplyPlan::~ApplyPlan() {
  for (void* p : owned)
    if (p) (void)hipFree(p);
}

template <typename T>
static T* own(ApplyPlan* P, size_t n) {
  T* p = dal<T>(n);
  P->owned.push_back((void*)p);
  return p;
}

int apply_plan_build(ApplyPlan* P, const ApplyPlanInput& in, hipStream_t s) {
  P->npatches = in.npatches;
  P->npts = in.npts;
  P->nptso = in.nptso;
  P->ntarg = in.targs ? in.ntarg : in.npts;
  P->nnz = in.nnz;
  P->helm = in.helm;
  P->zk[0] = in.zk[0]; P->zk[1] = in.zk[1];
  P->alpha[0] = in.alpha[0]; P->alpha[1] = in.alpha[1];
  P->beta[0] = in.beta[0]; P->beta[1] = in.beta[1];
  const int64_t npat = in.npatches, N = P->ntarg, NO = in.nptso;
  const double* tsrc = in.targs ? in.targs : in.srcvals;  // target coordinates
  const int64_t tstride = in.targs ? in.ndtarg : 12;
  // host-side tables: interpolation matrices per (norder, nover), patch of each
  // oversampled point, sorted near lists, uniform-p check
  std::vector<int64_t> ioff(npat);
  std::vector<int> np_of(npat), opatch(NO);
  std::vector<double> imat_all;
  std::vector<std::pair<int, int>> keys;
  std::vector<int64_t> koff;
  P->uniform_p = 0;
  int p0 = -1;
  bool uni = true;
  for (int64_t i = 0; i < npat; i++) {
    const int no = (int)in.norders[i], nv = (int)in.novers[i];
    int64_t off = -1;
    for (size_t k = 0; k < keys.size(); k++)
      if (keys[k].first == no && keys[k].second == nv) off = koff[k];
    if (off < 0) {
      std::vector<double> im;
      fmmhost::interp_matrix(no, nv, im);
      off = (int64_t)imat_all.size();
      imat_all.insert(imat_all.end(), im.begin(), im.end());
      keys.push_back({no, nv});
      koff.push_back(off);
    }
    ioff[i] = off;
    np_of[i] = (int)(in.ixyzs[i + 1] - in.ixyzs[i]);
    if (p0 < 0) p0 = np_of[i];
    if (np_of[i] != p0) uni = false;
    for (int64_t j = in.ixyzso[i] - 1; j < in.ixyzso[i + 1] - 1; j++) opatch[j] = (int)i;
  }
  P->uniform_p = uni ? p0 : 0;
  std::vector<int> nsorted(in.nnz);
  for (int64_t t = 0; t < N; t++) {
    const int64_t a = in.row_ptr[t] - 1, b = in.row_ptr[t + 1] - 1;
    for (int64_t j = a; j < b; j++) nsorted[j] = (int)(in.col_ind[j] - 1);
    std::sort(nsorted.begin() + a, nsorted.begin() + b);
  }
  // threshold: 2^-51 * max bounding-box edge of sources and targets
  double mn[3], mx[3];
  for (int m = 0; m < 3; m++) mn[m] = mx[m] = in.srcover[m];
  for (int64_t i = 0; i < NO; i++)
    for (int m = 0; m < 3; m++) {
      const double v = in.srcover[12 * i + m];
      mn[m] = v < mn[m] ? v : mn[m];
      mx[m] = v > mx[m] ? v : mx[m];
    }
  for (int64_t i = 0; i < N; i++)
    for (int m = 0; m < 3; m++) {
      const double v = tsrc[tstride * i + m];
      mn[m] = v < mn[m] ? v : mn[m];
      mx[m] = v > mx[m] ? v : mx[m];
    }
  double bs = mx[0] - mn[0];
  if (mx[1] - mn[1] > bs) bs = mx[1] - mn[1];
  if (mx[2] - mn[2] > bs) bs = mx[2] - mn[2];
  P->thresh = std::ldexp(1.0, -51) * bs;
  // largest FP32 phase |k| r / (2 pi) over the box diagonal: the hardware sine takes up to
  // 256 revolutions; beyond 200 the FP32 far sum reduces the phase first
  const double diag = std::sqrt((mx[0] - mn[0]) * (mx[0] - mn[0]) + (mx[1] - mn[1]) * (mx[1] - mn[1]) +
                                (mx[2] - mn[2]) * (mx[2] - mn[2]));
  P->phase_small = !in.helm || std::fabs(in.zk[0]) * diag * 0.15915494309189533577 < 200.0;
  // the FP64 kernel evaluations take sin/cos of k r for |k r| <= SINCOS_RED_MAX
  if (in.helm && std::fabs(in.zk[0]) * diag > SINCOS_RED_MAX) {
    fprintf(stderr, "fmm3dbie_amd: |Re k| x surface diameter = %g exceeds %g\n", std::fabs(in.zk[0]) * diag,
            SINCOS_RED_MAX);
    return -6;
  }

  P->ixyzs = own<int64_t>(P, npat + 1);
  P->ixyzso = own<int64_t>(P, npat + 1);
  P->row_ptr = own<int64_t>(P, N + 1);
  P->col_ind = own<int64_t>(P, in.nnz);
  P->iquad = own<int64_t>(P, in.nnz + 1);
  P->ioff = own<int64_t>(P, npat);
  P->np_of = own<int>(P, npat);
  P->opatch = own<int>(P, NO);
  P->near_sorted = own<int>(P, in.nnz);
  P->csr_src = own<int>(P, in.nnz);
  std::vector<int> csr_src(in.nnz);
  for (int64_t j = 0; j < in.nnz; j++) csr_src[j] = (int)(in.ixyzs[in.col_ind[j] - 1] - 1);
  P->imat = own<double>(P, imat_all.size());
  P->wover = own<double>(P, NO);
  P->wover32 = own<float>(P, NO);
  for (int c = 0; c < 6; c++) {
    P->src64[c] = own<double>(P, NO);
    P->src32[c] = own<float>(P, NO);
  }
  for (int c = 0; c < 3; c++) {
    P->trg64[c] = own<double>(P, N);
    P->trg32[c] = own<float>(P, N);
  }
  const int64_t nq = in.nquad;
  P->nquad = nq;
  P->wnear64 = own<double>(P, 2 * nq);
  P->wnear32 = own<float>(P, 2 * nq);
  double* d_srcover = dal<double>(12 * NO);
  double* d_srcvals = dal<double>(tstride * N);
  if (!d_srcover || !d_srcvals || !P->wnear32 || !P->wnear64) return -2;
  ACHK(hipMemcpyAsync(P->ixyzs, in.ixyzs, 8 * (npat + 1), hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->ixyzso, in.ixyzso, 8 * (npat + 1), hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->row_ptr, in.row_ptr, 8 * (N + 1), hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->col_ind, in.col_ind, 8 * in.nnz, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->iquad, in.iquad, 8 * (in.nnz + 1), hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->ioff, ioff.data(), 8 * npat, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->np_of, np_of.data(), 4 * npat, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->opatch, opatch.data(), 4 * NO, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->near_sorted, nsorted.data(), 4 * in.nnz, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->csr_src, csr_src.data(), 4 * in.nnz, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->imat, imat_all.data(), 8 * imat_all.size(), hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(P->wover, in.whtsover, 8 * NO, hipMemcpyHostToDevice, s));
  ACHK(hipMemcpyAsync(d_srcover, in.srcover, 8 * 12 * NO, hipMemcpyHostToDevice, s));
  if (N > 0) ACHK(hipMemcpyAsync(d_srcvals, tsrc, 8 * tstride * N, hipMemcpyHostToDevice, s));
  if (nq == 0) {
    // no near pairs: nothing to copy
  } else if (in.wnear_is_device) {
    if (in.helm) {
      ACHK(hipMemcpyAsync(P->wnear64, in.wnear, 16 * nq, hipMemcpyDeviceToDevice, s));
    } else {
      k_real_to_cplx<<<nblk(nq, 256), 256, 0, s>>>(nq, in.wnear, (V2<double>*)P->wnear64);
    }
  } else if (in.helm) {
    ACHK(hipMemcpyAsync(P->wnear64, in.wnear, 16 * nq, hipMemcpyHostToDevice, s));
  } else {
    double* tmp = dal<double>(nq);
    if (!tmp) return -2;
    ACHK(hipMemcpyAsync(tmp, in.wnear, 8 * nq, hipMemcpyHostToDevice, s));
    k_real_to_cplx<<<nblk(nq, 256), 256, 0, s>>>(nq, tmp, (V2<double>*)P->wnear64);
    ACHK(hipStreamSynchronize(s));
    (void)hipFree(tmp);
  }
  if (nq > 0) k_convert<double, float><<<nblk(2 * nq, 256), 256, 0, s>>>(2 * nq, P->wnear64, P->wnear32);
  if (NO > 0)
    k_split_src<<<nblk(NO, 256), 256, 0, s>>>(NO, d_srcover, P->src64[0], P->src64[1], P->src64[2], P->src64[3],
                                              P->src64[4], P->src64[5], 12);
  if (N > 0)
    k_split_src<<<nblk(N, 256), 256, 0, s>>>(N, d_srcvals, P->trg64[0], P->trg64[1], P->trg64[2], nullptr, nullptr,
                                            nullptr, tstride);
  if (NO > 0)
    for (int c = 0; c < 6; c++) k_convert<double, float><<<nblk(NO, 256), 256, 0, s>>>(NO, P->src64[c], P->src32[c]);
  if (N > 0)
    for (int c = 0; c < 3; c++) k_convert<double, float><<<nblk(N, 256), 256, 0, s>>>(N, P->trg64[c], P->trg32[c]);
  if (NO > 0) k_convert<double, float><<<nblk(NO, 256), 256, 0, s>>>(NO, P->wover, P->wover32);
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(s));
  (void)hipFree(d_srcover);
  (void)hipFree(d_srcvals);
  P->ws_cap = 0;
  return 0;
}

static int ensure_ws(ApplyPlan* P, int nrhs) {
  if (P->ws_cap >= nrhs) return 0;
  if (P->q64) { (void)hipFree(P->q64); (void)hipFree(P->m64); (void)hipFree(P->q32); (void)hipFree(P->m32); }
  if (P->part64) { (void)hipFree(P->part64); (void)hipFree(P->part32); }
  if (P->sig64c) (void)hipFree(P->sig64c);
  if (P->pot64c) (void)hipFree(P->pot64c);
  const size_t NO = P->nptso;
  P->q64 = dal<double>(2 * NO * nrhs);
  P->m64 = dal<double>(2 * NO * nrhs);
  P->q32 = dal<float>(2 * NO * nrhs);
  P->m32 = dal<float>(2 * NO * nrhs);
  P->sig64c = dal<double>(2 * (size_t)P->npts);
  P->pot64c = dal<double>(2 * (size_t)P->ntarg);
  // partial far sums: nsplit x nrhs x npts with nsplit x nrhs <= far_split_cap x 64
  const size_t npart = (size_t)P->far_split_cap * 64 * P->ntarg;
  P->part64 = dal<double>(2 * npart);
  P->part32 = dal<float>(2 * npart);
  if (!P->q64 || !P->m64 || !P->q32 || !P->m32 || !P->sig64c || !P->pot64c || !P->part64 || !P->part32) return -2;
  P->ws_cap = nrhs;
  return 0;
}

void apply_plan_free_ws(ApplyPlan* P) {
  void* ps[] = {P->q64, P->m64, P->q32, P->m32, P->sig64c, P->pot64c, P->part64, P->part32};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  P->q64 = P->m64 = P->sig64c = P->pot64c = P->part64 = nullptr;
  P->q32 = P->m32 = P->part32 = nullptr;
  P->ws_cap = 0;
}

// Far sum of nrhs right-hand sides in groups of NRHS (grid z), then the reduction over
// source splits fused with the CSR correction (uniform p) or followed by the general CSR
// kernel.  The split count keeps the grid at >= wg_target workgroups, within the partial
// buffer (far_split_cap x 64 x npts).
template <typename T, int NRHS>
static int launch_far(ApplyPlan* P, bool mask, int64_t wg_target, const V2<T>* q, const V2<T>* m, int64_t ldstr,
                      int nrhs, V2<T>* part, int64_t part_cap, int64_t* nsplit_out, hipStream_t s) {
  const T* const* src;
  const T* const* trg;
  if constexpr (sizeof(T) == 8) {
    src = (const T* const*)P->src64;
    trg = (const T* const*)P->trg64;
  } else {
    src = (const T* const*)P->src32;
    trg = (const T* const*)P->trg32;
  }
  const T kr = (T)(P->helm ? P->zk[0] : 0.0), ki = (T)(P->helm ? P->zk[1] : 0.0);
  const T krev = (T)(P->helm ? P->zk[0] * 0.15915494309189533577 : 0.0);  // k / (2 pi)
  const T th2 = (T)(P->thresh * P->thresh);
  const unsigned g = nblk(P->ntarg, TPB);
  const int ngrp = nrhs / NRHS;
  const int64_t ntiles = (P->nptso + TPB - 1) / TPB;
  int64_t nsplit = (wg_target + (int64_t)g * ngrp - 1) / ((int64_t)g * ngrp);
  // splits that fit the partial buffer (part_cap complex values)
  const int64_t cap = part_cap / ((int64_t)nrhs * (P->ntarg > 0 ? P->ntarg : 1));
  if (nsplit > cap) nsplit = cap;
  if (nsplit > ntiles) nsplit = ntiles;
  if (nsplit < 1) nsplit = 1;
  const int64_t chunk = ((ntiles + nsplit - 1) / nsplit) * TPB;
  nsplit = (P->nptso + chunk - 1) / chunk;
  const int64_t ldpart = (int64_t)nrhs * P->ntarg;
  dim3 grid(g, (unsigned)nsplit, (unsigned)ngrp);
#define FMM_FAR(MK, CKV)                                                                                       \
  k_farsum<T, MK, CKV, NRHS><<<grid, TPB, 0, s>>>(P->ntarg, trg[0], trg[1], trg[2], P->nptso, src[0], src[1], src[2], \
                                                  src[3], src[4], src[5], P->ixyzso, q, m, ldstr, P->row_ptr,       \
                                                  P->near_sorted, kr, krev, ki, th2, chunk, part, ldpart)
  const bool ck = ki != (T)0 || (sizeof(T) == 4 && !P->phase_small);
  if (mask) {
    if (ck) FMM_FAR(true, true);
    else FMM_FAR(true, false);
  } else {
    if (ck) FMM_FAR(false, true);
    else FMM_FAR(false, false);
  }
#undef FMM_FAR
  *nsplit_out = nsplit;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The matrix-core far sum: 16 right-hand sides of one group per launch column, FP32,
// near patches masked; 64 targets per workgroup, the source range split as for k_farsum.
static int launch_far_mfma(ApplyPlan* P, int64_t wg_target, const V2<float>* q, const V2<float>* m,
                           int64_t ldstr, int nrhs, V2<float>* part, int64_t part_cap, int64_t* nsplit_out,
                           hipStream_t s) {
  const float* const* src = (const float* const*)P->src32;
  const float* const* trg = (const float* const*)P->trg32;
  const float kr = (float)(P->helm ? P->zk[0] : 0.0), ki = (float)(P->helm ? P->zk[1] : 0.0);
  const float krev = (float)(P->helm ? P->zk[0] * 0.15915494309189533577 : 0.0);  // k / (2 pi)
  const unsigned g = nblk(P->ntarg, 16 * FW);
  const int ngrp = nrhs / 16;
  const int64_t ntiles = (P->nptso + FT - 1) / FT;
  int64_t nsplit = (wg_target + (int64_t)g * ngrp - 1) / ((int64_t)g * ngrp);
  const int64_t cap = part_cap / ((int64_t)nrhs * (P->ntarg > 0 ? P->ntarg : 1));
  if (nsplit > cap) nsplit = cap;
  if (nsplit > ntiles) nsplit = ntiles;
  if (nsplit < 1) nsplit = 1;
  const int64_t chunk = ((ntiles + nsplit - 1) / nsplit) * FT;
  nsplit = (P->nptso + chunk - 1) / chunk;
  const int64_t ldpart = (int64_t)nrhs * P->ntarg;
  dim3 grid(g, (unsigned)nsplit, (unsigned)ngrp);
#define FMM_FARM(CKV)                                                                                          \
  k_farsum_mfma<CKV><<<grid, 64 * FW, 0, s>>>(P->ntarg, trg[0], trg[1], trg[2], P->nptso, src[0], src[1], src[2], \
                                             src[3], src[4], src[5], P->ixyzso, q, m, ldstr, P->row_ptr,         \
                                             P->near_sorted, kr, krev, ki, chunk, part, ldpart)
  if (ki != 0.0f || !P->phase_small) FMM_FARM(true);
  else FMM_FARM(false);
#undef FMM_FARM
  *nsplit_out = nsplit;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
static int far_group(const ApplyPlan* P, int nrhs, int prec, int64_t* wg) {
  const auto it = P->tuned.find(prec * 1000 + nrhs);
  if (it != P->tuned.end()) {
    *wg = it->second.second;
    return it->second.first;
  }
  *wg = P->far_wg_target;
  if (sizeof(T) == 4 && nrhs % 16 == 0) return 16;  // the matrix-core far sum
  const int pref = sizeof(T) == 8 ? 4 : 8;
  for (int gsz = pref; gsz > 1; gsz >>= 1)
    if (nrhs % gsz == 0) return gsz;
  return 1;
}

template <typename T>
static int dispatch_far(ApplyPlan* P, bool mask, int grp, int64_t wg, const V2<T>* q, const V2<T>* m, int64_t ldstr,
                        int nrhs, V2<T>* part, int64_t part_cap, int64_t* nsplit, hipStream_t s) {
  switch (grp) {
    case 1: return launch_far<T, 1>(P, mask, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
    case 2: return launch_far<T, 2>(P, mask, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
    case 4: return launch_far<T, 4>(P, mask, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
    case 8: return launch_far<T, 8>(P, mask, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
    case 16:
      if constexpr (sizeof(T) == 4) {
        if (mask) return launch_far_mfma(P, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
        return launch_far<T, 16>(P, mask, wg, q, m, ldstr, nrhs, part, part_cap, nsplit, s);
      }
      return -6;
    default: return -6;
  }
}

// pot = far (+ CSR) for all right-hand sides
template <typename T>
static int far_and_csr(ApplyPlan* P, int prec, bool mask, const V2<T>* q, const V2<T>* m, const V2<T>* sg,
                       V2<T>* pt, int nrhs, V2<T>* part, int64_t part_cap, hipStream_t s, hipEvent_t ev_far) {
  const int64_t N = P->ntarg, NO = P->nptso, NS = P->npts;
  const V2<T>* wn = (const V2<T>*)(sizeof(T) == 8 ? (void*)P->wnear64 : (void*)P->wnear32);
  int64_t nsplit = 0;
  if (N == 0) return 0;
  if (NO > 0) {
    int64_t wg;
    const int grp = far_group<T>(P, nrhs, prec, &wg);
    const int rc = dispatch_far<T>(P, mask, grp, wg, q, m, NO, nrhs, part, part_cap, &nsplit, s);
    if (rc) return rc;
  }
  if (ev_far) ACHK(hipEventRecord(ev_far, s));
  if (nsplit > 4) {
    // many source splits: one coalesced pass over the partials before the per-target CSR
    k_split_sum<T><<<dim3((unsigned)((N + 63) / 64), (unsigned)nrhs), 256, 0, s>>>(N, (int)nsplit, part,
                                                                                  (int64_t)nrhs * N);
    nsplit = 1;
  }
  if (P->uniform_p > 0) {
#define FMM_CSR(PC)                                                                                             \
  if (nrhs == 1)                                                                                               \
    k_reduce_csr<T, T, 1, PC><<<nblk(N, 4), 256, 0, s>>>(N, nrhs, (int)nsplit, part, (int64_t)nrhs * N,         \
                                                         P->row_ptr, P->csr_src, P->iquad, P->uniform_p, wn, sg, \
                                                         NS, pt, N);                                             \
  else                                                                                                         \
    k_reduce_csr<T, T, 8, PC><<<nblk(N, 4), 256, 0, s>>>(N, nrhs, (int)nsplit, part, (int64_t)nrhs * N,         \
                                                         P->row_ptr, P->csr_src, P->iquad, P->uniform_p, wn, sg, \
                                                         NS, pt, N)
    switch (P->uniform_p) {
      case 3: FMM_CSR(3); break;
      case 6: FMM_CSR(6); break;
      case 10: FMM_CSR(10); break;
      case 15: FMM_CSR(15); break;
      case 21: FMM_CSR(21); break;
      case 28: FMM_CSR(28); break;
      case 36: FMM_CSR(36); break;
      case 45: FMM_CSR(45); break;
      case 55: FMM_CSR(55); break;
      case 66: FMM_CSR(66); break;
      default: FMM_CSR(0);
    }
#undef FMM_CSR
  } else {
    if (nsplit > 0)
      k_farsum_reduce<T><<<dim3(nblk(N, 256), nrhs), 256, 0, s>>>(N, nrhs, (int)nsplit, part, (int64_t)nrhs * N, pt,
                                                                  N);
    else
      k_fill<T><<<nblk(2 * N * nrhs, 256), 256, 0, s>>>(2 * N * nrhs, (T*)pt, (T)0);
    k_csr_gen<T, T><<<nblk(N, 256), 256, 0, s>>>(N, P->row_ptr, P->col_ind, P->iquad, P->ixyzs, wn, sg, NS, nrhs,
                                                 pt, N);
  }
  ACHK(hipGetLastError());
  return 0;
}

// Work buffers of one apply in the precision of the call: charge and dipole strengths
// at the oversampled nodes (nptso x nrhs complex each) and the far-sum partials
// (part_cap complex values, at least ntarg x nrhs: one source split).
struct ApplyWs {
  void *q, *m, *part;
  int64_t part_cap;
};

// sigma/pot: complex (interleaved) device arrays of length npts * nrhs, leading dim npts
static int apply_run_ws(ApplyPlan* P, int prec, int mode, const void* sigma, void* pot, int nrhs, hipStream_t s,
                        ApplyTimes* tm, const ApplyWs& W) {
  const int64_t N = P->npts, NT = P->ntarg, NO = P->nptso;
  const double ar = P->alpha[0], ai = P->helm ? P->alpha[1] : 0.0;
  const double br = P->beta[0], bi = P->helm ? P->beta[1] : 0.0;
  hipEvent_t ev[5];
  if (tm)
    for (auto& e : ev) ACHK(hipEventCreate(&e));
  if (tm) ACHK(hipEventRecord(ev[0], s));
  int rc = 0;
  if (prec == 64) {
    const V2<double>* sg = (const V2<double>*)sigma;
    V2<double>* pt = (V2<double>*)pot;
    if (NO > 0)
      k_oversample<double><<<nblk(NO, 256), 256, 0, s>>>(NO, P->opatch, P->ixyzs, P->ixyzso, P->ioff, P->np_of, P->imat,
                                                       sg, nrhs, N, P->wover, ar, ai, br, bi,
                                                       (V2<double>*)W.q, (V2<double>*)W.m, NO);
    if (tm) ACHK(hipEventRecord(ev[1], s));
    const bool mask = mode == APPLY_MASKED;
    auto* q = (const V2<double>*)W.q;
    auto* m = (const V2<double>*)W.m;
    rc = far_and_csr<double>(P, 64, mask, q, m, sg, pt, nrhs, (V2<double>*)W.part, W.part_cap, s,
                             tm ? ev[2] : nullptr);
    if (tm) ACHK(hipEventRecord(ev[3], s));
    if (!rc && !mask) {
      const double kr = P->helm ? P->zk[0] : 0.0, ki = P->helm ? P->zk[1] : 0.0;
      if (NT > 0)
        k_near_sub<<<nblk(NT, 4), 256, 0, s>>>(NT, P->trg64[0], P->trg64[1], P->trg64[2], P->src64[0], P->src64[1],
                                               P->src64[2], P->src64[3], P->src64[4], P->src64[5], q, m, NO, nrhs,
                                               P->row_ptr, P->col_ind, P->ixyzso, kr, ki, P->thresh * P->thresh, pt,
                                               NT);
    }
    if (tm) ACHK(hipEventRecord(ev[4], s));
  } else {
    const V2<float>* sg = (const V2<float>*)sigma;
    V2<float>* pt = (V2<float>*)pot;
    if (NO > 0)
      k_oversample<float><<<nblk(NO, 256), 256, 0, s>>>(NO, P->opatch, P->ixyzs, P->ixyzso, P->ioff, P->np_of, P->imat,
                                                      sg, nrhs, N, P->wover32, ar, ai, br, bi, (V2<float>*)W.q,
                                                      (V2<float>*)W.m, NO);
    if (tm) ACHK(hipEventRecord(ev[1], s));
    rc = far_and_csr<float>(P, 32, true, (const V2<float>*)W.q, (const V2<float>*)W.m, sg, pt, nrhs,
                            (V2<float>*)W.part, W.part_cap, s, tm ? ev[2] : nullptr);
    if (tm) ACHK(hipEventRecord(ev[3], s));
    if (tm) ACHK(hipEventRecord(ev[4], s));
  }
  if (rc) return rc;
  ACHK(hipGetLastError());
  if (tm) {
    ACHK(hipEventSynchronize(ev[4]));
    float ms;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]); tm->ms_over = ms;
    (void)hipEventElapsedTime(&ms, ev[1], ev[2]); tm->ms_far = ms;
    (void)hipEventElapsedTime(&ms, ev[2], ev[3]); tm->ms_csr = ms;
    (void)hipEventElapsedTime(&ms, ev[3], ev[4]); tm->ms_sub = ms;
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  return 0;
}

static int check_prec_nrhs(int prec, int nrhs) {
  if (nrhs < 1 || nrhs > 64) return -6;
  if (prec != 32 && prec != 64) {
    fprintf(stderr, "fmm3dbie_amd: apply precision %d (must be 32 or 64)\n", prec);
    return -6;
  }
  return 0;
}

static ApplyWs plan_ws(const ApplyPlan* P, int prec) {
  const int64_t cap = (int64_t)P->far_split_cap * 64 * P->ntarg;
  if (prec == 64) return {P->q64, P->m64, P->part64, cap};
  return {P->q32, P->m32, P->part32, cap};
}

// sigma/pot: complex (interleaved) device arrays of length npts * nrhs, leading dim npts;
// work buffers owned by the plan (allocated on the first call that needs them)
int apply_plan_run(ApplyPlan* P, int prec, int mode, const void* sigma, void* pot, int nrhs, hipStream_t s,
                   ApplyTimes* tm) {
  const int rc = check_prec_nrhs(prec, nrhs);
  if (rc) return rc;
  if (ensure_ws(P, nrhs) != 0) return -2;
  return apply_run_ws(P, prec, mode, sigma, pot, nrhs, s, tm, plan_ws(P, prec));
}

// ------------------------------------------------------------------ allocation-free apply
// The work buffers of apply_plan_run_noalloc live in this file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes `scratch` itself to
// a kernel (on the host the name is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns, looked up once per device.
constexpr size_t APPLY_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[APPLY_SCRATCH_BYTES];

static std::mutex g_scratch_mutex;
static std::map<int, void*> g_scratch_ptr;  // device -> address of its copy of `scratch`

static void* scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    fprintf(stderr, "fmm3dbie_amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
    return nullptr;
  }
  g_scratch_ptr[dev] = ptr;
  return ptr;
}

size_t apply_scratch_bytes() { return APPLY_SCRATCH_BYTES; }

int apply_plan_run_noalloc(ApplyPlan* P, int prec, int mode, const void* sigma, void* pot, int nrhs, hipStream_t s) {
  int rc = check_prec_nrhs(prec, nrhs);
  if (rc) return rc;
  const size_t cs = prec == 64 ? 2 * sizeof(double) : 2 * sizeof(float);  // one complex value
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t strength = up(cs * (size_t)(P->nptso > 0 ? P->nptso : 1) * nrhs);
  const size_t one_split = cs * (size_t)(P->ntarg > 0 ? P->ntarg : 1) * nrhs;
  if (2 * strength + one_split > APPLY_SCRATCH_BYTES) {
    fprintf(stderr,
            "fmm3dbie_amd: allocation-free apply: %zu bytes of work buffers needed (nptso %lld, targets %lld, "
            "nrhs %d, prec %d), the fixed scratch holds %zu\n",
            2 * strength + one_split, (long long)P->nptso, (long long)P->ntarg, nrhs, prec, APPLY_SCRATCH_BYTES);
    return -7;
  }
  unsigned char* base = (unsigned char*)scratch_address();
  if (!base) return -1;
  ApplyWs W;
  W.q = base;
  W.m = base + strength;
  W.part = base + 2 * strength;
  W.part_cap = (int64_t)((APPLY_SCRATCH_BYTES - 2 * strength) / cs);
  return apply_run_ws(P, prec, mode, sigma, pot, nrhs, s, nullptr, W);
}

// Autotuner of the far-sum launch shape for (prec, nrhs): every right-hand-side group
// size that divides nrhs (1..8 in FP64, 1..16 in FP32) and workgroup targets of 1024,
// 4096, 8192, 16384, 32768 are timed (one warm-up launch, then one launch each) on unit
// densities; the fastest is kept
// in the plan and used by every later apply with the same (prec, nrhs).
int apply_plan_tune(ApplyPlan* P, int prec, int nrhs, hipStream_t s, double* report, int nreport) {
  if (nrhs < 1 || nrhs > 64 || (prec != 32 && prec != 64)) return -6;
  if (ensure_ws(P, nrhs) != 0) return -2;
  if (P->ntarg == 0 || P->nptso == 0) return 0;
  const int64_t NO = P->nptso;
  if (prec == 64) {
    k_fill<double><<<nblk(2 * NO * nrhs, 256), 256, 0, s>>>(2 * NO * nrhs, P->q64, 1.0);
    k_fill<double><<<nblk(2 * NO * nrhs, 256), 256, 0, s>>>(2 * NO * nrhs, P->m64, 1.0);
  } else {
    k_fill<float><<<nblk(2 * NO * nrhs, 256), 256, 0, s>>>(2 * NO * nrhs, P->q32, 1.0f);
    k_fill<float><<<nblk(2 * NO * nrhs, 256), 256, 0, s>>>(2 * NO * nrhs, P->m32, 1.0f);
  }
  auto far = [&](int gsz, int64_t wg) {
    int64_t nsplit;
    const int64_t cap = (int64_t)P->far_split_cap * 64 * P->ntarg;
    return prec == 64 ? dispatch_far<double>(P, true, gsz, wg, (const V2<double>*)P->q64, (const V2<double>*)P->m64,
                                             NO, nrhs, (V2<double>*)P->part64, cap, &nsplit, s)
                      : dispatch_far<float>(P, true, gsz, wg, (const V2<float>*)P->q32, (const V2<float>*)P->m32, NO,
                                            nrhs, (V2<float>*)P->part32, cap, &nsplit, s);
  };
  hipEvent_t e0, e1;
  ACHK(hipEventCreate(&e0));
  ACHK(hipEventCreate(&e1));
  const int groups[5] = {1, 2, 4, 8, 16};
  const int64_t wgs[4] = {4096, 8192, 16384, 32768};
  float best = 1e30f;
  int bg = 1, nrep = 0;
  int64_t bw = P->far_wg_target;
  // one warm-up launch, then one timed launch per shape
  int rc = far(1, P->far_wg_target);
  if (rc) return rc;
  for (int gsz : groups) {
    if (nrhs % gsz != 0 || (prec == 64 && gsz > 8)) continue;
    for (int64_t wg : wgs) {
      ACHK(hipEventRecord(e0, s));
      rc = far(gsz, wg);
      if (rc) return rc;
      ACHK(hipEventRecord(e1, s));
      ACHK(hipEventSynchronize(e1));
      float ms;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (report && nrep + 3 <= nreport) {
        report[nrep++] = gsz;
        report[nrep++] = (double)wg;
        report[nrep++] = ms;
      }
      if (ms < best) {
        best = ms;
        bg = gsz;
        bw = wg;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P->tuned[prec * 1000 + nrhs] = {bg, bw};
  return nrep / 3;
}

// real (Laplace) densities: convert to complex, apply, take real part
int apply_plan_run_real(ApplyPlan* P, int mode, const double* sigma, double* pot, hipStream_t s) {
  if (ensure_ws(P, 1) != 0) return -2;
  const int64_t N = P->npts, NT = P->ntarg;
  k_real_to_cplx<<<nblk(N, 256), 256, 0, s>>>(N, sigma, (V2<double>*)P->sig64c);
  int r = apply_plan_run(P, 64, mode, P->sig64c, P->pot64c, 1, s, nullptr);
  if (r) return r;
  if (NT > 0) k_cplx_to_real<<<nblk(NT, 256), 256, 0, s>>>(NT, (const V2<double>*)P->pot64c, pot);
  ACHK(hipGetLastError());
  return 0;
}
