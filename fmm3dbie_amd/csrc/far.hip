// Far-field oversampling order on gfx950 (get_far_order / get_far_order_guru,
// far_field_routs.f90:10-539).  Per source patch: 25 probe targets -- the min(10, n/2)
// farthest of its n near targets (ties to the higher target index, the order of the
// reference's stable ascending sort read from the end) and points on the sphere of
// radius rfac*rad around the centroid at the hkrand angles; then, for the RV rules of
// orders 1, 2, 3, 4, 6, 8, 10, 12, 14, 16, the moments int K(t, y) P_l(y) dA against the
// patch basis (single layer if ikerorder = -1, else double layer, k = zk), and the first
// order whose moments agree with the previous one to epsp = eps * min_i |J_i|^(1/2)
// (max over probes of the l2 difference) selects the previous order; no agreement: 16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <cstdio>
#include <vector>

#include "common.hpp"
#include "far.hpp"
#include "host_basis.hpp"

namespace {

constexpr int NPROBE = 25, NFAR = 10, CH = 32;  // probes, candidate orders, node chunk
constexpr int kNfars[NFAR] = {1, 2, 3, 4, 6, 8, 10, 12, 14, 16};

// CSC of the near list: targets of each patch in ascending target order
__global__ void k_csc_count(int64_t nnz, const int64_t* __restrict__ col_ind, int* __restrict__ cnt) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < nnz) atomicAdd(cnt + (col_ind[j] - 1), 1);
}

__global__ void k_scan1(int64_t n, const int* __restrict__ cnt, int64_t* __restrict__ off) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x;
  const int64_t chunk = (n + blockDim.x - 1) / blockDim.x;
  const int64_t a = tid * chunk, b = a + chunk < n ? a + chunk : n;
  int64_t s = 0;
  for (int64_t i = a; i < b; i++) s += cnt[i];
  part[tid] = s;
  __syncthreads();
  if (tid == 0) {
    int64_t acc = 0;
    for (int i = 0; i < (int)blockDim.x; i++) {
      const int64_t v = part[i];
      part[i] = acc;
      acc += v;
    }
    off[n] = acc;
  }
  __syncthreads();
  int64_t acc = part[tid];
  for (int64_t i = a; i < b; i++) {
    off[i] = acc;
    acc += cnt[i];
  }
}

__global__ void k_csc_fill(int64_t ntarg, const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col_ind,
                           const int64_t* __restrict__ off, int* __restrict__ fill, int* __restrict__ rows) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ntarg) return;
  for (int64_t j = row_ptr[t] - 1; j < row_ptr[t + 1] - 1; j++) {
    const int64_t p = col_ind[j] - 1;
    rows[off[p] + atomicAdd(fill + p, 1)] = (int)t;
  }
}

struct FarTables {
  // per candidate order n: RV nodes (u, v, w) at ufar[off[n]..], basis of the patch order
  // at those nodes pfar[(off[n] + q) * p + l]; RV nodes of the patch order for epsp
  const double* ufar;
  const double* pfar;
  const double* prv;  // (p, p): basis at the patch's own RV nodes
  int off[NFAR + 1];
  double ang[NPROBE][2];  // (phi, theta) of the sphere probes
};

template <int NP>
__device__ __forceinline__ double far_geom(const double* __restrict__ sc, const double* __restrict__ P, double* g) {
  for (int m = 0; m < 9; m++) g[m] = 0.0;
  for (int l = 0; l < NP; l++)
    for (int m = 0; m < 9; m++) g[m] = fma(sc[9 * l + m], P[l], g[m]);
  return geom_finish(g);
}

// one workgroup (256 threads) per source patch
template <int N>
__global__ void __launch_bounds__(256) k_far_order(int64_t npat, const int64_t* __restrict__ ixyzs,
                                                   const double* __restrict__ srccoefs,
                                                   const double* __restrict__ cms, const double* __restrict__ rads,
                                                   double rfac, int64_t ndtarg, const double* __restrict__ targs,
                                                   const int64_t* __restrict__ coff, const int* __restrict__ rows,
                                                   FarTables T, double eps, int slp, double zkr, double zki,
                                                   int64_t* __restrict__ novers) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  __shared__ double probe[NPROBE][3];
  __shared__ double sgeo[CH][7];                  // node x, y, z, n, |J| w
  __shared__ double fre[NPROBE][CH], fim[NPROBE][CH];
  __shared__ double vc[2][NPROBE * NP], vi[2][NPROBE * NP];  // moments of the previous / current order
  __shared__ double red[256];
  __shared__ double s_epsp;
  __shared__ int s_done;
  const int tid = threadIdx.x;
  const int64_t ip = blockIdx.x;
  if (ip >= npat) return;  // (the grid is one workgroup per patch)
  const double* sc = srccoefs + 9 * (ixyzs[ip] - 1);
  const double* cm = cms + 3 * ip;
  // probes: the farthest near targets (ties: higher index first), then sphere points
  const int64_t a = coff[ip], n0 = coff[ip + 1] - coff[ip];
  const int i0 = (int)(n0 / 2 < 10 ? n0 / 2 : 10);
  if (tid == 0) {
    int prev_t = -1;
    double prev_d = 0.0;
    for (int i = 0; i < i0; i++) {
      // next in (distance, index) descending order after (prev_d, prev_t)
      int best = -1;
      double bd = -1.0;
      for (int64_t k = 0; k < n0; k++) {
        const int t = rows[a + k];
        const double* x = targs + ndtarg * (int64_t)t;
        const double d = (x[0] - cm[0]) * (x[0] - cm[0]) + (x[1] - cm[1]) * (x[1] - cm[1]) +
                         (x[2] - cm[2]) * (x[2] - cm[2]);
        const bool below_prev = prev_t < 0 || d < prev_d || (d == prev_d && t < prev_t);
        if (below_prev && (best < 0 || d > bd || (d == bd && t > best))) {
          best = t;
          bd = d;
        }
      }
      const double* x = targs + ndtarg * (int64_t)best;
      probe[i][0] = x[0];
      probe[i][1] = x[1];
      probe[i][2] = x[2];
      prev_t = best;
      prev_d = bd;
    }
    const double rad = rfac * rads[ip];
    for (int i = i0; i < NPROBE; i++) {
      const double phi = T.ang[i][0], th = T.ang[i][1];
      probe[i][0] = cm[0] + rad * sin(th) * cos(phi);
      probe[i][1] = cm[1] + rad * sin(th) * sin(phi);
      probe[i][2] = cm[2] + rad * cos(th);
    }
    // epsp = eps * (min over the patch's RV nodes of |J|^2)^(1/4)
    double rsc = 0.0;
    for (int i = 0; i < NP; i++) {
      double g[12];
      const double jj = far_geom<NP>(sc, T.prv + (size_t)i * NP, g);
      if (i == 0 || jj * jj < rsc) rsc = jj * jj;
    }
    s_epsp = pow(rsc, 0.25) * eps;
    s_done = 0;
  }
  __syncthreads();
  int chosen = NFAR;
  for (int nn = 0; nn < NFAR; nn++) {
    const int cur = nn & 1;
    const int npf = T.off[nn + 1] - T.off[nn];
    for (int o = tid; o < NPROBE * NP; o += 256) {
      vc[cur][o] = 0.0;
      vi[cur][o] = 0.0;
    }
    for (int q0 = 0; q0 < npf; q0 += CH) {
      const int nq = npf - q0 < CH ? npf - q0 : CH;
      __syncthreads();
      if (tid < nq) {
        const int q = T.off[nn] + q0 + tid;
        double g[12];
        const double jac = far_geom<NP>(sc, T.pfar + (size_t)q * NP, g);
        sgeo[tid][0] = g[0];
        sgeo[tid][1] = g[1];
        sgeo[tid][2] = g[2];
        sgeo[tid][3] = g[9];
        sgeo[tid][4] = g[10];
        sgeo[tid][5] = g[11];
        sgeo[tid][6] = jac * T.ufar[3 * q + 2];
      }
      __syncthreads();
      for (int o = tid; o < NPROBE * nq; o += 256) {
        const int t = o / nq, j = o - t * nq;
        const double dx = probe[t][0] - sgeo[j][0], dy = probe[t][1] - sgeo[j][1], dz = probe[t][2] - sgeo[j][2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double s, c;
        sincos(zkr * r, &s, &c);
        const double dec = zki != 0.0 ? exp(-zki * r) : 1.0;
        double vr, vv;
        if (slp) {
          // e^{ikr} / (4 pi r)
          vr = c * dec / r * OVER4PI;
          vv = s * dec / r * OVER4PI;
        } else {
          // (x - y).n (1 - ikr) e^{ikr} / (4 pi r^3), k = zkr + i zki
          const double dn = dx * sgeo[j][3] + dy * sgeo[j][4] + dz * sgeo[j][5];
          const double ar = 1.0 + zki * r, ai = -zkr * r;
          const double f = dn / (r * r * r) * OVER4PI * dec;
          vr = f * (ar * c - ai * s);
          vv = f * (ar * s + ai * c);
        }
        fre[t][j] = vr * sgeo[j][6];
        fim[t][j] = vv * sgeo[j][6];
      }
      __syncthreads();
      for (int o = tid; o < NPROBE * NP; o += 256) {
        const int t = o / NP, l = o - t * NP;
        double sr = vc[cur][o], si = vi[cur][o];
        for (int j = 0; j < nq; j++) {
          const double pl = T.pfar[(size_t)(T.off[nn] + q0 + j) * NP + l];
          sr = fma(fre[t][j], pl, sr);
          si = fma(fim[t][j], pl, si);
        }
        vc[cur][o] = sr;
        vi[cur][o] = si;
      }
    }
    __syncthreads();
    if (nn > 0) {
      // max over probes of the l2 norm (over basis functions) of the difference
      double e = 0.0;
      if (tid < NPROBE) {
        double e2 = 0.0;
        for (int l = 0; l < NP; l++) {
          const int o = tid * NP + l;
          const double dr = vc[cur][o] - vc[cur ^ 1][o], di = vi[cur][o] - vi[cur ^ 1][o];
          const double ee = sqrt(dr * dr + di * di);
          e2 += ee * ee;
        }
        e = sqrt(e2);
      }
      red[tid] = e;
      __syncthreads();
      if (tid == 0) {
        double m = 0.0;
        for (int i = 0; i < NPROBE; i++) m = red[i] > m ? red[i] : m;
        s_done = m < s_epsp ? 1 : 0;
      }
      __syncthreads();
      if (s_done) {
        chosen = nn;
        break;
      }
    }
  }
  if (tid == 0) novers[ip] = kNfars[chosen - 1 < 0 ? 0 : chosen - 1];
}

#define FCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "fmm3dbie_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return -1;                                                                  \
    }                                                                             \
  } while (0)

template <int N>
int run_far(const FarOrderArgs& A, hipStream_t s) {
  constexpr int NP = (N + 1) * (N + 2) / 2;
  // tables: RV nodes of each candidate order with the basis of order N at them
  FarTables T;
  std::vector<double> uf, pf, pr((size_t)NP * NP), pols(NP);
  T.off[0] = 0;
  for (int nn = 0; nn < NFAR; nn++) {
    const int nf = kNfars[nn], npf = fmmhost::npols(nf);
    std::vector<double> uv(2 * npf), w(npf);
    fmmhost::rv_nodes(nf, uv.data(), w.data());
    for (int q = 0; q < npf; q++) {
      uf.push_back(uv[2 * q]);
      uf.push_back(uv[2 * q + 1]);
      uf.push_back(w[q]);
      fmmhost::koorn_pols(uv[2 * q], uv[2 * q + 1], N, pols.data());
      pf.insert(pf.end(), pols.begin(), pols.end());
    }
    T.off[nn + 1] = T.off[nn] + npf;
  }
  {
    std::vector<double> uv(2 * NP), w(NP);
    fmmhost::rv_nodes(N, uv.data(), w.data());
    for (int i = 0; i < NP; i++) fmmhost::koorn_pols(uv[2 * i], uv[2 * i + 1], N, pr.data() + (size_t)i * NP);
  }
  const double pi = std::atan(1.0) * 4;
  for (int i = 0; i < NPROBE; i++) {
    T.ang[i][0] = fmmhost::hkrand(1001 + i + 1) * 2 * pi;
    T.ang[i][1] = fmmhost::hkrand(3042 + i + 1) * pi;
  }
  double *d_uf = nullptr, *d_pf = nullptr, *d_pr = nullptr;
  int *d_cnt = nullptr, *d_fill = nullptr, *d_rows = nullptr;
  int64_t* d_off = nullptr;
  const int64_t npat = A.npatches;
  FCHK(hipMalloc((void**)&d_uf, sizeof(double) * uf.size()));
  FCHK(hipMalloc((void**)&d_pf, sizeof(double) * pf.size()));
  FCHK(hipMalloc((void**)&d_pr, sizeof(double) * pr.size()));
  FCHK(hipMalloc((void**)&d_cnt, sizeof(int) * npat));
  FCHK(hipMalloc((void**)&d_fill, sizeof(int) * npat));
  FCHK(hipMalloc((void**)&d_off, sizeof(int64_t) * (npat + 1)));
  FCHK(hipMalloc((void**)&d_rows, sizeof(int) * (A.nnz > 0 ? A.nnz : 1)));
  FCHK(hipMemcpyAsync(d_uf, uf.data(), sizeof(double) * uf.size(), hipMemcpyHostToDevice, s));
  FCHK(hipMemcpyAsync(d_pf, pf.data(), sizeof(double) * pf.size(), hipMemcpyHostToDevice, s));
  FCHK(hipMemcpyAsync(d_pr, pr.data(), sizeof(double) * pr.size(), hipMemcpyHostToDevice, s));
  FCHK(hipMemsetAsync(d_cnt, 0, sizeof(int) * npat, s));
  FCHK(hipMemsetAsync(d_fill, 0, sizeof(int) * npat, s));
  if (A.nnz > 0) k_csc_count<<<(unsigned)((A.nnz + 255) / 256), 256, 0, s>>>(A.nnz, A.col_ind, d_cnt);
  k_scan1<<<1, 1024, 0, s>>>(npat, d_cnt, d_off);
  if (A.ntarg > 0)
    k_csc_fill<<<(unsigned)((A.ntarg + 255) / 256), 256, 0, s>>>(A.ntarg, A.row_ptr, A.col_ind, d_off, d_fill,
                                                                 d_rows);
  T.ufar = d_uf;
  T.pfar = d_pf;
  T.prv = d_pr;
  k_far_order<N><<<(unsigned)npat, 256, 0, s>>>(npat, A.ixyzs, A.srccoefs, A.cms, A.rads, A.rfac, A.ndtarg,
                                                A.targs, d_off, d_rows, T, A.eps, A.ikerorder == -1 ? 1 : 0,
                                                A.zk[0], A.zk[1], A.novers);
  FCHK(hipGetLastError());
  FCHK(hipStreamSynchronize(s));
  void* frees[] = {d_uf, d_pf, d_pr, d_cnt, d_fill, d_off, d_rows};
  for (void* p : frees) (void)hipFree(p);
  return 0;
}

}  // namespace

int far_order_device(const FarOrderArgs& A, hipStream_t s) {
  if (A.npatches <= 0) return 0;
  switch (A.norder) {
    case 1: return run_far<1>(A, s);
    case 2: return run_far<2>(A, s);
    case 3: return run_far<3>(A, s);
    case 4: return run_far<4>(A, s);
    case 5: return run_far<5>(A, s);
    case 6: return run_far<6>(A, s);
    case 7: return run_far<7>(A, s);
    case 8: return run_far<8>(A, s);
    default:
      fprintf(stderr, "fmm3dbie_amd: get_far_order on the GPU supports patch orders 1..8 (got %d)\n", A.norder);
      return -4;
  }
}

// ---------------------------------------------------------------- oversampled geometry
// oversample_geom (surf_routs.f90:304-449) and get_qwts (surf_routs.f90:1598-1651) on the
// GPU, one workgroup per patch: the geometry at the RV nodes of the patch's oversampling
// order nfar is the patch expansion times the basis at those nodes (the patch's own nodes
// are copied when nfar equals the patch order), with unit normals; the node shared by the
// two RV rules when both orders are multiples of 3 (get_clashing_indices,
// patch_basis_routs.f90:518-572: node 1, or 76 for order 15 and 106 for order 18) takes
// the patch's own values; the smooth weight of every node is |x_u x x_v| times the RV
// weight of order nfar.
namespace {

constexpr int OV_MAXORD = 20;

// Basis tables of one patch order: for every oversampling order nf <= OV_MAXORD, the
// npols(norder) basis values at each of the npols(nf) RV nodes of order nf (off[nf] is
// the first row) and the RV weights of order nf (woff[nf]).
struct OversampleTables {
  int norder = -1;
  double* pmat = nullptr;
  double* wts = nullptr;
  int64_t off[OV_MAXORD + 2] = {}, woff[OV_MAXORD + 2] = {};
};

__device__ __forceinline__ int clash_node(int n) { return n == 15 ? 75 : (n == 18 ? 105 : 0); }

__global__ void k_oversample_geom(int64_t npatches, int norder, int np, const int64_t* __restrict__ ixyzs,
                                  const double* __restrict__ srccoefs, const double* __restrict__ srcvals,
                                  const int64_t* __restrict__ nfars, const int64_t* __restrict__ ixyzso,
                                  const double* __restrict__ pmat, const int64_t* __restrict__ off,
                                  const double* __restrict__ rvw, const int64_t* __restrict__ woff,
                                  double* __restrict__ srcover, double* __restrict__ wover) {
  const int64_t ip = blockIdx.x;
  if (ip >= npatches) return;
  // norder < 0: every patch keeps its own nodes (the smooth weights of a surface whose
  // patch orders are nfars, get_qwts)
  const int nf = (int)nfars[ip];
  const bool own = norder < 0 || nf == norder;
  const int64_t i0 = ixyzs[ip] - 1, o0 = ixyzso[ip] - 1;
  const int npo = (int)(ixyzso[ip + 1] - ixyzso[ip]);
  const double* sc = srccoefs + 9 * i0;
  const bool clash = !own && norder % 3 == 0 && nf % 3 == 0;
  const int jc = clash_node(nf), ic = clash_node(norder);
  for (int j = threadIdx.x; j < npo; j += blockDim.x) {
    double g[12];
    if (own || (clash && j == jc)) {
      const double* sv = srcvals + 12 * (i0 + (own ? j : ic));
      for (int m = 0; m < 12; m++) g[m] = sv[m];
    } else {
      const double* P = pmat + (off[nf] + j) * (int64_t)np;
      for (int m = 0; m < 9; m++) g[m] = 0.0;
      for (int l = 0; l < np; l++)
        for (int m = 0; m < 9; m++) g[m] = fma(sc[9 * l + m], P[l], g[m]);
      const double nx = g[4] * g[8] - g[5] * g[7], ny = g[5] * g[6] - g[3] * g[8], nz = g[3] * g[7] - g[4] * g[6];
      const double rr = sqrt(nx * nx + ny * ny + nz * nz);
      g[9] = nx / rr;
      g[10] = ny / rr;
      g[11] = nz / rr;
    }
    double* o = srcover + 12 * (o0 + j);
    for (int m = 0; m < 12; m++) o[m] = g[m];
    const double nx = g[4] * g[8] - g[5] * g[7], ny = g[5] * g[6] - g[3] * g[8], nz = g[3] * g[7] - g[4] * g[6];
    wover[o0 + j] = sqrt(nx * nx + ny * ny + nz * nz) * rvw[woff[nf] + j];
  }
}

std::map<std::pair<int, int>, OversampleTables> g_ov;  // (device, patch order) -> tables

int oversample_tables(int norder, const OversampleTables** out, hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  auto it = g_ov.find({dev, norder});
  if (it == g_ov.end()) {
    OversampleTables T;
    T.norder = norder;
    const int np = (norder + 1) * (norder + 2) / 2;
    std::vector<double> pm, w, uv, rw, pols(np);
    for (int nf = 0; nf <= OV_MAXORD; nf++) {
      const int npf = (nf + 1) * (nf + 2) / 2;
      T.off[nf] = (int64_t)pm.size() / np;
      T.woff[nf] = (int64_t)w.size();
      uv.assign(2 * npf, 0.0);
      rw.assign(npf, 0.0);
      fmmhost::rv_nodes(nf, uv.data(), rw.data());
      for (int j = 0; j < npf; j++) {
        fmmhost::koorn_pols(uv[2 * j], uv[2 * j + 1], norder, pols.data());
        pm.insert(pm.end(), pols.begin(), pols.end());
        w.push_back(rw[j]);
      }
    }
    T.off[OV_MAXORD + 1] = (int64_t)pm.size() / np;
    T.woff[OV_MAXORD + 1] = (int64_t)w.size();
    if (hipMalloc(&T.pmat, sizeof(double) * pm.size()) != hipSuccess) return -2;
    if (hipMalloc(&T.wts, sizeof(double) * w.size()) != hipSuccess) return -2;
    if (hipMemcpyAsync(T.pmat, pm.data(), sizeof(double) * pm.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(T.wts, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -1;
    it = g_ov.emplace(std::make_pair(dev, norder), T).first;
  }
  *out = &it->second;
  return 0;
}

}  // namespace

int oversample_geom_device(int64_t npatches, int norder, const int64_t* ixyzs, const double* srccoefs,
                           const double* srcvals, const int64_t* nfars, const int64_t* ixyzso, double* srcover,
                           double* wover, hipStream_t s) {
  if (npatches == 0) return 0;
  if (norder > OV_MAXORD) return -4;
  const OversampleTables* T = nullptr;
  const int rc = oversample_tables(norder < 0 ? 0 : norder, &T, s);
  if (rc) return rc;
  int64_t* doff = nullptr;
  if (hipMalloc(&doff, sizeof(int64_t) * 2 * (OV_MAXORD + 2)) != hipSuccess) return -2;
  int64_t hoff[2 * (OV_MAXORD + 2)];
  for (int k = 0; k < OV_MAXORD + 2; k++) {
    hoff[k] = T->off[k];
    hoff[OV_MAXORD + 2 + k] = T->woff[k];
  }
  int err = 0;
  if (hipMemcpyAsync(doff, hoff, sizeof(hoff), hipMemcpyHostToDevice, s) != hipSuccess) err = -1;
  if (!err) {
    const int np = norder < 0 ? 1 : (norder + 1) * (norder + 2) / 2;
    k_oversample_geom<<<(unsigned)npatches, 64, 0, s>>>(npatches, norder, np, ixyzs, srccoefs, srcvals, nfars, ixyzso,
                                                        T->pmat, doff, T->wts, doff + OV_MAXORD + 2, srcover, wover);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) err = -1;
  }
  (void)hipFree(doff);
  return err;
}
