// Near-field finder on gfx950 (findnearmem / findnear, near_field_routs.f:193-757, with
// the result of the brute-force twin findnearslow, :38-186): target i is near source j
// iff |x_i - c_j|^2 <= r_j^2, the squared distance summed as ((dx^2 + dy^2) + dz^2)
// without fused multiply-adds so that boundary cases are decided as in the reference.
//
// Sources are binned on a uniform grid of cell size h = max_j r_j (slightly enlarged), so
// every near source of a target lies in the 3x3x3 cells around the target's cell.  One
// thread per target counts (pass 1) and then writes (pass 2) its sources; each row is
// sorted ascending, the order findnearslow produces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "near.hpp"

namespace {

struct Grid {
  double lo[3], inv_h;
  int n[3];
};

// cell index along one axis; -1 and n stand for "outside below/above" (a neighbour
// may still be inside); NaN maps to -1
__device__ __forceinline__ int cell_coord(double x, double lo, double inv_h, int n) {
  const double t = (x - lo) * inv_h;
  if (!(t >= 0.0)) return -1;
  if (t >= (double)n) return n;
  return (int)t;
}

__global__ void k_src_cell(int64_t ns, const double* __restrict__ xyzs, Grid G, int* __restrict__ cell,
                           int* __restrict__ cnt) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ns) return;
  int c[3];
  for (int d = 0; d < 3; d++) {
    int v = cell_coord(xyzs[3 * j + d], G.lo[d], G.inv_h, G.n[d]);
    c[d] = v < 0 ? 0 : (v >= G.n[d] ? G.n[d] - 1 : v);
  }
  const int id = (c[2] * G.n[1] + c[1]) * G.n[0] + c[0];
  cell[j] = id;
  atomicAdd(cnt + id, 1);
}

// exclusive scan of n ints into off (n + 1 entries) by one workgroup of 1024 threads
__global__ void k_scan_int(int64_t n, const int* __restrict__ cnt, int64_t* __restrict__ off, int64_t base) {
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
    off[n] = acc + base;
  }
  __syncthreads();
  int64_t acc = part[tid] + base;
  for (int64_t i = a; i < b; i++) {
    off[i] = acc;
    acc += cnt[i];
  }
}

__global__ void k_src_scatter(int64_t ns, const int* __restrict__ cell, const int64_t* __restrict__ coff,
                              int* __restrict__ fill, int* __restrict__ sorted) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const int id = cell[j];
  sorted[coff[id] + atomicAdd(fill + id, 1)] = (int)j;
}

__device__ __forceinline__ bool is_near(double tx, double ty, double tz, const double* __restrict__ c, double r) {
  const double dx = __dsub_rn(c[0], tx), dy = __dsub_rn(c[1], ty), dz = __dsub_rn(c[2], tz);
  const double rr = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  return rr <= __dmul_rn(r, r);
}

// pass 1 (out == nullptr): count per target; pass 2: write the 1-based source indices of
// target i at out[row_ptr[i] - 1 ...] and sort them
__global__ void k_near_targets(int64_t nt, int64_t ndt, const double* __restrict__ targets,
                               const double* __restrict__ xyzs, const double* __restrict__ rads, Grid G,
                               const int64_t* __restrict__ coff, const int* __restrict__ sorted,
                               int* __restrict__ count, const int64_t* __restrict__ row_ptr,
                               int64_t* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const double tx = targets[ndt * i], ty = targets[ndt * i + 1], tz = targets[ndt * i + 2];
  const int cx = cell_coord(tx, G.lo[0], G.inv_h, G.n[0]), cy = cell_coord(ty, G.lo[1], G.inv_h, G.n[1]),
            cz = cell_coord(tz, G.lo[2], G.inv_h, G.n[2]);
  int m = 0;
  int64_t* row = out ? out + (row_ptr[i] - 1) : nullptr;
  for (int z = cz - 1; z <= cz + 1; z++) {
    if (z < 0 || z >= G.n[2]) continue;
    for (int y = cy - 1; y <= cy + 1; y++) {
      if (y < 0 || y >= G.n[1]) continue;
      for (int x = cx - 1; x <= cx + 1; x++) {
        if (x < 0 || x >= G.n[0]) continue;
        const int id = (z * G.n[1] + y) * G.n[0] + x;
        for (int64_t k = coff[id]; k < coff[id + 1]; k++) {
          const int j = sorted[k];
          if (is_near(tx, ty, tz, xyzs + 3 * (int64_t)j, rads[j])) {
            if (row) {
              // insertion into the ascending row
              int64_t v = j + 1;
              int p = m;
              while (p > 0 && row[p - 1] > v) {
                row[p] = row[p - 1];
                p--;
              }
              row[p] = v;
            }
            m++;
          }
        }
      }
    }
  }
  if (!out) count[i] = m;
}

struct Finder {
  Grid G;
  int* cell = nullptr;
  int* ccnt = nullptr;
  int64_t* coff = nullptr;
  int* fill = nullptr;
  int* sorted = nullptr;
  int64_t ncell = 0;
  ~Finder() {
    void* ps[] = {cell, ccnt, coff, fill, sorted};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
};

#define NCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "fmm3dbie_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return -1;                                                                  \
    }                                                                             \
  } while (0)

// grid of cell size max(r) over the source bounding box (host reads the small inputs)
int build_grid(int64_t ns, const double* xyzs, const double* rads, Finder& F, hipStream_t s) {
  std::vector<double> hx(3 * ns), hr(ns);
  NCHK(hipMemcpyAsync(hx.data(), xyzs, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, s));
  NCHK(hipMemcpyAsync(hr.data(), rads, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
  NCHK(hipStreamSynchronize(s));
  double lo[3] = {hx[0], hx[1], hx[2]}, hi[3] = {hx[0], hx[1], hx[2]}, rmax = 0.0;
  for (int64_t j = 0; j < ns; j++) {
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], hx[3 * j + d]);
      hi[d] = std::max(hi[d], hx[3 * j + d]);
    }
    if (!(hr[j] >= 0.0) || !std::isfinite(hr[j])) {
      fprintf(stderr, "fmm3dbie_amd: findnear radius %lld is not a finite non-negative number\n", (long long)j);
      return -8;
    }
    rmax = std::max(rmax, hr[j]);
  }
  double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  // cell size >= rmax (so near sources are within one cell), at most 2^21 cells in total
  double h = std::max(rmax, ext * 1e-6) * (1.0 + 1e-9);
  if (h <= 0.0) h = 1.0;
  for (;;) {
    int64_t tot = 1;
    for (int d = 0; d < 3; d++) tot *= (int64_t)((hi[d] - lo[d]) / h) + 1;
    if (tot <= (int64_t)1 << 21) break;
    h *= 1.25;
  }
  for (int d = 0; d < 3; d++) {
    F.G.lo[d] = lo[d];
    F.G.n[d] = (int)((hi[d] - lo[d]) / h) + 1;
  }
  F.G.inv_h = 1.0 / h;
  F.ncell = (int64_t)F.G.n[0] * F.G.n[1] * F.G.n[2];
  NCHK(hipMalloc((void**)&F.cell, sizeof(int) * std::max<int64_t>(ns, 1)));
  NCHK(hipMalloc((void**)&F.ccnt, sizeof(int) * F.ncell));
  NCHK(hipMalloc((void**)&F.coff, sizeof(int64_t) * (F.ncell + 1)));
  NCHK(hipMalloc((void**)&F.fill, sizeof(int) * F.ncell));
  NCHK(hipMalloc((void**)&F.sorted, sizeof(int) * std::max<int64_t>(ns, 1)));
  NCHK(hipMemsetAsync(F.ccnt, 0, sizeof(int) * F.ncell, s));
  NCHK(hipMemsetAsync(F.fill, 0, sizeof(int) * F.ncell, s));
  const unsigned gs = (unsigned)((ns + 255) / 256);
  if (ns > 0) k_src_cell<<<gs, 256, 0, s>>>(ns, xyzs, F.G, F.cell, F.ccnt);
  k_scan_int<<<1, 1024, 0, s>>>(F.ncell, F.ccnt, F.coff, 0);
  if (ns > 0) k_src_scatter<<<gs, 256, 0, s>>>(ns, F.cell, F.coff, F.fill, F.sorted);
  NCHK(hipGetLastError());
  return 0;
}

}  // namespace

int findnear_count_device(int64_t ns, const double* xyzs, const double* rads, int64_t ndt, int64_t nt,
                          const double* targets, int64_t* row_ptr, int64_t* nnz, hipStream_t s) {
  if (ns < 0 || nt < 0 || ndt < 3) return -6;
  if (ns == 0 || nt == 0) {
    std::vector<int64_t> ones(nt + 1, 1);
    if (row_ptr) NCHK(hipMemcpyAsync(row_ptr, ones.data(), sizeof(int64_t) * (nt + 1), hipMemcpyHostToDevice, s));
    NCHK(hipStreamSynchronize(s));
    *nnz = 0;
    return 0;
  }
  Finder F;
  int rc = build_grid(ns, xyzs, rads, F, s);
  if (rc) return rc;
  int* cnt = nullptr;
  int64_t* rp = row_ptr;
  NCHK(hipMalloc((void**)&cnt, sizeof(int) * nt));
  if (!rp) NCHK(hipMalloc((void**)&rp, sizeof(int64_t) * (nt + 1)));
  k_near_targets<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, ndt, targets, xyzs, rads, F.G, F.coff, F.sorted,
                                                               cnt, nullptr, nullptr);
  k_scan_int<<<1, 1024, 0, s>>>(nt, cnt, rp, 1);
  NCHK(hipGetLastError());
  int64_t last = 0;
  NCHK(hipMemcpyAsync(&last, rp + nt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  NCHK(hipStreamSynchronize(s));
  *nnz = last - 1;
  (void)hipFree(cnt);
  if (!row_ptr) (void)hipFree(rp);
  return 0;
}

int findnear_device(int64_t ns, const double* xyzs, const double* rads, int64_t ndt, int64_t nt,
                    const double* targets, int64_t* row_ptr, int64_t* col_ind, hipStream_t s) {
  int64_t nnz = 0;
  int rc = findnear_count_device(ns, xyzs, rads, ndt, nt, targets, row_ptr, &nnz, s);
  if (rc || nnz == 0) return rc;
  Finder F;
  rc = build_grid(ns, xyzs, rads, F, s);
  if (rc) return rc;
  k_near_targets<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, ndt, targets, xyzs, rads, F.G, F.coff, F.sorted,
                                                               nullptr, row_ptr, col_ind);
  NCHK(hipGetLastError());
  NCHK(hipStreamSynchronize(s));
  return 0;
}
