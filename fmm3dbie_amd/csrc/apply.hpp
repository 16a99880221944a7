// Device-resident apply plan (apply.hip): everything lpcomp_*_addsub needs, uploaded
// once and reused across matvecs (GMRES) and right-hand sides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <vector>

enum { APPLY_ADDSUB = 0, APPLY_MASKED = 1 };

struct ApplyPlanInput {
  bool helm;
  int64_t npatches, npts, nptso, nnz, nquad;
  const int64_t *norders, *ixyzs, *novers, *ixyzso;  // host
  const int64_t *row_ptr, *col_ind, *iquad;          // host, 1-based
  const double *srcvals, *srcover, *whtsover;        // host
  const double* wnear;    // complex (interleaved) if helm, else real
  // targets (host, (ndtarg, ntarg)); null: the discretization points (srcvals)
  const double* targs = nullptr;
  int64_t ndtarg = 12, ntarg = -1;
  bool wnear_is_device;   // wnear already in device memory
  double zk[2], alpha[2], beta[2];
};

struct ApplyTimes {
  double ms_over, ms_far, ms_csr, ms_sub;
};

struct ApplyPlan {
  bool helm = true;
  int64_t npatches = 0, npts = 0, nptso = 0, nnz = 0, nquad = 0;
  int64_t ntarg = 0;  // targets (rows of the near list); npts for the on-surface operator
  double zk[2] = {0, 0}, alpha[2] = {0, 0}, beta[2] = {0, 0};
  double thresh = 0;
  bool phase_small = true;  // FP32 phases within the hardware sine's range (no reduction)
  int uniform_p = 0;
  int64_t *ixyzs = nullptr, *ixyzso = nullptr, *row_ptr = nullptr, *col_ind = nullptr, *iquad = nullptr,
          *ioff = nullptr;
  int *np_of = nullptr, *opatch = nullptr, *near_sorted = nullptr;
  int* csr_src = nullptr;  // per CSR entry: first point (0-based) of its source patch
  double* imat = nullptr;
  double* wover = nullptr;
  float* wover32 = nullptr;
  double* src64[6] = {};
  float* src32[6] = {};
  double* trg64[3] = {};
  float* trg32[3] = {};
  double* wnear64 = nullptr;  // complex128 (Laplace weights with zero imaginary part)
  float* wnear32 = nullptr;   // complex64 copy for the mixed-precision path
  // workspace
  int ws_cap = 0;
  double *q64 = nullptr, *m64 = nullptr, *sig64c = nullptr, *pot64c = nullptr, *part64 = nullptr;
  float *q32 = nullptr, *m32 = nullptr, *part32 = nullptr;
  // far-sum launch shape: target workgroups x source splits >= far_wg_target (many short
  // workgroups, ~128 per CU, balance the 256 CUs better than one long round; the partial
  // sums of the splits are added by one coalesced pass)
  int64_t far_wg_target = 32768;
  int far_split_cap = 8;
  // autotuned far-sum shape per (prec * 1000 + nrhs): (RHS group size, workgroup target)
  std::map<int, std::pair<int, int64_t>> tuned;
  std::vector<void*> owned;
  ~ApplyPlan();
};

int apply_plan_build(ApplyPlan* P, const ApplyPlanInput& in, hipStream_t s);
// prec 64: complex128 sigma/pot; prec 32: complex64 (always masked)
int apply_plan_run(ApplyPlan* P, int prec, int mode, const void* sigma, void* pot, int nrhs, hipStream_t s,
                   ApplyTimes* tm);
int apply_plan_run_real(ApplyPlan* P, int mode, const double* sigma, double* pot, hipStream_t s);
// apply_plan_run with its work buffers in a fixed per-device scratch array instead of
// plan-owned allocations: calls no allocator; -7 when the buffers do not fit
int apply_plan_run_noalloc(ApplyPlan* P, int prec, int mode, const void* sigma, void* pot, int nrhs, hipStream_t s);
size_t apply_scratch_bytes();
void apply_plan_free_ws(ApplyPlan* P);
// time the far-sum launch shapes for (prec, nrhs) and keep the fastest; report gets
// (group, workgroup target, ms) triples; returns the number of shapes timed
int apply_plan_tune(ApplyPlan* P, int prec, int nrhs, hipStream_t s, double* report, int nreport);
