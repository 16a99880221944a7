// Device-pointer interface of the near-quadrature pipeline (quad.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct NearQuadArgs {
  bool helm;              // Helmholtz (complex weights) or Laplace (real weights)
  int norder;             // patch order (uniform over patches)
  int norder_branch;      // order used by the parameter table (average patch order)
  int64_t npatches;
  const int64_t* ixyzs;   // device, 1-based (npatches+1)
  const double* srccoefs; // device (9, npts)
  const double* srcvals;  // device (12, npts)
  int64_t ndtarg, ntarg;
  const double* targs;    // device (ndtarg, ntarg)
  const int64_t* ipatch_id;  // device (ntarg) or null (all off-surface)
  const double* uvs_targ;    // device (2, ntarg) or null
  double eps;
  double zk[2], alpha[2], beta[2];
  int64_t nnz;
  const int64_t* row_ptr;  // device, 1-based
  const int64_t* col_ind;  // device, 1-based
  const int64_t* iquad;    // device, 1-based
  double rfac0;
  double* wnear;           // device, nquad complex (interleaved) or real
  int self_n;              // Gauss-Legendre points per direction of the self rule
  double self_tau_panel;   // max sinh-parameter panel length of the self rule
  int adaptive_waves;      // persistent wavefronts of the adaptive kernel (0 = default)
  bool want_nref;
  int* nref_host;          // optional: refinements per pair (nnz)
  int* class_host;         // optional: pair class (nnz): 0 self, 1 adaptive, 2 fixed
  // learned refinement-depth selector (orders 1..6): MLP parameters in HBM with the
  // feature normalisation folded into the first layer; null = plain adaptive path
  const float* mlp_params;
  float* feat_out;         // optional, device (nnz, 8): selector features of adaptive pairs
  int* pred_host;          // optional: predicted depth class per pair (-1 if not adaptive)
};

struct NearQuadStats {
  int nself, nadap, nfixed, overflow, nf, kq;
  double ms_prep, ms_fixed, ms_adap, ms_self;
  long long nref_total;
  // learned path: pairs accepted by the depth-1 / depth-2 cached-tree kernels and pairs
  // sent back to the adaptive kernel.  Times: host tables + uploads, features + inference
  // + bucketing, depth-0..2 geometry cache, cached-tree kernels; ms_adap is the adaptive
  // kernel alone.
  int nshallow0, nshallow1, nfallback;
  double ms_tables, ms_select, ms_tree, ms_shallow;
};

// Calls are serialised; device work buffers and per-order tables are kept between calls
// (grow-only) until fmm_nearquad_release().
int fmm_nearquad_device(const NearQuadArgs& a, NearQuadStats* st, hipStream_t s);
void fmm_nearquad_release();
// 1 (default): the fixed-rule and self kernels run on a second stream beside the tree
// cache and the adaptive kernels; 0: everything on the caller's stream, in order
void fmm_nearquad_set_side_stream(int enable);
int fmm_koorn_constants();
