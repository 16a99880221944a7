// Device-pointer interface of the near-quadrature pipeline (quad.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Default self rule: panels of at most SELF_TAU in the sinh parameter, SELF_NR x SELF_NT
// Gauss-Legendre points (radial x sinh parameter).  Against a 16 x 16 rule on panels of
// 0.5, the combined-field weights of every target of sampled patches agree to 1.7e-9
// (cube sphere order 4), 1.0e-8 (wavy torus order 6) and 2.3e-8 (deformed sphere order
// 6) of the target's largest weight, as with 8 x 8 points on panels of 1, with 24% fewer
// nodes.
constexpr int SELF_NR = 8;
constexpr int SELF_NT = 10;
constexpr double SELF_TAU = 2.0;

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
  int self_nr, self_nt;    // Gauss-Legendre points of the self rule's panels (radial, sinh parameter)
  double self_tau_panel;   // max sinh-parameter panel length of the self rule
  int adaptive_waves;      // persistent wavefronts of the adaptive kernel (0 = default)
  bool want_nref;
  int* nref_host;          // optional: refinements per pair (nnz)
  int* class_host;         // optional: pair class (nnz): 0 self, 1 adaptive, 2 fixed
  // learned order selection: MLP parameters in HBM (feature normalisation folded into
  // the first layer) predicting the rule class of every adaptive pair; null = plain path
  const float* mlp_params;
  float* feat_out;         // optional, device (nnz, 8): selector features of adaptive pairs
  int* pred_host;          // optional: predicted rule class per pair (-1 if not adaptive)
  const int* route_lmax;   // optional, host (6): deepest level of each rule class (default 20)
  int* lev_host;           // optional: deepest level of every adaptive pair's tree (nnz)
  int* unc_host;           // optional: refinement tests below the cached tree per pair (nnz)
};

struct NearQuadStats {
  int nself, nadap, nfixed, overflow, nf, kq;
  double ms_prep, ms_fixed, ms_adap, ms_self;
  long long nref_total;
  // learned path: rule classes with pairs, pairs per class, pairs sent back to the plain
  // integrator.  Times: host tables + uploads, features + inference + bucketing, the plain
  // tree cache, the routed classes' tree caches; ms_adap is the adaptive kernel launches.
  int nroute, nfallback, ncls[6];
  double ms_tables, ms_select, ms_tree, ms_shallow;
  int cache_tris;  // triangles of the plain rule's tree cache (85: depths 0..3, 341: 0..4)
};

// Calls are serialised; device work buffers and per-order tables are kept between calls
// (grow-only) until fmm_nearquad_release().
int fmm_nearquad_device(const NearQuadArgs& a, NearQuadStats* st, hipStream_t s);
void fmm_nearquad_release();
// 1 (default): the fixed-rule and self kernels run on a second stream beside the tree
// cache and the adaptive kernels; 0: everything on the caller's stream, in order
void fmm_nearquad_set_side_stream(int enable);
int fmm_koorn_constants();
