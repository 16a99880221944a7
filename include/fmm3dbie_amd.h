/*
 * fmm3dbie_amd: MI355X (gfx950) implementation of fmm3dbie's locally corrected
 * layer-potential hot path, exported as a C ABI.
 *
 * Reference-signature entry points keep the Fortran calling convention of the
 * reference (gfortran mangling: lower case + trailing underscore; every argument by
 * pointer; integer*8 -> int64_t; real*8 -> double; complex*16 -> two doubles (re, im);
 * arrays column-major; index arrays 1-based).  They take host arrays, run on the GPU
 * and return host arrays; the caller owns every buffer.  The reference has no error
 * return; here a nonzero status of the last call is available from
 * fmm3dbie_last_error().
 *
 * Supported: triangular RV patches (iptype = 1) of one order in 1..9, iquadtype = 1,
 * nker = 1, ndim = 1.
 */
#ifndef FMM3DBIE_AMD_H
#define FMM3DBIE_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- near quadrature --------------------------------------------------------------- */

/* src/helm_wrappers/helm_comb_dir.f:107-244  getnearquad_helm_comb_dir
 * zpars(3) = (k, alpha, beta); wnear(nquad) complex. */
void getnearquad_helm_comb_dir_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                const double *srcvals, const double *eps, const double *zpars,
                                const int64_t *iquadtype, const int64_t *nnz, const int64_t *row_ptr,
                                const int64_t *col_ind, const int64_t *iquad, const double *rfac0,
                                const int64_t *nquad, double *wnear);

/* src/helm_wrappers/helm_comb_dir.f:250-434  getnearquad_helm_comb_dir_eval */
void getnearquad_helm_comb_dir_eval_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                     const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                     const double *srcvals, const int64_t *ndtarg, const int64_t *ntarg,
                                     const double *targs, const int64_t *ipatch_id, const double *uvs_targ,
                                     const double *eps, const double *zpars, const int64_t *iquadtype,
                                     const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                                     const int64_t *iquad, const double *rfac0, const int64_t *nquad,
                                     double *wnear);

/* src/lap_wrappers/lap_comb_dir.f90:86-227  getnearquad_lap_comb_dir
 * dpars(2) = (alpha, beta); wnear(nquad) real. */
void getnearquad_lap_comb_dir_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                               const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                               const double *srcvals, const double *eps, const double *dpars,
                               const int64_t *iquadtype, const int64_t *nnz, const int64_t *row_ptr,
                               const int64_t *col_ind, const int64_t *iquad, const double *rfac0,
                               const int64_t *nquad, double *wnear);

/* src/lap_wrappers/lap_comb_dir.f90:228-406  getnearquad_lap_comb_dir_eval */
void getnearquad_lap_comb_dir_eval_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                    const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                    const double *srcvals, const int64_t *ndtarg, const int64_t *ntarg,
                                    const double *targs, const int64_t *ipatch_id, const double *uvs_targ,
                                    const double *eps, const double *dpars, const int64_t *iquadtype,
                                    const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                                    const int64_t *iquad, const double *rfac0, const int64_t *nquad,
                                    double *wnear);

/* ---- apply (the zgmres_guru/dgmres_guru fker signature) ----------------------------- */

/* src/helm_wrappers/helm_comb_dir.f:658-850  lpcomp_helm_comb_dir_addsub
 * sigma(npts), pot(npts) complex; FMM3D replaced by an exact tiled kernel sum. */
void lpcomp_helm_comb_dir_addsub_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                  const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                  const double *srcvals, const double *eps, const int64_t *ndd,
                                  const double *dpars, const int64_t *ndz, const double *zpars,
                                  const int64_t *ndi, const int64_t *ipars, const int64_t *nnz,
                                  const int64_t *row_ptr, const int64_t *col_ind, const int64_t *iquad,
                                  const int64_t *nquad, const int64_t *nker, const double *wnear,
                                  const int64_t *novers, const int64_t *nptso, const int64_t *ixyzso,
                                  const double *srcover, const double *whtsover, const int64_t *lwork,
                                  const double *work, const int64_t *ndim, const double *sigma, double *pot);

/* src/lap_wrappers/lap_comb_dir.f90:633-805  lpcomp_lap_comb_dir_addsub (dpars = alpha, beta) */
void lpcomp_lap_comb_dir_addsub_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                 const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                 const double *srcvals, const double *eps, const int64_t *ndd,
                                 const double *dpars, const int64_t *ndz, const double *zpars,
                                 const int64_t *ndi, const int64_t *ipars, const int64_t *nnz,
                                 const int64_t *row_ptr, const int64_t *col_ind, const int64_t *iquad,
                                 const int64_t *nquad, const int64_t *nker, const double *wnear,
                                 const int64_t *novers, const int64_t *nptso, const int64_t *ixyzso,
                                 const double *srcover, const double *whtsover, const int64_t *lwork,
                                 const double *work, const int64_t *ndim, const double *sigma, double *pot);

/* ---- solvers ------------------------------------------------------------------------ */

/* src/helm_wrappers/helm_comb_dir.f:2071-2229  helm_comb_dir_solver_guru (GMRES on the GPU) */
void helm_comb_dir_solver_guru_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                const double *srcvals, const double *eps, const double *zpars,
                                const int64_t *numit, const int64_t *ifinout, const double *rhs,
                                const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                                const int64_t *iquad, const int64_t *nquad, const int64_t *nker,
                                const double *wnear, const int64_t *novers, const int64_t *npts_over,
                                const int64_t *ixyzso, const double *srcover, const double *wover,
                                const double *eps_gmres, int64_t *niter, double *errs, double *rres,
                                double *soln);

/* src/lap_wrappers/lap_comb_dir.f90:1459-1650  lap_comb_dir_solver_guru */
void lap_comb_dir_solver_guru_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                               const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                               const double *srcvals, const double *eps, const double *dpars,
                               const int64_t *numit, const int64_t *ifinout, const double *rhs,
                               const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                               const int64_t *iquad, const int64_t *nquad, const int64_t *nker,
                               const double *wnear, const int64_t *novers, const int64_t *nptso,
                               const int64_t *ixyzso, const double *srcover, const double *whtsover,
                               const double *eps_gmres, int64_t *niter, double *errs, double *rres,
                               double *soln);

/* ---- bookkeeping utilities ---------------------------------------------------------- */

/* src/quadratures/near_field_routs.f:826-873  get_rfacs */
void get_rfacs_(const int64_t *norder, const int64_t *iptype, double *rfac, double *rfac0);
/* src/quadratures/near_field_routs.f:765-820  get_iquad_rsc */
void get_iquad_rsc_(const int64_t *npatches, const int64_t *ixyzs, const int64_t *npts, const int64_t *nnz,
                    const int64_t *row_ptr, const int64_t *col_ind, int64_t *iquad);
/* src/quadratures/near_field_routs.f:998-1057  get_near_corr_max */
void get_near_corr_max_(const int64_t *ntarg, const int64_t *row_ptr, const int64_t *nnz,
                        const int64_t *col_ind, const int64_t *npatches, const int64_t *ixyzs, int64_t *nmax);
/* src/helm_wrappers/helm_comb_dir.f:856-1208  helm_comb_dir_eval_addsub: the apply at
 * arbitrary targets (ndtarg, ntarg) with precomputed near weights; nker = ndim = ndim_p = 1,
 * idensflag = ipotflag = 0 */
void helm_comb_dir_eval_addsub_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                                const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                                const double *srcvals, const int64_t *ndtarg, const int64_t *ntarg,
                                const double *targs, const double *eps, const int64_t *ndd, const double *dpars,
                                const int64_t *ndz, const double *zpars, const int64_t *ndi, const int64_t *ipars,
                                const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                                const int64_t *iquad, const int64_t *nquad, const int64_t *nker, const double *wnear,
                                const int64_t *novers, const int64_t *nptso, const int64_t *ixyzso,
                                const double *srcover, const double *whtsover, const int64_t *lwork,
                                const double *work, const int64_t *idensflag, const int64_t *ndim,
                                const double *sigma, const int64_t *ipotflag, const int64_t *ndim_p, double *pot);
/* src/lap_wrappers/lap_comb_dir.f90:806-1188  lap_comb_dir_eval_addsub */
void lap_comb_dir_eval_addsub_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                               const int64_t *iptype, const int64_t *npts, const double *srccoefs,
                               const double *srcvals, const int64_t *ndtarg, const int64_t *ntarg,
                               const double *targs, const double *eps, const int64_t *ndd, const double *dpars,
                               const int64_t *ndz, const double *zpars, const int64_t *ndi, const int64_t *ipars,
                               const int64_t *nnz, const int64_t *row_ptr, const int64_t *col_ind,
                               const int64_t *iquad, const int64_t *nquad, const int64_t *nker, const double *wnear,
                               const int64_t *novers, const int64_t *nptso, const int64_t *ixyzso,
                               const double *srcover, const double *whtsover, const int64_t *lwork,
                               const double *work, const int64_t *idensflag, const int64_t *ndim_s,
                               const double *sigma, const int64_t *ipotflag, const int64_t *ndim_p, double *pot);
/* src/helm_wrappers/helm_comb_dir.f:440-652  helm_comb_dir_eval: layer potential at
 * targets (near list, far orders, quadrature and apply on the GPU); on-surface targets
 * give the principal value part */
void helm_comb_dir_eval_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs, const int64_t *iptype,
                         const int64_t *npts, const double *srccoefs, const double *srcvals, const int64_t *ndtarg,
                         const int64_t *ntarg, const double *targs, const int64_t *ipatch_id, const double *uvs_targ,
                         const double *eps, const double *zpars, const double *sigma, double *pot);
/* src/lap_wrappers/lap_comb_dir.f90:408-632  lap_comb_dir_eval */
void lap_comb_dir_eval_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs, const int64_t *iptype,
                        const int64_t *npts, const double *srccoefs, const double *srcvals, const int64_t *ndtarg,
                        const int64_t *ntarg, const double *targs, const int64_t *ipatch_id, const double *uvs_targ,
                        const double *eps, const double *dpars, const double *sigma, double *pot);
/* src/helm_wrappers/helm_comb_dir.f:1808-2064  helm_comb_dir_solver: setup on the GPU,
 * then the device GMRES of helm_comb_dir_solver_guru */
void helm_comb_dir_solver_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                           const int64_t *iptype, const int64_t *npts, const double *srccoefs, const double *srcvals,
                           const double *eps, const double *zpars, const int64_t *numit, const int64_t *ifinout,
                           const double *rhs, const double *eps_gmres, int64_t *niter, double *errs, double *rres,
                           double *soln);
/* src/lap_wrappers/lap_comb_dir.f90:1189-1458  lap_comb_dir_solver */
void lap_comb_dir_solver_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs, const int64_t *iptype,
                          const int64_t *npts, const double *srccoefs, const double *srcvals, const double *eps,
                          const double *dpars, const int64_t *numit, const int64_t *ifinout, const double *rhs,
                          const double *eps_gmres, int64_t *niter, double *errs, double *rres, double *soln);
/* src/quadratures/near_field_routs.f:193-423  findnearmem: nnz of the near list
 * (sources j with |x_i - c_j|^2 <= rads_j^2), computed on the GPU */
void findnearmem_(const double *xyzs, const int64_t *ns, const double *rads, const int64_t *ndt,
                  const double *targets, const int64_t *nt, int64_t *nnz);
/* src/quadratures/near_field_routs.f:424-757  findnear: row_ptr (nt+1), col_ind (nnz),
 * 1-based, ascending within a row (the order of findnearslow, :38-131) */
void findnear_(const double *xyzs, const int64_t *ns, const double *rads, const int64_t *ndt,
               const double *targets, const int64_t *nt, int64_t *row_ptr, int64_t *col_ind);
/* src/quadratures/far_field_routs.f90:10-171  get_far_order (computed on the GPU):
 * oversampling order per patch and ixyzso; zk complex (2 doubles); patch orders 1..8 */
void get_far_order_(const double *eps, const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs,
                    const int64_t *iptype, const double *cms, const double *rads, const int64_t *npts,
                    const double *srccoefs, const int64_t *ndtarg, const int64_t *ntarg, const double *targvals,
                    const int64_t *ikerorder, const double *zk, const int64_t *nnz, const int64_t *row_ptr,
                    const int64_t *col_ind, const double *rfac, int64_t *nfars, int64_t *ixyzso);
/* src/surface_routs/surf_routs.f90:304-449  oversample_geom (computed on the GPU):
 * srcover(12, nptso) at the RV nodes of order nfars(i) on patch i (ixyzso offsets) */
void oversample_geom_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs, const int64_t *iptype,
                      const int64_t *npts, const double *srccoefs, const double *srcvals, const int64_t *nfars,
                      const int64_t *ixyzso, const int64_t *nptso, double *srcover);
/* src/surface_routs/surf_routs.f90:1598-1651  get_qwts (computed on the GPU) */
void get_qwts_(const int64_t *npatches, const int64_t *norders, const int64_t *ixyzs, const int64_t *iptype,
               const int64_t *npts, const double *srcvals, double *qwts);
/* src/common/get_fmm_thresh.f:1-90  get_fmm_thresh */
void get_fmm_thresh_(const int64_t *nds, const int64_t *ns, const double *src, const int64_t *ndt,
                     const int64_t *nt, const double *trg, double *thresh);
/* src/common/sparse_reps.f:38-72  rsc_to_csc */
void rsc_to_csc_(const int64_t *ncol, const int64_t *nrow, const int64_t *nnz, const int64_t *row_ptr,
                 const int64_t *col_ind, int64_t *col_ptr, int64_t *row_ind, int64_t *iper);

/* ---- discretization tables (host; no GPU needed) ------------------------------------- */

/* src/tria_routs/koornexps.f90:57-154  get_vioreanu_nodes/get_vioreanu_wts: returns p */
int fmm3dbie_rv_nodes(int norder, double *uvs, double *wts);
/* src/tria_routs/koornexps.f90:568-601  koorn_vals2coefs: umat (p x p, column-major) */
int fmm3dbie_rv_umat(int norder, double *umat);
/* src/tria_routs/koornexps.f90:909-1016  koornf_pols */
void fmm3dbie_koorn_pols(double u, double v, int nmax, double *pols);
/* src/tria_routs/triasymq.f:89-158  triasymq on the simplex: returns the node count */
int fmm3dbie_xg_rule(int deg, double *uvs, double *wts);
/* No reference counterpart: the sine and cosine the Helmholtz kernel evaluations use (the
 * device function evaluated on the host), for |x| <= 1.6e6; returns -1 beyond that. */
int fmm3dbie_sincos_red(int64_t n, const double *x, double *s, double *c);

/* ---- device-pointer interface (inputs resident in HBM, caller's stream) -------------- */
/* No single reference counterpart: the same computations as the entries above with
 * device pointers, for callers that keep the discretization on the GPU. */

/* Status of the last call (0 = success). */
int fmm3dbie_last_error(void);

/* getnearquad_*_comb_dir(_eval) on device buffers.  kpar6 = (k re, k im, alpha re,
 * alpha im, beta re, beta im); helm = 0 selects the Laplace kernel (real weights).
 * self_nr, self_nt, self_tau_panel: Gauss-Legendre points of the self rule's panels in
 * the radial and the sinh-parameter directions, and the panels' largest sinh-parameter
 * length (<= 0: 8, 10 and 2.0).
 * stats (optional, 24 doubles): nself, nadap, nfixed, overflow, nf, kq; times in ms of
 * prep, fixed, adaptive, self, host tables, selector, plain tree cache, routed tree
 * caches; rule classes used, 0, pairs sent back to the plain integrator, pairs per rule
 * class (6), triangles of the plain rule's tree cache (85: depths 0..3; 341: depths 0..4,
 * used from 300 near pairs per patch).  nref_host/class_host/lev_host/unc_host (optional, nnz ints): refinement
 * tests per pair, pair class (0 self, 1 adaptive, 2 fixed), deepest level of the pair's
 * tree, refinement tests below the cached depths (children geometry computed).
 * mlp_params (optional, device, fmm3dbie_mlp_nparam() floats): learned order selection.
 * Every adaptive pair (ggq-quads.f:382-386) gets a rule class c = 0..5 from the network:
 * the adaptive integrator of triaintrouts.f:1269-1464 with the XG rule of degree
 * q0 + (-4, -2, 0, +4, +8, +12)[c] (q0 = the reference's degree, class 2 = the plain
 * rule; degrees kept within 8..30) and level cap route_lmax[c] (optional host array of 6; default 20); a pair whose
 * tree needs a level past its cap is integrated again by the plain integrator.
 * feat_out (optional, device, nnz x 8 floats): selector features of the adaptive pairs.
 * pred_host (optional, nnz ints): predicted class (-1 for non-adaptive pairs). */
int fmm3dbie_nearquad_dev(int helm, int norder, int norder_branch, int64_t npatches, const int64_t *ixyzs,
                          const double *srccoefs, const double *srcvals, int64_t ndtarg, int64_t ntarg,
                          const double *targs, const int64_t *ipatch_id, const double *uvs_targ, double eps,
                          const double *kpar6, int64_t nnz, const int64_t *row_ptr, const int64_t *col_ind,
                          const int64_t *iquad, double rfac0, double *wnear, int self_nr, int self_nt,
                          double self_tau_panel, void *stream, double *stats, int *nref_host, int *class_host, const float *mlp_params,
                          float *feat_out, int *pred_host, const int *route_lmax, int *lev_host, int *unc_host);

/* findnearmem / findnear on device arrays (xyzs (3,ns), rads (ns), targets (ndt,nt));
 * row_ptr may be null for the count. */
int fmm3dbie_findnearmem_dev(int64_t ns, const double *xyzs, const double *rads, int64_t ndt, int64_t nt,
                             const double *targets, int64_t *row_ptr, int64_t *nnz, void *stream);
int fmm3dbie_findnear_dev(int64_t ns, const double *xyzs, const double *rads, int64_t ndt, int64_t nt,
                          const double *targets, int64_t *row_ptr, int64_t *col_ind, void *stream);

/* get_far_order on device arrays (uniform patch order norder); novers out (npatches) */
int fmm3dbie_far_order_dev(double eps, int64_t npatches, int norder, const int64_t *ixyzs, const double *cms,
                           const double *rads, const double *srccoefs, int64_t ndtarg, int64_t ntarg,
                           const double *targs, int ikerorder, const double *zk, int64_t nnz,
                           const int64_t *row_ptr, const int64_t *col_ind, double rfac, int64_t *novers,
                           void *stream);

/* oversample_geom and the smooth weights of the oversampled nodes (get_qwts applied to
 * srcover, surf_routs.f90:1598-1651) on device arrays; uniform patch order norder <= 20,
 * nfars (npatches) <= 20. */
int fmm3dbie_oversample_geom_dev(int64_t npatches, int norder, const int64_t *ixyzs, const double *srccoefs,
                                 const double *srcvals, const int64_t *nfars, const int64_t *ixyzso,
                                 double *srcover, double *whtsover, void *stream);

/* Near-quadrature calls keep their device work buffers (grow-only, per device) and
 * per-order tables between calls and are serialised by a lock; this frees the buffers.
 * No reference counterpart. */
void fmm3dbie_release_workspace(void);
/* The host lpcomp_*_addsub_ / *_eval_addsub_ drop-ins keep their last apply plan (device
 * copies of the geometry, near weights and bookkeeping) and reuse it while every array's
 * address, length and 257 sampled entries are unchanged, as when zgmres_guru calls them
 * once per iteration; fmm3dbie_release_workspace() frees it.  This counts the plans they
 * built.  No reference counterpart. */
int64_t fmm3dbie_host_plan_builds(void);

/* Near-quadrature concurrency (no reference counterpart): with enable = 1 (default) and
 * surfaces of up to 4096 patches, the tree cache, fixed-rule and self-quadrature kernels
 * run on a library-owned second stream beside the adaptive kernels (disjoint pairs and
 * weights; the call returns after both streams); 0 runs every kernel in order on the
 * caller's stream (per-kernel timings undisturbed). */
void fmm3dbie_set_side_stream(int enable);

/* GMRES progress (the reference's zgmres_guru prints "iter=", it, errs(it) every
 * iteration, gmres_routs.f90:303): every > 0 prints the iteration number and the largest
 * relative residual estimate over the systems to stderr every `every` iterations, and a
 * line per refinement of fmm3dbie_gmres_ir_dev; 0 (default) prints nothing. */
void fmm3dbie_set_gmres_print(int every);

/* ---- learned order selection (no reference counterpart) ------------------------------
 * MLP 8 -> 64 -> 64 -> 6 (ReLU, softmax), FP32, parameters packed as W1[64][8] b1[64]
 * W2[64][64] b2[64] W3[6][64] b3[6]; X is n x 8 row-major, y class ids 0..5 (the rule
 * class of fmm3dbie_nearquad_dev whose adaptive integration of the pair evaluates the
 * fewest kernel values). */
int fmm3dbie_mlp_nparam(void);
/* mean softmax cross-entropy over n samples and its gradient (device pointers) */
int fmm3dbie_mlp_grad(const float *params, const float *X, const int *y, int64_t n, float *grad, float *loss,
                      void *stream);
/* minibatch Adam (batch a multiple of 64), reshuffled every epoch from seed;
 * loss_hist (host, epochs floats, optional): last minibatch loss of each epoch */
int fmm3dbie_mlp_train(float *params, const float *X, const int *y, int64_t n, int epochs, int batch, float lr,
                       unsigned seed, float *loss_hist, void *stream);
/* argmax class (and optionally the n x 6 logits) */
int fmm3dbie_mlp_predict(const float *params, const float *X, int64_t n, int *cls, float *logits, void *stream);

/* Device-resident apply plan for lpcomp_*_addsub (host bookkeeping arrays; wnear on the
 * host or, with wnear_on_device = 1, already in HBM). */
void *fmm3dbie_apply_plan_create(int helm, int64_t npatches, const int64_t *norders, const int64_t *ixyzs,
                                 int64_t npts, const double *srcvals, const double *kpar6, int64_t nnz,
                                 const int64_t *row_ptr, const int64_t *col_ind, const int64_t *iquad,
                                 int64_t nquad, const double *wnear, int wnear_on_device, const int64_t *novers,
                                 int64_t nptso, const int64_t *ixyzso, const double *srcover,
                                 const double *whtsover, void *stream);
/* prec 64: complex128 sigma/pot; prec 32: complex64 (near patches masked in the tiled
 * sum).  mode 0: add/subtract as the reference; 1: masked.  nrhs <= 64 columns of
 * length npts.  times (optional, 4 doubles): oversample, far, csr, subtract ms. */
int fmm3dbie_apply_plan_apply(void *plan, int prec, int mode, const void *sigma, void *pot, int nrhs,
                              void *stream, double *times);
/* No reference counterpart: fmm3dbie_apply_plan_apply for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc/hipFree or allocating library routine: the strength
 * and partial-sum buffers live in one fixed per-device array of
 * fmm3dbie_apply_scratch_bytes() bytes (64 MiB), held by each GPU the library runs on
 * until the process ends.  It needs 2 x nptso x nrhs + ntarg x nrhs complex values (ntarg
 * = the plan's targets; 16 B each at prec 64, 8 B at prec 32) and returns -7 when they do not fit.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (different streams or
 * threads) race; calls in one stream are safe.  The plan itself is built by the
 * allocating fmm3dbie_apply_plan_create.  No timing output. */
int fmm3dbie_apply_plan_apply_noalloc(void *plan, int prec, int mode, const void *sigma, void *pot, int nrhs,
                                      void *stream);
size_t fmm3dbie_apply_scratch_bytes(void);
int fmm3dbie_apply_plan_info(void *plan, int64_t *out4);
void fmm3dbie_apply_plan_destroy(void *plan);
/* Autotuner of the all-pairs far-sum launch (no reference counterpart): times every
 * right-hand-side group size dividing nrhs (FP64 1..8, FP32 1..16) and four workgroup
 * targets (4096..32768), keeps the fastest for later applies with the same (prec, nrhs); report gets
 * (group, workgroup target, ms) triples (nreport doubles).  Returns the number timed. */
int fmm3dbie_apply_plan_tune(void *plan, int prec, int nrhs, void *stream, double *report, int nreport);

/* src/common/gmres_routs.f90:13-371  zgmres_guru on a device plan (nrhs systems in
 * lockstep, one batched apply per iteration). */
int fmm3dbie_gmres_dev(void *plan, int prec, int mode, const double *zid, const void *rhs, void *soln,
                       const double *wts, int nrhs, int numit, double eps_gmres, int *niter, double *errs,
                       double *rres, void *stream, int *stats);
/* No reference counterpart: FP32/complex64 inner GMRES with FP64 residual refinement. */
int fmm3dbie_gmres_ir_dev(void *plan, int mode, const double *zid, const void *rhs, void *soln,
                          const double *wts, int nrhs, int numit, double eps_inner, double tol, int max_outer,
                          int *outer, int *inner_total, double *rres, void *stream, int *stats);

#ifdef __cplusplus
}
#endif
#endif
