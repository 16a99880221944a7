/*
 * fmm3dbie CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C (C99 + OpenMP) FP64 restatement of the reference's locally corrected
 * layer-potential path.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only as the checker or as the timed CPU baseline.
 * The product (fmm3dbie_amd/) never links, loads or calls anything here.
 *
 * Conventions: arrays are column-major as in the reference; index arrays passed in
 * (ixyzs, row_ptr, col_ind, iquad, ixyzso) are 1-based exactly as the reference's
 * Fortran callers pass them.  Complex values are interleaved (re, im) doubles.
 *
 * Parity status of each function is stated in its comment in fmm3dbie_oracle.c.
 */
#ifndef FMM3DBIE_ORACLE_H
#define FMM3DBIE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel kinds */
enum {
  ORC_HELM_COMB = 0, ORC_HELM_SLP = 1, ORC_HELM_DLP = 2,
  ORC_LAP_COMB = 3, ORC_LAP_SLP = 4, ORC_LAP_DLP = 5,
  ORC_ONE_OVER_R = 6
};
/* self-quadrature modes */
enum { ORC_SELF_ADAP = 0, ORC_SELF_POLAR = 1 };

int orc_npols(int norder);
void orc_koorn_pols(double u, double v, int nmax, double *pols);
void orc_koorn_pols_many(int nmax, int64_t npts, const double *uv, double *pols);
void orc_set_num_threads(int n);
int orc_max_threads(void);
int orc_rv_nodes(int norder, double *uvs, double *wts);
int orc_rv_umat(int norder, double *umat);
int orc_xg_rule(int deg, double *uvs, double *wts);
int orc_xg_unif_nodes(int nlev, int deg, double *uvs, double *wts);
void orc_quadparams_adap(double eps, int norder, int *nqorder, double *eps_adap,
                         int *nlev, int *nqorder_f);
void orc_rfacs(int norder, double *rfac, double *rfac0);

void orc_kernel(int kind, const double *zpars, const double *src, const double *targ,
                double *val);

void orc_triaints_wnodes(int norder, int npols, const double *srccoefs, int ndtarg,
                         int ntarg, const double *targs, int nq, const double *qnodes,
                         const double *qwts, int kind, const double *zpars, double *cint);
int orc_triaints_adap(double eps, int nqorder, int norder, int npols,
                      const double *srccoefs, int ndtarg, const double *targ, int kind,
                      const double *zpars, double *cint, int *nrefined, int *maxlev);

void orc_self_adap(double epsp, int norder, int npols, const double *uv,
                   const double *srccoefs, int ndtarg, const double *targ, int kind,
                   const double *zpars, int nqorder, double *zquad);
void orc_self_polar(int norder, int npols, const double *uv, const double *srccoefs,
                    int ndtarg, const double *targ, int kind, const double *zpars,
                    int nr, int nt, double *zquad);

void orc_centroid_rads(int64_t npatches, const int64_t *norders, const int64_t *ixyzs,
                       const double *srccoefs, double *cms, double *rads);

void orc_getnearquad(int64_t npatches, const int64_t *norders, const int64_t *ixyzs,
                     const int64_t *iptype, int64_t npts, const double *srccoefs,
                     const double *srcvals, int64_t ndtarg, int64_t ntarg,
                     const double *targs, const int64_t *ipatch_id, const double *uvs_targ,
                     double eps, int kind, const double *zpars, int64_t nnz,
                     const int64_t *row_ptr, const int64_t *col_ind, const int64_t *iquad,
                     double rfac0, int64_t nquad, double *wnear, int selfmode,
                     int self_nr, int self_nt, int64_t *pair_class, int32_t *pair_nref);

void orc_oversample_fun(int nd, int64_t npatches, const int64_t *norders,
                        const int64_t *ixyzs, const double *u, const int64_t *novers,
                        const int64_t *ixyzso, double *uover);
void orc_oversample_geom(int64_t npatches, const int64_t *norders, const int64_t *ixyzs,
                         const double *srccoefs, const double *srcvals,
                         const int64_t *novers, const int64_t *ixyzso, double *srcover,
                         double *whtsover);

void orc_lpcomp_addsub(int cplx, int64_t npatches, const int64_t *norders,
                       const int64_t *ixyzs, int64_t npts, const double *srcvals,
                       int64_t ndtarg, int64_t ntarg, const double *targs,
                       const double *zpars, int64_t nnz, const int64_t *row_ptr,
                       const int64_t *col_ind, const int64_t *iquad, const double *wnear,
                       const int64_t *novers, int64_t nptso, const int64_t *ixyzso,
                       const double *srcover, const double *whtsover, const double *sigma,
                       double *pot, int64_t tgt_stride);

int orc_gmres(int cplx, int64_t npatches, const int64_t *norders, const int64_t *ixyzs,
              int64_t npts, const double *srcvals, const double *wts, const double *zpars,
              int64_t nnz, const int64_t *row_ptr, const int64_t *col_ind,
              const int64_t *iquad, const double *wnear, const int64_t *novers,
              int64_t nptso, const int64_t *ixyzso, const double *srcover,
              const double *whtsover, const double *zid, const double *rhs, int numit,
              double eps_gmres, int *niter, double *errs, double *rres, double *soln);

void orc_findnear(const double *cms, int64_t ns, const double *rads, int64_t ndt,
                  const double *targs, int64_t nt, int64_t *row_ptr, int64_t *col_ind,
                  int64_t nnz_cap, int64_t *nnz_out);
double orc_dlaran(int *iseed);
double orc_hkrand(int seed);
void orc_far_order(double eps, int64_t npatches, const int64_t *norders,
                   const int64_t *ixyzs, const double *cms, const double *rads,
                   const double *srccoefs, int64_t ndtarg, int64_t ntarg,
                   const double *targs, int ikerorder, const double *zk, int64_t nnz,
                   const int64_t *row_ptr, const int64_t *col_ind, double rfac,
                   int64_t *novers, int64_t *ixyzso);

#ifdef __cplusplus
}
#endif
#endif
