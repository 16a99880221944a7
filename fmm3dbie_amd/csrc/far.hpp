// Far-field oversampling order (far.hip), get_far_order semantics on device arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct FarOrderArgs {
  double eps;
  int64_t npatches;
  int norder;                 // uniform patch order (1..8)
  const int64_t* ixyzs;       // device (npatches + 1), 1-based
  const double* cms;          // device (3, npatches)
  const double* rads;         // device (npatches)
  const double* srccoefs;     // device (9, npts)
  int64_t ndtarg, ntarg;
  const double* targs;        // device (ndtarg, ntarg)
  int ikerorder;              // -1: single layer, else double layer
  double zk[2];
  int64_t nnz;
  const int64_t* row_ptr;     // device (ntarg + 1), 1-based
  const int64_t* col_ind;     // device (nnz), 1-based
  double rfac;
  int64_t* novers;            // device out (npatches)
};

int far_order_device(const FarOrderArgs& a, hipStream_t s);

// oversample_geom + the smooth weights of the oversampled nodes (get_qwts on srcover) on
// device arrays; nfars (npatches) are the oversampling orders (<= 20), ixyzso their point
// offsets (1-based); srcover (12, nptso), wover (nptso) out.  norder < 0: every patch
// copies its own nodes (nfars = the patch orders; wover = get_qwts).
int oversample_geom_device(int64_t npatches, int norder, const int64_t* ixyzs, const double* srccoefs,
                           const double* srcvals, const int64_t* nfars, const int64_t* ixyzso, double* srcover,
                           double* wover, hipStream_t s);
