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
