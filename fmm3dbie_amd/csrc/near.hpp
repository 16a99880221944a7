// Near-field finder (near.hip): CSR of the sources within r_j of each target.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

// Device arrays: xyzs (3, ns), rads (ns), targets (ndt, nt).  row_ptr (nt + 1, 1-based;
// may be null for the count) and col_ind (nnz, 1-based, ascending within a row).
int findnear_count_device(int64_t ns, const double* xyzs, const double* rads, int64_t ndt, int64_t nt,
                          const double* targets, int64_t* row_ptr, int64_t* nnz, hipStream_t s);
int findnear_device(int64_t ns, const double* xyzs, const double* rads, int64_t ndt, int64_t nt,
                    const double* targets, int64_t* row_ptr, int64_t* col_ind, hipStream_t s);
