// Device GMRES and mixed-precision iterative refinement (gmres.hip).
#pragma once
#include <hip/hip_runtime.h>

struct ApplyPlan;

struct GmresStats {
  int napply = 0;    // applies at the solver's precision
  int napply64 = 0;  // FP64 residual applies of iterative refinement
};

// rhs/soln: complex (prec 64: complex128, prec 32: complex64), npts x nrhs, device.
// wts: quadrature weights (npts, device, double).  niter/rres: nrhs; errs: nrhs*numit.
int gmres_device(ApplyPlan* P, int prec, int mode, const double* zid, const void* rhs, void* soln,
                 const double* wts, int nrhs, int numit, double eps_gmres, int* niter, double* errs, double* rres,
                 hipStream_t s, GmresStats* gs);
int gmres_ir_device(ApplyPlan* P, int mode, const double* zid, const void* rhs, void* soln, const double* wts,
                    int nrhs, int numit, double eps_inner, double tol, int max_outer, int* outer, int* inner_total,
                    double* rres, hipStream_t s, GmresStats* gs);
// iteration lines on stderr every `every` iterations (0: none)
void gmres_set_print(int every);
