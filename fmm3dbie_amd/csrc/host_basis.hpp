// Host-side construction of the discretization tables used by the kernels:
// RV nodes/weights, the values->coefficients matrix (umat), XG rules, the fixed
// oversampled rule, Gauss-Legendre rules and the Koornwinder recurrence constants.
#pragma once
#include <vector>

#include "common.hpp"

namespace fmmhost {

inline int npols(int n) { return (n + 1) * (n + 2) / 2; }

// Koornwinder polynomials at (u,v), all degrees <= n (koornf_pols ordering).
void koorn_pols(double u, double v, int n, double* pols);
// RV nodes (2*p, interleaved u,v) and weights of order n (get_vioreanu_nodes/wts).
int rv_nodes(int n, double* uvs, double* wts);
// umat(i,j) column-major p x p: inverse of V(i,j) = P_j(node_i).
int rv_umat(int n, std::vector<double>& umat);
// XG rule of degree deg on the simplex.
int xg_rule(int deg, std::vector<double>& uv, std::vector<double>& w);
// gen_xg_unif_nodes_tri: XG(deg) on 4^nlev congruent sub-triangles.
int xg_unif_nodes(int nlev, int deg, std::vector<double>& uv, std::vector<double>& w);
// Gauss-Legendre on [0,1].
void gauss_legendre01(int n, double* x, double* w);
// get_quadparams_adap (near_field_routs.f:883-992) with the branch variable = norder.
void quadparams_adap(double eps, int norder, int* nqorder, double* eps_adap, int* nlev,
                     int* nqorder_f);
void rfacs(int norder, double* rfac, double* rfac0);
// Recurrence constants for the device Koornwinder evaluation.
void fill_koorn_const(KoornConst* kc);
// Interpolation from order-norder RV values to order-nover RV nodes, row-major
// (po x p): out[i*p + l].
void interp_matrix(int norder, int nover, std::vector<double>& imat);
// LAPACK's dlaran (48-bit multiplicative congruential generator, 12-bit limbs) and the
// reference's hkrand wrapper for a nonzero seed (src/common/hkrand.f:1-19).
double dlaran(int iseed[4]);
double hkrand(int seed);

}  // namespace fmmhost
