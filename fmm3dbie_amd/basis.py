"""Host-side discretization tables of the product library (RV nodes/weights, the
values->coefficients matrix, Koornwinder polynomials, XG rules), read through the
C-ABI of libfmm3dbie_amd.so.  No GPU is needed for these."""
import ctypes as C

import numpy as np

from ._lib import lib


def npols(n):
    return (n + 1) * (n + 2) // 2


def rv_nodes(norder):
    p = npols(norder)
    uvs = np.zeros((p, 2))
    w = np.zeros(p)
    if lib().fmm3dbie_rv_nodes(norder, uvs.ctypes.data, w.ctypes.data) != p:
        raise ValueError("RV order %d out of range" % norder)
    return uvs, w


def rv_umat(norder):
    """umat[i, j] = umat(i+1, j+1): maps values at RV nodes to Koornwinder coefficients."""
    p = npols(norder)
    out = np.zeros(p * p)
    if lib().fmm3dbie_rv_umat(norder, out.ctypes.data) != 0:
        raise ValueError("singular Vandermonde at order %d" % norder)
    return out.reshape(p, p).T.copy()


def koorn_pols(u, v, nmax):
    out = np.zeros(npols(nmax))
    lib().fmm3dbie_koorn_pols(C.c_double(u), C.c_double(v), nmax, out.ctypes.data)
    return out


def xg_rule(deg):
    uv = np.zeros((600, 2))
    w = np.zeros(600)
    n = lib().fmm3dbie_xg_rule(deg, uv.ctypes.data, w.ctypes.data)
    if n < 0:
        raise ValueError("XG degree %d out of range" % deg)
    return uv[:n].copy(), w[:n].copy()
