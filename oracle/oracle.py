"""ctypes front-end of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
Array convention: a Fortran column-major array ``a(m, n)`` is a C-contiguous numpy
array of shape ``(n, m)``; complex arrays are numpy complex128.  Index arrays follow
the reference (1-based).
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

HELM_COMB, HELM_SLP, HELM_DLP, LAP_COMB, LAP_SLP, LAP_DLP, ONE_OVER_R = range(7)
SELF_ADAP, SELF_POLAR = 0, 1

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", str(_HERE)])


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = C.CDLL(str(path))
        _sig(_LIB)
    return _LIB


def _sig(L):
    vp = C.c_void_p
    L.orc_koorn_pols.argtypes = [C.c_double, C.c_double, C.c_int, _d]
    L.orc_koorn_pols_many.argtypes = [C.c_int, C.c_int64, _d, _d]
    L.orc_set_num_threads.argtypes = [C.c_int]
    L.orc_max_threads.restype = C.c_int
    L.orc_rv_nodes.argtypes = [C.c_int, vp, vp]
    L.orc_rv_umat.argtypes = [C.c_int, _d]
    L.orc_xg_rule.argtypes = [C.c_int, _d, _d]
    L.orc_xg_unif_nodes.argtypes = [C.c_int, C.c_int, _d, _d]
    L.orc_quadparams_adap.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_kernel.argtypes = [C.c_int, _d, _d, _d, _d]
    L.orc_triaints_wnodes.argtypes = [C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_int, _d, _d,
                                      C.c_int, _d, _d]
    L.orc_triaints_adap.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, _d, C.c_int, _d, C.c_int,
                                    _d, _d, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_self_adap.argtypes = [C.c_double, C.c_int, C.c_int, _d, _d, C.c_int, _d, C.c_int, _d,
                                C.c_int, _d]
    L.orc_self_polar.argtypes = [C.c_int, C.c_int, _d, _d, C.c_int, _d, C.c_int, _d, C.c_int,
                                 C.c_int, _d]
    L.orc_centroid_rads.argtypes = [C.c_int64, _i64, _i64, _d, _d, _d]
    L.orc_getnearquad.argtypes = [C.c_int64, _i64, _i64, _i64, C.c_int64, _d, _d, C.c_int64,
                                  C.c_int64, _d, _i64, _d, C.c_double, C.c_int, _d, C.c_int64,
                                  _i64, _i64, _i64, C.c_double, C.c_int64, _d, C.c_int, C.c_int,
                                  C.c_int, vp, vp]
    L.orc_oversample_fun.argtypes = [C.c_int, C.c_int64, _i64, _i64, _d, _i64, _i64, _d]
    L.orc_oversample_geom.argtypes = [C.c_int64, _i64, _i64, _d, _d, _i64, _i64, _d, _d]
    L.orc_lpcomp_addsub.argtypes = [C.c_int, C.c_int64, _i64, _i64, C.c_int64, _d, C.c_int64,
                                    C.c_int64, _d, _d, C.c_int64, _i64, _i64, _i64, _d, _i64,
                                    C.c_int64, _i64, _d, _d, _d, _d, C.c_int64]
    L.orc_gmres.argtypes = [C.c_int, C.c_int64, _i64, _i64, C.c_int64, _d, _d, _d, C.c_int64,
                            _i64, _i64, _i64, _d, _i64, C.c_int64, _i64, _d, _d, _d, _d, C.c_int,
                            C.c_double, C.POINTER(C.c_int), _d, C.POINTER(C.c_double), _d]
    L.orc_findnear.argtypes = [_d, C.c_int64, _d, C.c_int64, _d, C.c_int64, _i64, vp, C.c_int64,
                               C.POINTER(C.c_int64)]
    L.orc_dlaran.argtypes = [np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
    L.orc_dlaran.restype = C.c_double
    L.orc_hkrand.argtypes = [C.c_int]
    L.orc_hkrand.restype = C.c_double
    L.orc_far_order.argtypes = [C.c_double, C.c_int64, _i64, _i64, _d, _d, _d, C.c_int64,
                                C.c_int64, _d, C.c_int, _d, C.c_int64, _i64, _i64, C.c_double,
                                _i64, _i64]


def npols(n):
    return (n + 1) * (n + 2) // 2


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c2d(z):
    return np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).view(np.float64)


def koorn_pols(u, v, nmax):
    out = np.zeros(npols(nmax))
    lib().orc_koorn_pols(u, v, nmax, out)
    return out


def koorn_pols_many(uv, nmax):
    """(npts, npols) basis values at the rows of uv (npts, 2)."""
    uv = _f(uv)
    out = np.zeros((len(uv), npols(nmax)))
    lib().orc_koorn_pols_many(nmax, len(uv), uv, out)
    return out


def rv_nodes(norder):
    p = npols(norder)
    uvs = np.zeros((p, 2)); w = np.zeros(p)
    lib().orc_rv_nodes(norder, uvs.ctypes.data, w.ctypes.data)
    return uvs, w


def rv_umat(norder):
    """umat as numpy (p, p) with umat[i, j] = umat(i+1, j+1)."""
    p = npols(norder)
    out = np.zeros(p * p)
    lib().orc_rv_umat(norder, out)
    return out.reshape(p, p).T.copy()


def xg_rule(deg):
    n = 500
    uv = np.zeros((n, 2)); w = np.zeros(n)
    k = lib().orc_xg_rule(deg, uv, w)
    return uv[:k].copy(), w[:k].copy()


def xg_unif_nodes(nlev, deg):
    n = 4 ** nlev * 500
    uv = np.zeros((n, 2)); w = np.zeros(n)
    k = lib().orc_xg_unif_nodes(nlev, deg, uv, w)
    return uv[:k].copy(), w[:k].copy()


def quadparams_adap(eps, norder):
    a, c, d = C.c_int(), C.c_int(), C.c_int()
    b = C.c_double()
    lib().orc_quadparams_adap(eps, norder, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return a.value, b.value, c.value, d.value


def kernel(kind, zpars, src, targ):
    out = np.zeros(2)
    lib().orc_kernel(kind, _c2d(zpars) if np.iscomplexobj(zpars) else _f(zpars), _f(src), _f(targ), out)
    return out[0] + 1j * out[1]


def _kp(kind, zpars):
    if kind in (LAP_COMB, LAP_SLP, LAP_DLP, ONE_OVER_R):
        return _f(np.real(np.atleast_1d(zpars)) if zpars is not None else np.zeros(2))
    return _c2d(zpars)


def triaints_wnodes(norder, srccoefs, targs, qnodes, qwts, kind, zpars):
    p = npols(norder)
    targs = _f(targs)
    nt, ndt = targs.shape
    out = np.zeros((nt, p, 2))
    lib().orc_triaints_wnodes(norder, p, _f(srccoefs), ndt, nt, targs, len(qwts), _f(qnodes),
                              _f(qwts), kind, _kp(kind, zpars), out)
    return out[..., 0] + 1j * out[..., 1]


def triaints_adap(eps, nqorder, norder, srccoefs, targ, kind, zpars):
    p = npols(norder)
    targ = _f(targ)
    out = np.zeros((p, 2))
    nref, ml = C.c_int(), C.c_int()
    lib().orc_triaints_adap(eps, nqorder, norder, p, _f(srccoefs), len(targ), targ, kind,
                            _kp(kind, zpars), out, C.byref(nref), C.byref(ml))
    return out[:, 0] + 1j * out[:, 1], nref.value, ml.value


def self_adap(epsp, norder, uv, srccoefs, targ, kind, zpars, nqorder):
    p = npols(norder)
    out = np.zeros((p, 2))
    targ = _f(targ)
    lib().orc_self_adap(epsp, norder, p, _f(uv), _f(srccoefs), len(targ), targ, kind,
                        _kp(kind, zpars), nqorder, out)
    return out[:, 0] + 1j * out[:, 1]


def _set_self_tau(tau):
    C.c_double.in_dll(lib(), "orc_self_tau_panel").value = tau


def self_polar(norder, uv, srccoefs, targ, kind, zpars, nr, nt, tau=2.0):
    """The self rule at one target: nr x nt Gauss-Legendre points on panels of at most tau
    in the sinh-graded edge parameter."""
    p = npols(norder)
    out = np.zeros((p, 2))
    targ = _f(targ)
    _set_self_tau(tau)
    lib().orc_self_polar(norder, p, _f(uv), _f(srccoefs), len(targ), targ, kind,
                         _kp(kind, zpars), nr, nt, out)
    return out[:, 0] + 1j * out[:, 1]


def centroid_rads(norders, ixyzs, srccoefs):
    npat = len(norders)
    cms = np.zeros((npat, 3)); rads = np.zeros(npat)
    lib().orc_centroid_rads(npat, np.asarray(norders, np.int64), np.asarray(ixyzs, np.int64),
                            _f(srccoefs), cms, rads)
    return cms, rads


def getnearquad(geom, targs, ipatch_id, uvs_targ, eps, kind, zpars, row_ptr, col_ind, iquad,
                rfac0, selfmode=SELF_POLAR, self_nr=0, self_nt=0, want_class=False, self_tau=2.0):
    """geom: dict(norders, ixyzs, iptype, srccoefs (N,9), srcvals (N,12)).  The polar self
    rule uses self_nr x self_nt points on panels of at most self_tau."""
    _set_self_tau(self_tau)
    norders = np.asarray(geom["norders"], np.int64)
    ixyzs = np.asarray(geom["ixyzs"], np.int64)
    iptype = np.asarray(geom["iptype"], np.int64)
    targs = _f(targs)
    nt, ndt = targs.shape
    nnz = len(col_ind)
    nquad = int(iquad[-1]) - 1
    real = kind in (LAP_COMB, LAP_SLP, LAP_DLP)
    wnear = np.zeros(nquad if real else 2 * nquad)
    cls = np.zeros(nnz, np.int64); nref = np.zeros(nnz, np.int32)
    lib().orc_getnearquad(len(norders), norders, ixyzs, iptype, int(ixyzs[-1]) - 1,
                          _f(geom["srccoefs"]), _f(geom["srcvals"]), ndt, nt, targs,
                          np.asarray(ipatch_id, np.int64), _f(uvs_targ), eps, kind,
                          _kp(kind, zpars), nnz, np.asarray(row_ptr, np.int64),
                          np.asarray(col_ind, np.int64), np.asarray(iquad, np.int64), rfac0,
                          nquad, wnear, selfmode, self_nr, self_nt, cls.ctypes.data,
                          nref.ctypes.data)
    w = wnear if real else wnear.view(np.complex128)
    if want_class:
        return w, cls, nref
    return w


def oversample_fun(nd, norders, ixyzs, u, novers, ixyzso):
    ixyzso = np.asarray(ixyzso, np.int64)
    out = np.zeros((int(ixyzso[-1]) - 1) * nd)
    lib().orc_oversample_fun(nd, len(norders), np.asarray(norders, np.int64),
                             np.asarray(ixyzs, np.int64), _f(u), np.asarray(novers, np.int64),
                             ixyzso, out)
    return out


def oversample_geom(geom, novers, ixyzso):
    ixyzso = np.asarray(ixyzso, np.int64)
    nso = int(ixyzso[-1]) - 1
    srcover = np.zeros((nso, 12)); wover = np.zeros(nso)
    lib().orc_oversample_geom(len(geom["norders"]), np.asarray(geom["norders"], np.int64),
                              np.asarray(geom["ixyzs"], np.int64), _f(geom["srccoefs"]),
                              _f(geom["srcvals"]), np.asarray(novers, np.int64), ixyzso,
                              srcover, wover)
    return srcover, wover


def lpcomp_addsub(cplx, geom, targs, zpars, row_ptr, col_ind, iquad, wnear, novers, ixyzso,
                  srcover, whtsover, sigma):
    targs = _f(targs)
    nt, ndt = targs.shape
    norders = np.asarray(geom["norders"], np.int64)
    ixyzs = np.asarray(geom["ixyzs"], np.int64)
    w = _c2d(wnear) if cplx else _f(wnear)
    s = _c2d(sigma) if cplx else _f(sigma)
    zp = _c2d(zpars) if cplx else _f(zpars)
    pot = np.zeros(nt * (2 if cplx else 1))
    lib().orc_lpcomp_addsub(cplx, len(norders), norders, ixyzs, int(ixyzs[-1]) - 1,
                            _f(geom["srcvals"]), ndt, nt, targs, zp, len(col_ind),
                            np.asarray(row_ptr, np.int64), np.asarray(col_ind, np.int64),
                            np.asarray(iquad, np.int64), w, np.asarray(novers, np.int64),
                            len(whtsover), np.asarray(ixyzso, np.int64), _f(srcover),
                            _f(whtsover), s, pot, ndt)
    return pot.view(np.complex128) if cplx else pot


def gmres(cplx, geom, wts, zpars, row_ptr, col_ind, iquad, wnear, novers, ixyzso, srcover,
          whtsover, zid, rhs, numit, eps_gmres):
    norders = np.asarray(geom["norders"], np.int64)
    ixyzs = np.asarray(geom["ixyzs"], np.int64)
    N = int(ixyzs[-1]) - 1
    niter = C.c_int(); rres = C.c_double()
    errs = np.zeros(numit + 1)
    soln = np.zeros(N * (2 if cplx else 1))
    lib().orc_gmres(cplx, len(norders), norders, ixyzs, N, _f(geom["srcvals"]), _f(wts),
                    _c2d(zpars) if cplx else _f(zpars), len(col_ind),
                    np.asarray(row_ptr, np.int64), np.asarray(col_ind, np.int64),
                    np.asarray(iquad, np.int64), _c2d(wnear) if cplx else _f(wnear),
                    np.asarray(novers, np.int64), len(whtsover), np.asarray(ixyzso, np.int64),
                    _f(srcover), _f(whtsover), _c2d([zid]) if cplx else _f([np.real(zid)]),
                    _c2d(rhs) if cplx else _f(rhs), numit, eps_gmres, C.byref(niter), errs,
                    C.byref(rres), soln)
    s = soln.view(np.complex128) if cplx else soln
    return s, niter.value, errs[: niter.value], rres.value


def findnear(cms, rads, targs):
    targs = _f(targs)
    nt, ndt = targs.shape
    row_ptr = np.zeros(nt + 1, np.int64)
    nnz = C.c_int64()
    lib().orc_findnear(_f(cms), len(rads), _f(rads), ndt, targs, nt, row_ptr, None, 0, C.byref(nnz))
    col = np.zeros(max(nnz.value, 1), np.int64)
    lib().orc_findnear(_f(cms), len(rads), _f(rads), ndt, targs, nt, row_ptr, col.ctypes.data,
                       nnz.value, C.byref(nnz))
    return row_ptr, col[: nnz.value]


def dlaran(iseed):
    s = np.asarray(iseed, np.int32).copy()
    v = lib().orc_dlaran(s)
    return v, s


def hkrand(seed):
    return lib().orc_hkrand(seed)


def far_order(eps, geom, cms, rads, targs, ikerorder, zk, row_ptr, col_ind, rfac):
    norders = np.asarray(geom["norders"], np.int64)
    npat = len(norders)
    targs = _f(targs)
    nt, ndt = targs.shape
    novers = np.zeros(npat, np.int64); ixyzso = np.zeros(npat + 1, np.int64)
    lib().orc_far_order(eps, npat, norders, np.asarray(geom["ixyzs"], np.int64), _f(cms),
                        _f(rads), _f(geom["srccoefs"]), ndt, nt, targs, ikerorder,
                        _c2d([zk]), len(col_ind), np.asarray(row_ptr, np.int64),
                        np.asarray(col_ind, np.int64), rfac, novers, ixyzso)
    return novers, ixyzso


def set_num_threads(n):
    """OpenMP threads of the oracle's loops (effective even after the runtime started)."""
    lib().orc_set_num_threads(int(n))
    return int(lib().orc_max_threads())


def set_threads(n):
    os.environ["OMP_NUM_THREADS"] = str(n)
