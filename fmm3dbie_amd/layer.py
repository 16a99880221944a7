"""Python mirror of the reference's layer-potential interface.

Host drop-ins (numpy in, numpy out) call the reference-signature C-ABI symbols of
libfmm3dbie_amd.so, which run on the GPU:

* :func:`getnearquad_helm_comb_dir`, :func:`getnearquad_helm_comb_dir_eval`
  (src/helm_wrappers/helm_comb_dir.f:107-434);
* :func:`getnearquad_lap_comb_dir`, :func:`getnearquad_lap_comb_dir_eval`
  (src/lap_wrappers/lap_comb_dir.f90:86-406);
* :func:`lpcomp_helm_comb_dir_addsub`, :func:`lpcomp_lap_comb_dir_addsub`
  (helm_comb_dir.f:658-850, lap_comb_dir.f90:633-805);
* :func:`helm_comb_dir_solver_guru`, :func:`lap_comb_dir_solver_guru`;
* :func:`helm_comb_dir_solver`, :func:`lap_comb_dir_solver` (setup included,
  helm_comb_dir.f:1808-2064, lap_comb_dir.f90:1189-1458);
* :func:`helm_comb_dir_eval`, :func:`lap_comb_dir_eval` (potential at arbitrary targets,
  helm_comb_dir.f:440-652, lap_comb_dir.f90:408-632).

Device paths (torch tensors resident in HBM) are :class:`NearQuadDevice`,
:class:`ApplyPlan`, :func:`gmres` and :func:`gmres_ir`.

Errors: the reference prints and returns; here every entry point checks the library's
status and raises ``RuntimeError``.
"""
import ctypes as C

import numpy as np

from ._lib import lib
from . import near


# The self rule: panels of at most SELF_TAU in the sinh-graded edge parameter with
# SELF_NR x SELF_NT Gauss-Legendre points (radial x edge parameter).  It integrates the
# self-patch weights of the benchmark surfaces to the same 1e-8 (relative to the largest
# weight of the target) as 8 x 8 points on panels of 1 or 16 x 16 on panels of 0.5, with
# 24% fewer nodes than the former.
SELF_NR = 8
SELF_NT = 10
SELF_TAU = 2.0


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _i8(v):
    return C.byref(C.c_int64(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def _check(what):
    rc = lib().fmm3dbie_last_error()
    if rc != 0:
        raise RuntimeError("%s failed (status %d)" % (what, rc))


def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _z3(zpars):
    return _arr(np.asarray(zpars, np.complex128), np.complex128).view(np.float64)


# ------------------------------------------------------------------ host drop-ins

def getnearquad_helm_comb_dir_eval(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg, ntarg,
                                   targs, ipatch_id, uvs_targ, eps, zpars, iquadtype, nnz, row_ptr, col_ind,
                                   iquad, rfac0, nquad):
    """wnear (nquad,) complex128 for u = alpha S_k + beta D_k, zpars = (k, alpha, beta)."""
    wnear = np.zeros(int(nquad), np.complex128)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _arr(targs, np.float64), _arr(ipatch_id, np.int64),
            _arr(uvs_targ, np.float64), _z3(zpars), _arr(row_ptr, np.int64), _arr(col_ind, np.int64),
            _arr(iquad, np.int64)]
    f = lib().getnearquad_helm_comb_dir_eval_
    f(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]), _p(keep[4]), _i8(ndtarg),
      _i8(ntarg), _p(keep[5]), _p(keep[6]), _p(keep[7]), _d(eps), _p(keep[8]), _i8(iquadtype), _i8(nnz),
      _p(keep[9]), _p(keep[10]), _p(keep[11]), _d(rfac0), _i8(nquad), _p(wnear))
    _check("getnearquad_helm_comb_dir_eval")
    return wnear


def getnearquad_helm_comb_dir(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, zpars, iquadtype,
                              nnz, row_ptr, col_ind, iquad, rfac0, nquad):
    wnear = np.zeros(int(nquad), np.complex128)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _z3(zpars), _arr(row_ptr, np.int64), _arr(col_ind, np.int64),
            _arr(iquad, np.int64)]
    lib().getnearquad_helm_comb_dir_(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts),
                                     _p(keep[3]), _p(keep[4]), _d(eps), _p(keep[5]), _i8(iquadtype), _i8(nnz),
                                     _p(keep[6]), _p(keep[7]), _p(keep[8]), _d(rfac0), _i8(nquad), _p(wnear))
    _check("getnearquad_helm_comb_dir")
    return wnear


def getnearquad_lap_comb_dir_eval(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg, ntarg, targs,
                                  ipatch_id, uvs_targ, eps, dpars, iquadtype, nnz, row_ptr, col_ind, iquad, rfac0,
                                  nquad):
    """wnear (nquad,) float64 for u = alpha S + beta D, dpars = (alpha, beta)."""
    wnear = np.zeros(int(nquad), np.float64)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _arr(targs, np.float64), _arr(ipatch_id, np.int64),
            _arr(uvs_targ, np.float64), _arr(dpars, np.float64), _arr(row_ptr, np.int64),
            _arr(col_ind, np.int64), _arr(iquad, np.int64)]
    lib().getnearquad_lap_comb_dir_eval_(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts),
                                         _p(keep[3]), _p(keep[4]), _i8(ndtarg), _i8(ntarg), _p(keep[5]),
                                         _p(keep[6]), _p(keep[7]), _d(eps), _p(keep[8]), _i8(iquadtype), _i8(nnz),
                                         _p(keep[9]), _p(keep[10]), _p(keep[11]), _d(rfac0), _i8(nquad), _p(wnear))
    _check("getnearquad_lap_comb_dir_eval")
    return wnear


def getnearquad_lap_comb_dir(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, dpars, iquadtype, nnz,
                             row_ptr, col_ind, iquad, rfac0, nquad):
    wnear = np.zeros(int(nquad), np.float64)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _arr(dpars, np.float64), _arr(row_ptr, np.int64), _arr(col_ind, np.int64),
            _arr(iquad, np.int64)]
    lib().getnearquad_lap_comb_dir_(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]),
                                    _p(keep[4]), _d(eps), _p(keep[5]), _i8(iquadtype), _i8(nnz), _p(keep[6]),
                                    _p(keep[7]), _p(keep[8]), _d(rfac0), _i8(nquad), _p(wnear))
    _check("getnearquad_lap_comb_dir")
    return wnear


def _lpcomp(sym, cplx, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, pars, nnz, row_ptr, col_ind,
            iquad, nquad, wnear, novers, nptso, ixyzso, srcover, whtsover, sigma):
    pot = np.zeros(int(npts), np.complex128 if cplx else np.float64)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _z3(pars) if cplx else _arr(pars, np.float64), _arr(row_ptr, np.int64),
            _arr(col_ind, np.int64), _arr(iquad, np.int64),
            _arr(wnear, np.complex128 if cplx else np.float64), _arr(novers, np.int64), _arr(ixyzso, np.int64),
            _arr(srcover, np.float64), _arr(whtsover, np.float64),
            _arr(sigma, np.complex128 if cplx else np.float64)]
    dummy = np.zeros(4)
    zero = _i8(0)
    dp = keep[5] if not cplx else dummy
    zp = keep[5] if cplx else dummy
    getattr(lib(), sym)(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]), _p(keep[4]),
                        _d(eps), _i8(0 if cplx else 2), _p(dp), _i8(3 if cplx else 0), _p(zp), zero, _p(dummy),
                        _i8(nnz), _p(keep[6]), _p(keep[7]), _p(keep[8]), _i8(nquad), _i8(1), _p(keep[9]),
                        _p(keep[10]), _i8(nptso), _p(keep[11]), _p(keep[12]), _p(keep[13]), _i8(0), _p(dummy),
                        _i8(1), _p(keep[14]), _p(pot))
    _check(sym)
    return pot


def lpcomp_helm_comb_dir_addsub(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, zpars, nnz,
                                row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover, whtsover,
                                sigma):
    return _lpcomp("lpcomp_helm_comb_dir_addsub_", True, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals,
                   eps, zpars, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover,
                   whtsover, sigma)


def lpcomp_lap_comb_dir_addsub(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, dpars, nnz, row_ptr,
                               col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover, whtsover, sigma):
    return _lpcomp("lpcomp_lap_comb_dir_addsub_", False, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals,
                   eps, dpars, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover,
                   whtsover, sigma)


def _solver(sym, cplx, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, pars, numit, ifinout, rhs,
            nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover, wover, eps_gmres):
    soln = np.zeros(int(npts), np.complex128 if cplx else np.float64)
    errs = np.zeros(int(numit) + 1)
    niter = C.c_int64(0)
    rres = C.c_double(0)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _z3(pars) if cplx else _arr(pars, np.float64),
            _arr(rhs, np.complex128 if cplx else np.float64), _arr(row_ptr, np.int64), _arr(col_ind, np.int64),
            _arr(iquad, np.int64), _arr(wnear, np.complex128 if cplx else np.float64), _arr(novers, np.int64),
            _arr(ixyzso, np.int64), _arr(srcover, np.float64), _arr(wover, np.float64)]
    getattr(lib(), sym)(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]), _p(keep[4]),
                        _d(eps), _p(keep[5]), _i8(numit), _i8(ifinout), _p(keep[6]), _i8(nnz), _p(keep[7]),
                        _p(keep[8]), _p(keep[9]), _i8(nquad), _i8(1), _p(keep[10]), _p(keep[11]), _i8(nptso),
                        _p(keep[12]), _p(keep[13]), _p(keep[14]), _d(eps_gmres), C.byref(niter), _p(errs),
                        C.byref(rres), _p(soln))
    _check(sym)
    return niter.value, errs[: niter.value], rres.value, soln


def helm_comb_dir_solver_guru(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, zpars, numit, ifinout,
                              rhs, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover, wover,
                              eps_gmres):
    """(niter, errs, rres, soln) -- exterior problems use ifinout=1 (zid = beta/2)."""
    return _solver("helm_comb_dir_solver_guru_", True, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals,
                   eps, zpars, numit, ifinout, rhs, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso,
                   ixyzso, srcover, wover, eps_gmres)


def lap_comb_dir_solver_guru(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, dpars, numit, ifinout,
                             rhs, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso, ixyzso, srcover, wover,
                             eps_gmres):
    return _solver("lap_comb_dir_solver_guru_", False, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals,
                   eps, dpars, numit, ifinout, rhs, nnz, row_ptr, col_ind, iquad, nquad, wnear, novers, nptso,
                   ixyzso, srcover, wover, eps_gmres)


def _eval(sym, cplx, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg, ntarg, targs, ipatch_id,
          uvs_targ, eps, pars, sigma):
    pot = np.zeros(int(ntarg), np.complex128 if cplx else np.float64)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _arr(targs, np.float64), _arr(ipatch_id, np.int64),
            _arr(uvs_targ, np.float64), _z3(pars) if cplx else _arr(pars, np.float64),
            _arr(sigma, np.complex128 if cplx else np.float64)]
    getattr(lib(), sym)(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]), _p(keep[4]),
                        _i8(ndtarg), _i8(ntarg), _p(keep[5]), _p(keep[6]), _p(keep[7]), _d(eps), _p(keep[8]),
                        _p(keep[9]), _p(pot))
    _check(sym)
    return pot


def helm_comb_dir_eval(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg, ntarg, targs, ipatch_id,
                       uvs_targ, eps, zpars, sigma):
    """(alpha S_k + beta D_k)[sigma] at the targets (helm_comb_dir.f:440-652); off-surface
    targets have ipatch_id = -1."""
    return _eval("helm_comb_dir_eval_", True, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg,
                 ntarg, targs, ipatch_id, uvs_targ, eps, zpars, sigma)


def lap_comb_dir_eval(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg, ntarg, targs, ipatch_id,
                      uvs_targ, eps, dpars, sigma):
    return _eval("lap_comb_dir_eval_", False, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, ndtarg,
                 ntarg, targs, ipatch_id, uvs_targ, eps, dpars, sigma)


def _solver_full(sym, cplx, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, pars, numit, ifinout,
                 rhs, eps_gmres):
    soln = np.zeros(int(npts), np.complex128 if cplx else np.float64)
    errs = np.zeros(int(numit) + 1)
    niter = C.c_int64(0)
    rres = C.c_double(0)
    keep = [_arr(norders, np.int64), _arr(ixyzs, np.int64), _arr(iptype, np.int64), _arr(srccoefs, np.float64),
            _arr(srcvals, np.float64), _z3(pars) if cplx else _arr(pars, np.float64),
            _arr(rhs, np.complex128 if cplx else np.float64)]
    getattr(lib(), sym)(_i8(npatches), _p(keep[0]), _p(keep[1]), _p(keep[2]), _i8(npts), _p(keep[3]), _p(keep[4]),
                        _d(eps), _p(keep[5]), _i8(numit), _i8(ifinout), _p(keep[6]), _d(eps_gmres), C.byref(niter),
                        _p(errs), C.byref(rres), _p(soln))
    _check(sym)
    return int(niter.value), errs[:niter.value], float(rres.value), soln


def helm_comb_dir_solver(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, zpars, numit, ifinout, rhs,
                         eps_gmres):
    """(niter, errs, rres, soln): setup (near list, far orders, quadrature) and GMRES on the
    GPU (helm_comb_dir.f:1808-2064)."""
    return _solver_full("helm_comb_dir_solver_", True, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps,
                        zpars, numit, ifinout, rhs, eps_gmres)


def lap_comb_dir_solver(npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps, dpars, numit, ifinout, rhs,
                        eps_gmres):
    return _solver_full("lap_comb_dir_solver_", False, npatches, norders, ixyzs, iptype, npts, srccoefs, srcvals, eps,
                        dpars, numit, ifinout, rhs, eps_gmres)


# ------------------------------------------------------------------ device paths

def _torch():
    import torch
    return torch


def _dptr(t):
    return C.c_void_p(t.data_ptr())


def kpar6(helm, pars):
    if helm:
        z = np.asarray(pars, np.complex128)
        return np.array([z[0].real, z[0].imag, z[1].real, z[1].imag, z[2].real, z[2].imag])
    a, b = pars
    return np.array([0.0, 0.0, float(a), 0.0, float(b), 0.0])


# the adaptive and self kernels hold one complex moment component per lane of a 64-wide
# wavefront: (p + 1)(p + 2) / 2 <= 64
MAX_QUAD_ORDER = 9


def set_side_stream(enable=True):
    """Near-quadrature concurrency: True (default) runs the fixed-rule and self kernels on
    a second stream beside the adaptive kernels; False keeps every kernel in order on the
    caller's stream, so that per-kernel times are those of the kernel alone."""
    global _SIDE
    _SIDE = bool(enable)
    lib().fmm3dbie_set_side_stream(int(bool(enable)))


_SIDE = True


def set_gmres_print(every=1):
    """GMRES progress on stderr: the iteration and residual estimate every `every`
    iterations (as the reference's zgmres_guru prints each iteration), and a line per
    refinement of gmres_ir; 0 turns it off (the default)."""
    lib().fmm3dbie_set_gmres_print(int(every))


def side_stream_enabled():
    """The current side-stream setting (set_side_stream; on by default)."""
    return _SIDE


class NearQuadDevice:
    """Near quadrature with device-resident inputs (the benchmarked path).

    geom: surface dict; targets default to the on-surface nodes (getnearquad_*_comb_dir).
    """

    def __init__(self, geom, helm, pars, eps, row_ptr, col_ind, iquad, rfac0=1.25, device="cuda", targets=None,
                 ipatch_id=None, uvs_targ=None):
        torch = _torch()
        self.helm = bool(helm)
        self.kp = kpar6(helm, pars)
        self.eps = float(eps)
        self.rfac0 = float(rfac0)
        norders = np.asarray(geom["norders"], np.int64)
        if not np.all(norders == norders[0]):
            raise ValueError("patches of one order only")
        self.norder = int(norders[0])
        if not 1 <= self.norder <= MAX_QUAD_ORDER:
            raise ValueError("patch order %d not supported by the near quadrature (1..%d)"
                             % (self.norder, MAX_QUAD_ORDER))
        self.norder_branch = near.norder_avg(geom)
        self.npatches = len(norders)
        if targets is None:
            targets = geom["srcvals"]
            ipatch_id, uvs_targ = near.get_patch_id_uvs(geom)
        self.ntarg, self.ndtarg = targets.shape
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device=device)  # noqa: E731
        self.d_ixyzs = t(geom["ixyzs"], np.int64)
        self.d_srccoefs = t(geom["srccoefs"], np.float64)
        self.d_srcvals = t(geom["srcvals"], np.float64)
        self.d_targs = t(targets, np.float64)
        self.d_pid = t(ipatch_id, np.int64) if ipatch_id is not None else None
        self.d_uvs = t(uvs_targ, np.float64) if uvs_targ is not None else None
        self.d_row = t(row_ptr, np.int64)
        self.d_col = t(col_ind, np.int64)
        self.d_iquad = t(iquad, np.int64)
        self.nnz = len(col_ind)
        self.nquad = int(iquad[-1]) - 1
        self.wnear = torch.zeros(self.nquad, dtype=torch.complex128 if self.helm else torch.float64, device=device)
        self.stats = None

    def run(self, stream=None, want_detail=False, selector=None, features=None, route_lmax=None):
        """Fill self.wnear.  selector: device float32 tensor of MLP parameters (learned
        order selection: every adaptive pair integrated with the rule class the network
        predicts), or None for the plain adaptive path.  route_lmax: the deepest level of
        each of the 6 rule classes (a pair whose tree needs more is integrated again by
        the plain integrator); default 20.  features: device float32 tensor (nnz, 8) that
        receives the selector features of the adaptive pairs.  want_detail adds per-pair
        class, refinement-test count, deepest level, tests below the cached tree and
        (with a selector) predicted rule class to self.stats."""
        torch = _torch()
        st = np.zeros(24)
        nref = np.zeros(self.nnz, np.int32) if want_detail else None
        cls = np.zeros(self.nnz, np.int32) if want_detail else None
        lev = np.zeros(self.nnz, np.int32) if want_detail else None
        unc = np.zeros(self.nnz, np.int32) if want_detail else None
        pred = np.zeros(self.nnz, np.int32) if (want_detail and selector is not None) else None
        if selector is not None and (selector.dtype != torch.float32 or not selector.is_cuda
                                     or selector.numel() != lib().fmm3dbie_mlp_nparam()):
            raise ValueError("selector must be a device float32 tensor of %d parameters"
                             % lib().fmm3dbie_mlp_nparam())
        if features is not None and (features.dtype != torch.float32 or not features.is_cuda
                                     or tuple(features.shape) != (self.nnz, 8) or not features.is_contiguous()):
            raise ValueError("features must be a contiguous device float32 tensor of shape (nnz, 8)")
        lm = None
        if route_lmax is not None:
            lm = np.ascontiguousarray(route_lmax, dtype=np.int32)
            if lm.shape != (6,):
                raise ValueError("route_lmax must hold 6 levels")
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().fmm3dbie_nearquad_dev(
            int(self.helm), self.norder, self.norder_branch, self.npatches, _dptr(self.d_ixyzs),
            _dptr(self.d_srccoefs), _dptr(self.d_srcvals), self.ndtarg, self.ntarg, _dptr(self.d_targs),
            _dptr(self.d_pid) if self.d_pid is not None else None,
            _dptr(self.d_uvs) if self.d_uvs is not None else None, self.eps, _p(self.kp), self.nnz,
            _dptr(self.d_row), _dptr(self.d_col), _dptr(self.d_iquad), self.rfac0, _dptr(self.wnear),
            SELF_NR, SELF_NT, SELF_TAU,
            C.c_void_p(s.cuda_stream), _p(st), _p(nref) if want_detail else None, _p(cls) if want_detail else None,
            _dptr(selector) if selector is not None else None, _dptr(features) if features is not None else None,
            _p(pred) if pred is not None else None, _p(lm) if lm is not None else None,
            _p(lev) if want_detail else None, _p(unc) if want_detail else None)
        if rc != 0:
            raise RuntimeError("fmm3dbie_nearquad_dev failed (status %d)" % rc)
        self.stats = dict(nself=int(st[0]), nadap=int(st[1]), nfixed=int(st[2]), overflow=int(st[3]), nf=int(st[4]),
                          kq=int(st[5]), ms_prep=st[6], ms_fixed=st[7], ms_adap=st[8], ms_self=st[9],
                          ms_tables=st[10], ms_select=st[11], ms_tree=st[12], ms_tree_routed=st[13],
                          nroute=int(st[14]), nfallback=int(st[16]), ncls=[int(v) for v in st[17:23]],
                          cache_tris=int(st[23]))
        if want_detail:
            self.stats["nref"] = nref
            self.stats["cls"] = cls
            self.stats["lev"] = lev
            self.stats["unc"] = unc
            if pred is not None:
                self.stats["pred"] = pred
        return self.wnear


def _dens(t, npts, prec, what):
    """Checks a density / right-hand-side tensor handed to the kernels: on the GPU,
    contiguous, complex128 (prec 64) or complex64 (prec 32), shape (npts,) or
    (nrhs, npts) with 1 <= nrhs <= 64.  Returns nrhs."""
    torch = _torch()
    if prec not in (32, 64):
        raise ValueError("prec must be 32 or 64, got %r" % (prec,))
    want = torch.complex128 if prec == 64 else torch.complex64
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError("%s must be a GPU tensor" % what)
    if t.dtype != want:
        raise ValueError("%s must be %s for prec %d, got %s" % (what, want, prec, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % what)
    if t.dim() == 1 and t.shape[0] == npts:
        return 1
    if t.dim() == 2 and t.shape[1] == npts and 1 <= t.shape[0] <= 64:
        return int(t.shape[0])
    raise ValueError("%s must have shape (%d,) or (nrhs, %d) with nrhs <= 64, got %s"
                     % (what, npts, npts, tuple(t.shape)))


def _weights(w, npts):
    torch = _torch()
    if (not isinstance(w, torch.Tensor) or not w.is_cuda or w.dtype != torch.float64 or not w.is_contiguous()
            or tuple(w.shape) != (npts,)):
        raise ValueError("wts must be a contiguous GPU float64 tensor of shape (%d,)" % npts)


class ApplyPlan:
    """Device-resident apply (lpcomp_*_addsub) reused across matvecs and RHS."""

    def __init__(self, geom, helm, pars, row_ptr, col_ind, iquad, wnear, novers, ixyzso, srcover, whtsover,
                 stream=None):
        torch = _torch()
        self.helm = bool(helm)
        self.kp = kpar6(helm, pars)
        self.npts = int(geom["ixyzs"][-1]) - 1
        keep = [_arr(geom["norders"], np.int64), _arr(geom["ixyzs"], np.int64), _arr(geom["srcvals"], np.float64),
                _arr(row_ptr, np.int64), _arr(col_ind, np.int64), _arr(iquad, np.int64), _arr(novers, np.int64),
                _arr(ixyzso, np.int64), _arr(srcover, np.float64), _arr(whtsover, np.float64)]
        if isinstance(wnear, np.ndarray):
            wk = _arr(wnear, np.complex128 if self.helm else np.float64)
            wptr, ondev = _p(wk), 0
        else:
            wk = wnear.contiguous()
            wptr, ondev = _dptr(wk), 1
        s = stream if stream is not None else torch.cuda.current_stream()
        self.handle = lib().fmm3dbie_apply_plan_create(
            int(self.helm), len(keep[0]), _p(keep[0]), _p(keep[1]), self.npts, _p(keep[2]), _p(self.kp),
            len(keep[4]), _p(keep[3]), _p(keep[4]), _p(keep[5]), int(iquad[-1]) - 1, wptr, ondev, _p(keep[6]),
            len(whtsover), _p(keep[7]), _p(keep[8]), _p(keep[9]), C.c_void_p(s.cuda_stream))
        if not self.handle:
            raise RuntimeError("apply plan creation failed (status %d)" % lib().fmm3dbie_last_error())

    def apply(self, sigma, prec=64, mode=0, stream=None, times=None):
        """sigma: complex tensor (npts,) or (nrhs, npts) on the GPU, complex128 for
        prec 64 and complex64 for prec 32; returns pot of the same shape."""
        torch = _torch()
        if mode not in (0, 1):
            raise ValueError("mode must be 0 (add/subtract) or 1 (masked)")
        nrhs = _dens(sigma, self.npts, prec, "sigma")
        sig = sigma
        pot = torch.empty_like(sig)
        tm = np.zeros(4)
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().fmm3dbie_apply_plan_apply(self.handle, prec, mode, _dptr(sig), _dptr(pot), nrhs,
                                             C.c_void_p(s.cuda_stream), _p(tm) if times is not None else None)
        if rc != 0:
            raise RuntimeError("apply failed (status %d)" % rc)
        if times is not None:
            times[:] = tm
        return pot

    def apply_noalloc(self, sigma, pot, prec=64, mode=0, stream=None):
        """The apply with no device allocation (fmm3dbie_apply_plan_apply_noalloc): work
        buffers in the library's fixed per-device scratch array; pot must be preallocated
        like sigma.  Raises RuntimeError when the work buffers exceed that array."""
        torch = _torch()
        if mode not in (0, 1):
            raise ValueError("mode must be 0 (add/subtract) or 1 (masked)")
        nrhs = _dens(sigma, self.npts, prec, "sigma")
        if _dens(pot, self.npts, prec, "pot") != nrhs:
            raise ValueError("pot must have the shape of sigma")
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().fmm3dbie_apply_plan_apply_noalloc(self.handle, prec, mode, _dptr(sigma), _dptr(pot), nrhs,
                                                     C.c_void_p(s.cuda_stream))
        if rc != 0:
            raise RuntimeError("allocation-free apply failed (status %d)" % rc)
        return pot

    def info(self):
        """(npts, nptso, nnz, uniform patch size or 0)."""
        out = np.zeros(4, np.int64)
        lib().fmm3dbie_apply_plan_info(self.handle, _p(out))
        return tuple(int(x) for x in out)

    def tune(self, prec=32, nrhs=1, stream=None):
        """Autotune the far-sum launch for (prec, nrhs); returns [(group, wg_target, ms)]."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        rep = np.zeros(3 * 32)
        n = lib().fmm3dbie_apply_plan_tune(self.handle, int(prec), int(nrhs), C.c_void_p(s.cuda_stream), _p(rep),
                                           len(rep))
        if n < 0:
            raise RuntimeError("apply plan tuning failed (status %d)" % n)
        return [(int(rep[3 * k]), int(rep[3 * k + 1]), float(rep[3 * k + 2])) for k in range(n)]

    def close(self):
        h = getattr(self, "handle", None)
        if h:
            lib().fmm3dbie_apply_plan_destroy(h)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


def gmres(plan, rhs, zid, wts, numit=200, eps_gmres=1e-7, prec=64, mode=0, stream=None):
    """GMRES for (zid I + K) x = rhs on the device.  rhs (npts,) or (nrhs, npts) complex.
    Returns (x, niter[nrhs], errs[nrhs, numit], rres[nrhs], applies)."""
    torch = _torch()
    nrhs = _dens(rhs, plan.npts, prec, "rhs")
    _weights(wts, plan.npts)
    if mode not in (0, 1) or int(numit) < 1:
        raise ValueError("mode must be 0 or 1 and numit >= 1")
    r = rhs
    x = torch.empty_like(r)
    niter = np.zeros(nrhs, np.int32)
    errs = np.zeros((nrhs, numit))
    rres = np.zeros(nrhs)
    st = np.zeros(2, np.int32)
    z = np.array([complex(zid).real, complex(zid).imag])
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().fmm3dbie_gmres_dev(plan.handle, prec, mode, _p(z), _dptr(r), _dptr(x), _dptr(wts), nrhs, numit,
                                  eps_gmres, _p(niter), _p(errs), _p(rres), C.c_void_p(s.cuda_stream), _p(st))
    if rc != 0:
        raise RuntimeError("gmres failed (status %d)" % rc)
    return x, niter, errs, rres, int(st[0])


def gmres_ir(plan, rhs, zid, wts, numit=200, eps_inner=1e-4, tol=1e-6, max_outer=10, mode=0, stream=None):
    """Mixed precision: FP32 inner GMRES + FP64 residual refinement.
    Returns (x, outer, inner_total, rres[nrhs], fp32 applies, fp64 applies)."""
    torch = _torch()
    nrhs = _dens(rhs, plan.npts, 64, "rhs")
    _weights(wts, plan.npts)
    if mode not in (0, 1) or int(numit) < 1 or int(max_outer) < 1:
        raise ValueError("mode must be 0 or 1, numit >= 1 and max_outer >= 1")
    r = rhs
    x = torch.empty_like(r)
    outer = C.c_int(0)
    inner = C.c_int(0)
    rres = np.zeros(nrhs)
    st = np.zeros(2, np.int32)
    z = np.array([complex(zid).real, complex(zid).imag])
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().fmm3dbie_gmres_ir_dev(plan.handle, mode, _p(z), _dptr(r), _dptr(x), _dptr(wts), nrhs, numit,
                                     eps_inner, tol, max_outer, C.byref(outer), C.byref(inner), _p(rres),
                                     C.c_void_p(s.cuda_stream), _p(st))
    if rc != 0:
        raise RuntimeError("gmres_ir failed (status %d)" % rc)
    return x, outer.value, inner.value, rres, int(st[0]), int(st[1])
