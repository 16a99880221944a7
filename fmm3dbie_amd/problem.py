"""Set-up of a combined-field Dirichlet problem the way the reference's solver does it
(helm_comb_dir_solver, helm_comb_dir.f:1808-2065; lap_comb_dir_solver,
lap_comb_dir.f90:1189-1458): radii -> near lists -> quadrature offsets -> far order ->
oversampled geometry -> smooth weights.  Host-side numpy; the quadrature itself and the
apply run on the GPU through :mod:`fmm3dbie_amd.layer`."""
import numpy as np

from . import near


def setup(geom, eps, helm, pars, far_order=True, nover_fixed=None, gpu_near=False):
    """Returns a dict with row_ptr, col_ind, iquad (1-based), rfac, rfac0, cms, rads,
    novers, ixyzso, srcover, whtsover, wts, ipatch_id, uvs_targ, nnz, nquad.
    gpu_near: build the near list, the far-field orders, the oversampled geometry and the
    smooth weights on the GPU (same result)."""
    norder = near.norder_avg(geom)
    rfac, rfac0 = near.get_rfacs(norder)
    cms, rads = near.get_centroid_rads(geom)
    targs = geom["srcvals"][:, :3]
    finder = near.findnear_gpu if gpu_near else near.findnear
    row_ptr, col_ind = finder(cms, rads * rfac, targs)
    iquad = near.get_iquad_rsc(geom, row_ptr, col_ind)
    if helm:
        # the complex wavenumber zpars(1), as the reference passes it to get_far_order
        # (helm_comb_dir.f:596, 2010)
        zk = complex(np.asarray(pars, np.complex128)[0])
        beta = complex(np.asarray(pars, np.complex128)[2])
    else:
        zk = 0.0
        beta = float(pars[1])
    ikerorder = 0 if abs(beta) > 1e-16 else -1
    if nover_fixed is not None:
        novers = np.full(len(geom["norders"]), int(nover_fixed), np.int64)
    elif far_order:
        far = near.get_far_order_gpu if gpu_near else near.get_far_order
        novers, _ = far(eps, geom, cms, rads, targs, ikerorder, zk if helm else 0.0, row_ptr, col_ind, rfac)
    else:
        novers = np.asarray(geom["norders"], np.int64).copy()
    if gpu_near:
        srcover, whtsover, ixyzso = near.oversample_geom_gpu(geom, novers)
        wts = near.get_qwts_gpu(geom)
    else:
        srcover, whtsover, ixyzso = near.oversample_geom(geom, novers)
        wts = near.get_qwts(geom)
    pid, uvs = near.get_patch_id_uvs(geom)
    return dict(row_ptr=row_ptr, col_ind=col_ind, iquad=iquad, rfac=rfac, rfac0=rfac0, cms=cms, rads=rads,
                novers=novers, ixyzso=ixyzso, srcover=srcover, whtsover=whtsover, wts=wts,
                ipatch_id=pid, uvs_targ=uvs, nnz=len(col_ind), nquad=int(iquad[-1]) - 1)


def point_source(xyz, x0, zk=None):
    """e^{ik|x-x0|}/(4 pi |x-x0|) (Helmholtz) or 1/(4 pi |x-x0|) (Laplace)."""
    r = np.linalg.norm(xyz - np.asarray(x0)[None, :], axis=1)
    if zk is None:
        return 1.0 / (4 * np.pi * r)
    return np.exp(1j * zk * r) / (4 * np.pi * r)


def plane_wave_dirs(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def plane_waves(xyz, dirs, zk):
    """u_inc[r, i] = exp(i k d_r . x_i) for unit directions d_r (rows)."""
    return np.exp(1j * zk * (np.asarray(dirs) @ np.asarray(xyz).T))


def mie_sound_soft(targets, dirs, zk, a=1.0, nmax=None):
    """Field scattered by a sound-soft sphere of radius a centred at the origin for the
    plane waves exp(i k d . x): u_s(x) = -sum_n (2n+1) i^n j_n(ka)/h_n(ka) h_n(k|x|)
    P_n(d . x/|x|), h_n = j_n + i y_n (partial-wave expansion of the plane wave; the
    series converges for |x| >= a once n exceeds ka by a margin)."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn
    targets = np.asarray(targets, float)
    dirs = np.asarray(dirs, float)
    r = np.linalg.norm(targets, axis=1)
    ct = (dirs @ targets.T) / r[None, :]
    ka = zk * a
    if nmax is None:
        nmax = int(ka + 10 * max(1.0, ka) ** (1 / 3) + 20)
    u = np.zeros(ct.shape, complex)
    for n in range(nmax + 1):
        hna = spherical_jn(n, ka) + 1j * spherical_yn(n, ka)
        hnr = spherical_jn(n, zk * r) + 1j * spherical_yn(n, zk * r)
        u -= (2 * n + 1) * (1j ** n) * (spherical_jn(n, ka) / hna) * hnr[None, :] * eval_legendre(n, ct)
    return u
