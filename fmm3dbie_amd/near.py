"""Near-field bookkeeping on the host: near lists, quadrature offsets, radii, far order.

* :func:`get_rfacs` -- near_field_routs.f:826-873.
* :func:`get_centroid_rads` -- surf_routs.f90:195-302 (vertex images).
* :func:`findnear` -- CSR of targets within rfac*rad of each patch centroid; the same
  set as ``findnearslow`` (near_field_routs.f:38-125), columns ascending per row.
* :func:`get_iquad_rsc` -- near_field_routs.f:765-820.
* :func:`get_far_order` -- far_field_routs.f90:10-539 (oversampling order per patch),
  vectorised over patches, with the reference's pseudo-random probes (hkrand/dlaran).
* :func:`get_patch_id_uvs`, :func:`get_qwts`, :func:`oversample_geom`.
"""
import numpy as np
from scipy.spatial import cKDTree

from . import basis


def get_rfacs(norder, iptype=1):
    rfac = 1.25
    if norder <= 2:
        rfac = 2.75
    if 2 < norder <= 6:
        rfac = 2.0
    if norder > 6:
        rfac = 1.25
    return rfac, 1.25


def norder_avg(geom):
    return int(np.floor(np.sum(geom["norders"]) / (len(geom["norders"]) + 0.0)))


def get_centroid_rads(geom):
    norders = geom["norders"]
    ixyzs = geom["ixyzs"]
    npat = len(norders)
    cms = np.zeros((npat, 3))
    rads = np.zeros(npat)
    sc = geom["srccoefs"]
    for n in np.unique(norders):
        idx = np.nonzero(norders == n)[0]
        P = np.stack([basis.koorn_pols(u, v, int(n)) for (u, v) in [(0, 0), (1, 0), (0, 1)]])  # (3, p)
        p = P.shape[1]
        starts = ixyzs[idx] - 1
        coefs = sc[starts[:, None] + np.arange(p)[None, :], :3]  # (m, p, 3)
        xyz = np.einsum("vl,mld->mvd", P, coefs)  # (m, 3 vertices, 3)
        c = xyz.sum(1) / 3
        r2 = ((xyz - c[:, None, :]) ** 2).sum(-1)
        cms[idx] = c
        rads[idx] = np.sqrt(r2.max(1))
    return cms, rads


def findnear(cms, rads, targs):
    """row_ptr (nt+1), col_ind (nnz), 1-based; target i is near patch j iff
    |x_i - c_j|^2 <= rads_j^2."""
    targs = np.ascontiguousarray(targs[:, :3])
    tree = cKDTree(targs)
    rmax = float(rads.max())
    lists = tree.query_ball_point(cms, rmax * (1 + 1e-12))
    rows, cols = [], []
    for j, lst in enumerate(lists):
        if not lst:
            continue
        lst = np.asarray(lst)
        d2 = ((targs[lst] - cms[j]) ** 2).sum(1)
        ok = lst[d2 <= rads[j] ** 2]
        rows.append(ok)
        cols.append(np.full(ok.size, j, np.int64))
    if rows:
        r = np.concatenate(rows)
        c = np.concatenate(cols)
    else:
        r = np.zeros(0, np.int64)
        c = np.zeros(0, np.int64)
    order = np.lexsort((c, r))
    r = r[order]
    c = c[order]
    nt = len(targs)
    row_ptr = np.zeros(nt + 1, np.int64)
    np.add.at(row_ptr, r + 1, 1)
    row_ptr = np.cumsum(row_ptr) + 1
    return row_ptr.astype(np.int64), (c + 1).astype(np.int64)


def findnear_gpu(cms, rads, targs):
    """findnear on the GPU through the findnearmem_/findnear_ drop-ins (same result as
    findnear: 1-based CSR, ascending within a row)."""
    import ctypes as C
    from ._lib import lib
    xyz = np.ascontiguousarray(cms, np.float64)
    r = np.ascontiguousarray(rads, np.float64)
    t = np.ascontiguousarray(targs, np.float64)
    ns, nt, ndt = len(xyz), len(t), t.shape[1] if t.ndim == 2 else 3
    i8 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    nnz = C.c_int64(0)
    L = lib()
    L.findnearmem_(p(xyz), i8(ns), p(r), i8(ndt), p(t), i8(nt), C.byref(nnz))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("findnearmem failed (status %d)" % L.fmm3dbie_last_error())
    row_ptr = np.zeros(nt + 1, np.int64)
    col_ind = np.zeros(max(nnz.value, 1), np.int64)
    L.findnear_(p(xyz), i8(ns), p(r), i8(ndt), p(t), i8(nt), p(row_ptr), p(col_ind))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("findnear failed (status %d)" % L.fmm3dbie_last_error())
    return row_ptr, col_ind[:nnz.value]


def get_far_order_gpu(eps, geom, cms, rads, targs, ikerorder, zk, row_ptr, col_ind, rfac):
    """get_far_order on the GPU through the get_far_order_ drop-in (same result as
    get_far_order); returns (novers, ixyzso)."""
    import ctypes as C
    from ._lib import lib
    npat = len(geom["norders"])
    t = np.ascontiguousarray(np.asarray(targs)[:, :3], np.float64)
    arr = lambda a, dt: np.ascontiguousarray(a, dt)  # noqa: E731
    keep = [arr(geom["norders"], np.int64), arr(geom["ixyzs"], np.int64), arr(geom.get("iptype", np.ones(npat)),
                                                                          np.int64),
            arr(cms, np.float64), arr(rads, np.float64), arr(geom["srccoefs"], np.float64), t,
            arr(row_ptr, np.int64), arr(col_ind, np.int64), np.array([complex(zk).real, complex(zk).imag])]
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    i8 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    novers = np.zeros(npat, np.int64)
    ixyzso = np.zeros(npat + 1, np.int64)
    L = lib()
    L.get_far_order_(C.byref(C.c_double(eps)), i8(npat), p(keep[0]), p(keep[1]), p(keep[2]), p(keep[3]), p(keep[4]),
                     i8(len(geom["srccoefs"])), p(keep[5]), i8(3), i8(len(t)), p(keep[6]), i8(int(ikerorder)),
                     p(keep[9]), i8(len(keep[8])), p(keep[7]), p(keep[8]), C.byref(C.c_double(rfac)), p(novers),
                     p(ixyzso))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("get_far_order failed (status %d)" % L.fmm3dbie_last_error())
    return novers, ixyzso


def oversample_geom_gpu(geom, novers):
    """oversample_geom and the oversampled nodes' smooth weights on the GPU through the
    oversample_geom_ / get_qwts_ drop-ins (same result as oversample_geom); returns
    (srcover (nptso, 12), whtsover (nptso,), ixyzso (npatches + 1,))."""
    import ctypes as C
    from ._lib import lib
    npat = len(geom["norders"])
    novers = np.ascontiguousarray(novers, np.int64)
    ixyzso = np.ones(npat + 1, np.int64)
    ixyzso[1:] += np.cumsum([_npols(int(n)) for n in novers])
    nso = int(ixyzso[-1]) - 1
    arr = lambda a, dt: np.ascontiguousarray(a, dt)  # noqa: E731
    keep = [arr(geom["norders"], np.int64), arr(geom["ixyzs"], np.int64),
            arr(geom.get("iptype", np.ones(npat)), np.int64), arr(geom["srccoefs"], np.float64),
            arr(geom["srcvals"], np.float64)]
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    i8 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    srcover = np.zeros((nso, 12))
    wover = np.zeros(nso)
    L = lib()
    L.oversample_geom_(i8(npat), p(keep[0]), p(keep[1]), p(keep[2]), i8(len(keep[4])), p(keep[3]), p(keep[4]),
                       p(novers), p(ixyzso), i8(nso), p(srcover))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("oversample_geom failed (status %d)" % L.fmm3dbie_last_error())
    ones = np.ones(npat, np.int64)  # iptype of the oversampled patches (kept alive for the call)
    L.get_qwts_(i8(npat), p(novers), p(ixyzso), p(ones), i8(nso), p(srcover), p(wover))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("get_qwts failed (status %d)" % L.fmm3dbie_last_error())
    return srcover, wover, ixyzso


def get_qwts_gpu(geom):
    """get_qwts on the GPU through the get_qwts_ drop-in."""
    import ctypes as C
    from ._lib import lib
    npat = len(geom["norders"])
    arr = lambda a, dt: np.ascontiguousarray(a, dt)  # noqa: E731
    keep = [arr(geom["norders"], np.int64), arr(geom["ixyzs"], np.int64),
            arr(geom.get("iptype", np.ones(npat)), np.int64), arr(geom["srcvals"], np.float64)]
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    w = np.zeros(len(keep[3]))
    L = lib()
    L.get_qwts_(C.byref(C.c_int64(npat)), p(keep[0]), p(keep[1]), p(keep[2]), C.byref(C.c_int64(len(w))), p(keep[3]),
                p(w))
    if L.fmm3dbie_last_error() != 0:
        raise RuntimeError("get_qwts failed (status %d)" % L.fmm3dbie_last_error())
    return w


def get_iquad_rsc(geom, row_ptr, col_ind):
    npols = np.diff(geom["ixyzs"])
    iquad = np.ones(len(col_ind) + 1, np.int64)
    iquad[1:] += np.cumsum(npols[col_ind - 1])
    return iquad


def get_patch_id_uvs(geom):
    norders = geom["norders"]
    ixyzs = geom["ixyzs"]
    npts = int(ixyzs[-1]) - 1
    pid = np.zeros(npts, np.int64)
    uvs = np.zeros((npts, 2))
    for ip in range(len(norders)):
        a, b = ixyzs[ip] - 1, ixyzs[ip + 1] - 1
        pid[a:b] = ip + 1
    for n in np.unique(norders):
        u, _ = basis.rv_nodes(int(n))
        for ip in np.nonzero(norders == n)[0]:
            uvs[ixyzs[ip] - 1: ixyzs[ip + 1] - 1] = u
    return pid, uvs


def get_qwts(geom):
    """Smooth quadrature weights |x_u x x_v| * w_RV (surf_routs.f90:1598-1651)."""
    sv = geom["srcvals"]
    n = np.cross(sv[:, 3:6], sv[:, 6:9])
    jac = np.linalg.norm(n, axis=1)
    w = np.zeros(len(sv))
    for ip, no in enumerate(geom["norders"]):
        pass
    for no in np.unique(geom["norders"]):
        _, rw = basis.rv_nodes(int(no))
        for ip in np.nonzero(geom["norders"] == no)[0]:
            a, b = geom["ixyzs"][ip] - 1, geom["ixyzs"][ip + 1] - 1
            w[a:b] = rw
    return jac * w


def _npols(n):
    return (n + 1) * (n + 2) // 2


def oversample_geom(geom, novers):
    """srcover (nptso, 12), whtsover (nptso,), ixyzso (npatches+1,): geometry at RV nodes
    of order novers[i] on every patch (oversample_geom, surf_routs.f90:304-449, and
    get_qwts), centroid node shared when both orders are multiples of 3
    (get_clashing_indices, patch_basis_routs.f90:518-572)."""
    norders = geom["norders"]
    ixyzs = geom["ixyzs"]
    npat = len(norders)
    novers = np.asarray(novers, np.int64)
    ixyzso = np.ones(npat + 1, np.int64)
    ixyzso[1:] += np.cumsum([_npols(int(n)) for n in novers])
    nso = int(ixyzso[-1]) - 1
    srcover = np.zeros((nso, 12))
    wover = np.zeros(nso)
    sc = geom["srccoefs"]
    sv = geom["srcvals"]
    for (no, nv) in sorted(set(zip(norders.tolist(), novers.tolist()))):
        idx = np.nonzero((norders == no) & (novers == nv))[0]
        p, po = _npols(no), _npols(nv)
        uvo, wo = basis.rv_nodes(nv)
        if nv == no:
            for ip in idx:
                srcover[ixyzso[ip] - 1: ixyzso[ip + 1] - 1] = sv[ixyzs[ip] - 1: ixyzs[ip + 1] - 1]
        else:
            P = np.stack([basis.koorn_pols(u, v, no) for (u, v) in uvo])  # (po, p)
            starts = ixyzs[idx] - 1
            coefs = sc[starts[:, None] + np.arange(p)[None, :], :]  # (m, p, 9)
            vals = np.einsum("ql,mld->mqd", P, coefs)  # (m, po, 9)
            n = np.cross(vals[..., 3:6], vals[..., 6:9])
            n /= np.linalg.norm(n, axis=-1, keepdims=True)
            out = np.concatenate([vals, n], -1)
            if no % 3 == 0 and nv % 3 == 0:
                ic = {15: 75, 18: 105}.get(no, 0)
                icf = {15: 75, 18: 105}.get(nv, 0)
                out[:, icf, :] = sv[starts + ic]
            for k, ip in enumerate(idx):
                srcover[ixyzso[ip] - 1: ixyzso[ip + 1] - 1] = out[k]
        for ip in idx:
            a, b = ixyzso[ip] - 1, ixyzso[ip + 1] - 1
            nn = np.cross(srcover[a:b, 3:6], srcover[a:b, 6:9])
            wover[a:b] = np.linalg.norm(nn, axis=1) * wo
    return srcover, wover, ixyzso


# --- get_far_order ----------------------------------------------------------------

def _dlaran(iseed):
    """LAPACK dlaran: 48-bit multiplicative congruential generator (base 4096 digits)."""
    m1, m2, m3, m4 = 494, 322, 2508, 2549
    ipw2 = 4096
    r = 1.0 / ipw2
    s = list(iseed)
    while True:
        it4 = s[3] * m4
        it3 = it4 // ipw2
        it4 -= ipw2 * it3
        it3 += s[2] * m4 + s[3] * m3
        it2 = it3 // ipw2
        it3 -= ipw2 * it2
        it2 += s[1] * m4 + s[2] * m3 + s[3] * m2
        it1 = it2 // ipw2
        it2 -= ipw2 * it1
        it1 += s[0] * m4 + s[1] * m3 + s[2] * m2 + s[3] * m1
        it1 %= ipw2
        s = [it1, it2, it3, it4]
        out = r * (it1 + r * (it2 + r * (it3 + r * it4)))
        if out != 1.0:
            return out, s


def hkrand(seed):
    """hkrand with a nonzero seed (hkrand.f:1-19): reseed then one dlaran draw."""
    return _dlaran([0, 0, 0, (2 * seed + 1) % 4096])[0]


_NFARS = [1, 2, 3, 4, 6, 8, 10, 12, 14, 16]


def _far_order_chunk(eps, geom, idx, no, pr, ikerorder, zk):
    ixyzs = geom["ixyzs"]
    p = _npols(no)
    coefs = geom["srccoefs"][(ixyzs[idx] - 1)[:, None] + np.arange(p)[None, :], :]  # (m, p, 9)
    uv, w = basis.rv_nodes(no)
    P0 = np.stack([basis.koorn_pols(u, v, no) for (u, v) in uv])
    v0 = np.einsum("ql,mld->mqd", P0, coefs)
    j0 = np.linalg.norm(np.cross(v0[..., 3:6], v0[..., 6:9]), axis=-1)
    epsp = (j0 ** 2).min(1) ** 0.25 * eps

    def moments(nf):
        uvf, wf = basis.rv_nodes(nf)
        Pf = np.stack([basis.koorn_pols(u, v, no) for (u, v) in uvf])  # (q, p)
        vf = np.einsum("ql,mld->mqd", Pf, coefs)
        nrm = np.cross(vf[..., 3:6], vf[..., 6:9])
        jac = np.linalg.norm(nrm, axis=-1)
        nrm /= jac[..., None]
        qw = wf[None, :] * jac
        d = pr[:, :, None, :] - vf[:, None, :, :3]  # (m, t, q, 3)
        r = np.linalg.norm(d, axis=-1)
        if ikerorder == -1:
            val = np.exp(1j * zk * r) / r * 0.07957747154594767
        else:
            dd = (d * nrm[:, None, :, :]).sum(-1)
            val = dd * (1.0 - 1j * zk * r) * np.exp(1j * zk * r) / r ** 3 * 0.07957747154594767
        return np.einsum("mtq,ql,mq->mtl", val, Pf, qw)

    vcomp = moments(_NFARS[0])
    chosen = np.full(len(idx), 9)
    done = np.zeros(len(idx), bool)
    for nn in range(1, 10):
        vref = moments(_NFARS[nn])
        err = np.sqrt((np.abs(vref - vcomp) ** 2).sum(-1)).max(-1)
        hit = (~done) & (err < epsp)
        chosen[hit] = nn - 1
        done |= hit
        if done.all():
            break
        vcomp = vref
    return np.array(_NFARS)[chosen]


def get_far_order(eps, geom, cms, rads, targs, ikerorder, zk, row_ptr, col_ind, rfac):
    """novers (npatches,), ixyzso (npatches+1,) -- get_far_order(_guru)."""
    norders = geom["norders"]
    ixyzs = geom["ixyzs"]
    npat = len(norders)
    targs = np.ascontiguousarray(targs[:, :3])
    nt = len(targs)
    rows = np.repeat(np.arange(nt), np.diff(row_ptr))
    cols = col_ind - 1
    order = np.argsort(cols, kind="stable")
    rows_c = rows[order]
    col_ptr = np.zeros(npat + 1, np.int64)
    np.add.at(col_ptr, cols + 1, 1)
    col_ptr = np.cumsum(col_ptr)
    nmax = 25
    pi = np.arctan(1.0) * 4
    rnd = np.array([[hkrand(1001 + i + 1) * 2 * pi, hkrand(3042 + i + 1) * pi] for i in range(nmax)])
    probes = np.zeros((npat, nmax, 3))
    for ip in range(npat):
        a, b = col_ptr[ip], col_ptr[ip + 1]
        tl = rows_c[a:b]
        n0 = len(tl)
        i0 = min(10, n0 // 2)
        if i0:
            d = ((targs[tl] - cms[ip]) ** 2).sum(1)
            srt = tl[np.argsort(d, kind="stable")]
            probes[ip, :i0] = targs[srt[::-1][:i0]]
        rad = rfac * rads[ip]
        phi = rnd[i0:, 0]
        th = rnd[i0:, 1]
        probes[ip, i0:] = cms[ip] + rad * np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi),
                                                    np.cos(th)], 1)
    novers = np.zeros(npat, np.int64)
    for no in np.unique(norders):
        allidx = np.nonzero(norders == no)[0]
        for c0 in range(0, len(allidx), 256):
            idx = allidx[c0:c0 + 256]
            novers[idx] = _far_order_chunk(eps, geom, idx, int(no), probes[idx], ikerorder, zk)
    ixyzso = np.ones(npat + 1, np.int64)
    ixyzso[1:] += np.cumsum([_npols(int(n)) for n in novers])
    return novers, ixyzso
