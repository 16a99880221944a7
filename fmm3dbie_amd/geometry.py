"""Surface generators and the surface data layout of the reference.

A surface is a dict with the reference's arrays (Fortran column-major arrays are stored
as C-contiguous numpy arrays with the leading Fortran dimension last):

* ``norders`` (npatches,) int64, ``ixyzs`` (npatches+1,) int64 1-based, ``iptype``;
* ``srcvals`` (npts, 12): xyz, dxyz/du, dxyz/dv, unit normal at the RV nodes;
* ``srccoefs`` (npts, 9): Koornwinder coefficients of xyz, dxyz/du, dxyz/dv per patch.

Generators:

* :func:`get_sphere_npat` / :func:`get_ellipsoid_npat` -- cube skeleton projected
  radially (``get_ellipsoid_npat``, src/surface_routs/analytic_geometry_routs.f90:150-386,
  ``xtri_get_rectparapiped``/``xtri_rectmesh_3d``/``xtri_ellipsoid_eval``,
  src/surface_routs/xtri_parameterizations.f90:247-339,1020-1217) followed by
  ``getgeominfo`` (src/surface_routs/surf_routs.f90:3200-3279);
* :func:`get_deformed_sphere` -- the same sphere scaled radially by a smooth function;
* :func:`get_wavy_torus` -- a torus with a tube radius that varies along the surface;
* :func:`get_platonic_sphere` -- the refined-cube sphere of the reference's
  ``geometries/sphere_*_o03.go3`` fixtures (``xtri_platonic``, xtri_parameterizations.f90).
"""
import numpy as np

from . import basis


def _finish(skel_eval, npatches, norder):
    uvs, _ = basis.rv_nodes(norder)
    umat = basis.rv_umat(norder)
    p = len(uvs)
    srcvals = np.zeros((npatches, p, 12))
    xyz, du, dv = skel_eval(uvs[:, 0], uvs[:, 1])  # (npatches, p, 3) each
    srcvals[:, :, 0:3] = xyz
    srcvals[:, :, 3:6] = du
    srcvals[:, :, 6:9] = dv
    n = np.cross(du, dv)
    srcvals[:, :, 9:12] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    srccoefs = np.einsum("il,pld->pid", umat, srcvals[:, :, :9])
    return {
        "norders": np.full(npatches, norder, np.int64),
        "ixyzs": np.arange(npatches + 1, dtype=np.int64) * p + 1,
        "iptype": np.ones(npatches, np.int64),
        "srcvals": np.ascontiguousarray(srcvals.reshape(-1, 12)),
        "srccoefs": np.ascontiguousarray(srccoefs.reshape(-1, 9)),
    }


def _rectmesh(v1, v2, v3, v4, nu, nv):
    """2 nu nv flat triangles of the quadrilateral v1 v2 v3 v4 (xtri_rectmesh_3d,
    xtri_parameterizations.f90:1020-1077): grid points X(s, t) = (1 - t)((1 - s) v1 + s v2)
    + t((1 - s) v4 + s v3) at s = j / nu, t = i / nv; cell (i, j), rows i outer and
    columns j inner, gives (X00, X10, X01) then (X11, X01, X10) -- the reference's
    triangle order, which fixes the order of the surface's points."""
    v1, v2, v3, v4 = (np.asarray(v, np.float64) for v in (v1, v2, v3, v4))
    s = np.arange(nu + 1) / nu
    t = np.arange(nv + 1) / nv
    S, T = s[None, :, None], t[:, None, None]
    X = (1 - T) * ((1 - S) * v1 + S * v2) + T * ((1 - S) * v4 + S * v3)  # (nv+1, nu+1, 3)
    x00, x10, x01, x11 = X[:-1, :-1], X[:-1, 1:], X[1:, :-1], X[1:, 1:]
    ta = np.stack([x00, x10, x01], axis=2)  # (nv, nu, 3 vertices, 3)
    tb = np.stack([x11, x01, x10], axis=2)
    return list(np.stack([ta, tb], axis=2).reshape(-1, 3, 3))


def cube_skeleton(a, b, c, na, nb, nc):
    """xtri_get_rectparapiped: triangles of the six faces of a box."""
    vc = np.array([[-a, -b, -c], [a, -b, -c], [a, b, -c], [-a, b, -c],
                   [-a, -b, c], [a, -b, c], [a, b, c], [-a, b, c]], dtype=np.float64)
    tris = []
    tris += _rectmesh(vc[0], vc[3], vc[2], vc[1], nb, na)
    tris += _rectmesh(vc[4], vc[5], vc[6], vc[7], na, nb)
    tris += _rectmesh(vc[0], vc[1], vc[5], vc[4], na, nc)
    tris += _rectmesh(vc[3], vc[7], vc[6], vc[2], nc, na)
    tris += _rectmesh(vc[0], vc[4], vc[7], vc[3], nc, nb)
    tris += _rectmesh(vc[1], vc[2], vc[6], vc[5], nb, nc)
    return np.array(tris)  # (ntri, 3, 3): vertex index, xyz


def _projected_eval(skel, abc, c0, radial=None):
    """xtri_ellipsoid_eval (optionally followed by a radial scaling R(theta,phi)
    applied to the unit-sphere point, with chain-rule derivatives)."""
    p0 = skel[:, 0, :][:, None, :]
    e1 = (skel[:, 1, :] - skel[:, 0, :])[:, None, :]
    e2 = (skel[:, 2, :] - skel[:, 0, :])[:, None, :]
    abc = np.asarray(abc, np.float64)
    c0 = np.asarray(c0, np.float64)

    def ev(u, v):
        u = u[None, :, None]
        v = v[None, :, None]
        x = p0 + u * e1 + v * e2
        r = np.linalg.norm(x, axis=-1, keepdims=True)
        drdu = np.sum(x * e1, -1, keepdims=True) / r
        drdv = np.sum(x * e2, -1, keepdims=True) / r
        s = x / r
        su = (r * e1 - x * drdu) / r ** 2
        sv = (r * e2 - x * drdv) / r ** 2
        if radial is not None:
            R, Rs = radial(s)  # R (..,1), gradient of R wrt the unit vector (..,3)
            Ru = np.sum(Rs * su, -1, keepdims=True)
            Rv = np.sum(Rs * sv, -1, keepdims=True)
            su = R * su + Ru * s
            sv = R * sv + Rv * s
            s = R * s
        return abc * s + c0, abc * su, abc * sv

    return ev


def get_ellipsoid_npat(abc, nabc, c0, norder):
    skel = cube_skeleton(1 / np.sqrt(3.0), 1 / np.sqrt(3.0), 1 / np.sqrt(3.0), *nabc)
    return _finish(_projected_eval(skel, abc, c0), len(skel), norder)


def get_sphere_npat(a, na, c0, norder):
    """get_sphere_npat(a, na, c0, norder, iptype=1): 12*na^2 triangles."""
    return get_ellipsoid_npat((a, a, a), (na, na, na), c0, norder)


def get_deformed_sphere(na, norder, amp=0.2):
    """Unit cube-projected sphere scaled radially by 1 + amp*sin(3 theta)*cos(2 phi)
    (theta polar angle from +z, phi = atan2(y, x))."""

    def radial(s):
        x, y, z = s[..., 0:1], s[..., 1:2], s[..., 2:3]
        rho = np.sqrt(x * x + y * y)
        th = np.arctan2(rho, z)
        ph = np.arctan2(y, x)
        R = 1 + amp * np.sin(3 * th) * np.cos(2 * ph)
        dRth = 3 * amp * np.cos(3 * th) * np.cos(2 * ph)
        dRph = -2 * amp * np.sin(3 * th) * np.sin(2 * ph)
        # gradient on the sphere: dR/dth * e_th + dR/dph / sin(th) * e_ph
        st = np.maximum(rho, 1e-300)
        e_th = np.concatenate([z * x / st, z * y / st, -st], -1)
        e_ph = np.concatenate([-y / st, x / st, np.zeros_like(x)], -1)
        g = dRth * e_th + dRph / st * e_ph
        return R, g

    skel = cube_skeleton(1 / np.sqrt(3.0), 1 / np.sqrt(3.0), 1 / np.sqrt(3.0), na, na, na)
    return _finish(_projected_eval(skel, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), radial), len(skel), norder)


def get_wavy_torus(nu, nv, norder, R0=2.0, r0=0.5, amp=0.15, m=3):
    """x(u,v) = ((R0 + r cos v) cos u, (R0 + r cos v) sin u, r sin v),
    r = r0 + amp*cos(m u + v), on an nu x nv grid of (u, v) cells, 2 triangles per cell
    (split as xtri_rectmesh0_3d does)."""
    tri = []
    hu = 2 * np.pi / nu
    hv = 2 * np.pi / nv
    for i in range(nv):
        for j in range(nu):
            u0, v0 = j * hu, i * hv
            # quad corners in (u,v): v1=(u0,v0) v2=(u0+hu,v0) v3=(u0+hu,v0+hv) v4=(u0,v0+hv)
            c1, c2, c3, c4 = (u0, v0), (u0 + hu, v0), (u0 + hu, v0 + hv), (u0, v0 + hv)
            tri.append((c1, c2, c4))
            tri.append((c3, c4, c2))
    tri = np.array(tri)  # (ntri, 3, 2)
    p0 = tri[:, 0, :][:, None, :]
    e1 = (tri[:, 1, :] - tri[:, 0, :])[:, None, :]
    e2 = (tri[:, 2, :] - tri[:, 0, :])[:, None, :]

    def ev(s, t):
        s = s[None, :, None]
        t = t[None, :, None]
        q = p0 + s * e1 + t * e2
        u, v = q[..., 0:1], q[..., 1:2]
        r = r0 + amp * np.cos(m * u + v)
        ru = -amp * m * np.sin(m * u + v)
        rv = -amp * np.sin(m * u + v)
        rho = R0 + r * np.cos(v)
        X = np.concatenate([rho * np.cos(u), rho * np.sin(u), r * np.sin(v)], -1)
        Xu = np.concatenate([ru * np.cos(v) * np.cos(u) - rho * np.sin(u),
                             ru * np.cos(v) * np.sin(u) + rho * np.cos(u), ru * np.sin(v)], -1)
        Xv = np.concatenate([(rv * np.cos(v) - r * np.sin(v)) * np.cos(u),
                             (rv * np.cos(v) - r * np.sin(v)) * np.sin(u),
                             rv * np.sin(v) + r * np.cos(v)], -1)
        dus = e1[..., 0:1] * Xu + e1[..., 1:2] * Xv
        dvt = e2[..., 0:1] * Xu + e2[..., 1:2] * Xv
        return X, dus, dvt

    g = _finish(ev, len(tri), norder)
    return g


def _platonic_cube():
    v = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    f = np.array([[2, 1, 3], [4, 3, 1], [1, 2, 6], [6, 5, 1], [2, 3, 7], [7, 6, 2],
                  [3, 4, 8], [8, 7, 3], [4, 1, 5], [5, 8, 4], [5, 6, 7], [7, 8, 5]]) - 1
    v = v / np.linalg.norm(v[0])
    return v[f]


def get_platonic_sphere(nover, norder):
    """Cube refined nover times by midpoint subdivision (xtri_platonic with the
    refine4_flat child order), mapped radially to the unit sphere (xtri_sphere_eval)."""
    tri = list(_platonic_cube())
    for _ in range(nover):
        n = len(tri)
        for i in range(n):
            a, b, c = tri[i]
            m12, m23, m13 = (a + b) / 2, (b + c) / 2, (a + c) / 2
            tri[i] = np.array([a, m12, m13])
            tri.append(np.array([m23, m13, m12]))
            tri.append(np.array([m12, b, m23]))
            tri.append(np.array([m13, m23, c]))
    skel = np.array(tri)
    return _finish(_projected_eval(skel, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), len(skel), norder)


def read_go3(path):
    """Read a .go3 file (write_go3, src/surface_routs/write_go3.f90:1-24):
    norder, npatches, then srcvals(12, npts) written column by column (all x, all y...)."""
    with open(path) as fh:
        norder = int(fh.readline())
        npat = int(fh.readline())
        vals = np.array([float(x) for x in fh.read().split()])
    npts = vals.size // 12
    return norder, npat, vals.reshape(12, npts).T.copy()


def write_go3(path, norder, npatches, srcvals):
    with open(path, "w") as fh:
        fh.write("%d\n%d\n" % (norder, npatches))
        for l in range(12):
            for v in srcvals[:, l]:
                fh.write("%.17g\n" % v)


def geometry_from_srcvals(norder, npatches, srcvals):
    """Rebuild srccoefs from srcvals sampled at RV nodes (getgeominfo's last step)."""
    umat = basis.rv_umat(norder)
    p = umat.shape[0]
    sv = srcvals.reshape(npatches, p, 12)
    sc = np.einsum("il,pld->pid", umat, sv[:, :, :9])
    return {
        "norders": np.full(npatches, norder, np.int64),
        "ixyzs": np.arange(npatches + 1, dtype=np.int64) * p + 1,
        "iptype": np.ones(npatches, np.int64),
        "srcvals": np.ascontiguousarray(srcvals),
        "srccoefs": np.ascontiguousarray(sc.reshape(-1, 9)),
    }
