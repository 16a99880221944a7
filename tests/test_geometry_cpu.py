"""CPU checks of the geometry generators and the .go3 format (in_go3.f90:15-146,
write_go3.f90:1-24)."""
from pathlib import Path

import numpy as np

GOLD = Path(__file__).resolve().parent / "golden"


def test_go3_round_trip_and_reference_format(tmp_path):
    """write_go3 -> read_go3 gives the same srcvals bit for bit; a file in the Fortran
    list-directed layout of write_go3 (one value per line with leading blanks, the order
    then the patch count, then srcvals column by column) reads back to the fixture data,
    and geometry_from_srcvals rebuilds the cube-sphere coefficients."""
    from fmm3dbie_amd import geometry as G
    d = np.load(GOLD / "sphere_192_o03.npz")
    norder, npat, sv = int(d["norder"]), int(d["npatches"]), d["srcvals"]
    f = tmp_path / "a.go3"
    G.write_go3(f, norder, npat, sv)
    n2, p2, sv2 = G.read_go3(f)
    assert (n2, p2) == (norder, npat) and np.array_equal(sv2, sv)
    # list-directed layout as gfortran writes it: leading blanks, 17 significant digits,
    # E exponents for small values
    g = tmp_path / "b.go3"
    with open(g, "w") as fh:
        fh.write("%20d\n%20d\n" % (norder, npat))
        for l in range(12):
            for v in sv[:, l]:
                fh.write("  %s     \n" % (("%.17E" % v) if abs(v) < 1e-2 and v != 0 else repr(float(v))))
    n3, p3, sv3 = G.read_go3(g)
    assert (n3, p3) == (norder, npat) and np.array_equal(sv3, sv)
    geo = G.geometry_from_srcvals(n3, p3, sv3)
    ref = G.get_platonic_sphere(2, norder)
    assert np.abs(geo["srccoefs"] - ref["srccoefs"]).max() < 1e-13
    assert np.array_equal(geo["ixyzs"], ref["ixyzs"])


def test_rectmesh_grid_and_triangle_order():
    """xtri_rectmesh_3d: the vertices are the bilinear grid of the quadrilateral; cell
    (row i, column j) gives (X00, X10, X01), (X11, X01, X10), rows outer."""
    from fmm3dbie_amd import geometry as G
    v1, v2, v3, v4 = np.array([0.0, 0, 0]), np.array([2.0, 0, 0]), np.array([2.0, 1, 0.5]), np.array([0.0, 1, 0])
    tri = np.array(G._rectmesh(v1, v2, v3, v4, 2, 3))
    assert tri.shape == (12, 3, 3)

    def X(j, i):  # column j of 2, row i of 3
        s, t = j / 2, i / 3
        return (1 - t) * ((1 - s) * v1 + s * v2) + t * ((1 - s) * v4 + s * v3)

    k = 0
    for i in range(3):
        for j in range(2):
            assert np.allclose(tri[k], [X(j, i), X(j + 1, i), X(j, i + 1)], atol=1e-15)
            assert np.allclose(tri[k + 1], [X(j + 1, i + 1), X(j, i + 1), X(j + 1, i)], atol=1e-15)
            k += 2


def test_sphere_npat_on_the_sphere_with_outward_normals():
    """get_sphere_npat (analytic_geometry_routs.f90:299-386): 12 na^2 triangles, every
    node on the sphere, unit outward normals, consistent derivative data (the normal of
    the derivative data), area 4 pi a^2 to the discretization error, each cube face
    covered once (the triangles of one face tile it)."""
    from fmm3dbie_amd import geometry as G, near
    a, na, n = 1.3, 3, 4
    g = G.get_sphere_npat(a, na, (0.1, -0.2, 0.3), n)
    c = np.array([0.1, -0.2, 0.3])
    assert len(g["norders"]) == 12 * na * na
    x = g["srcvals"][:, :3] - c
    r = np.linalg.norm(x, axis=1)
    assert np.abs(r - a).max() < 1e-13
    nrm = g["srcvals"][:, 9:12]
    assert np.abs(nrm - x / r[:, None]).max() < 1e-12
    cr = np.cross(g["srcvals"][:, 3:6], g["srcvals"][:, 6:9])
    assert np.abs(cr / np.linalg.norm(cr, axis=1)[:, None] - nrm).max() < 1e-12
    area = near.get_qwts(g).sum()
    assert abs(area - 4 * np.pi * a * a) < 1e-4 * 4 * np.pi * a * a
    skel = G.cube_skeleton(1, 1, 1, na, na, na)
    ar = 0.5 * np.linalg.norm(np.cross(skel[:, 1] - skel[:, 0], skel[:, 2] - skel[:, 0]), axis=1)
    assert abs(ar.sum() - 6 * 4.0) < 1e-12
