"""GPU far-field order (get_far_order_) against the C oracle's get_far_order and the
host restatement: identical orders per patch."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geo,norder,zk,iker", [("sphere", 4, 1.1, 0), ("sphere", 3, 0.0, 0), ("torus", 6, 3.0, 0),
                                                ("deformed", 4, 10.0, -1)])
def test_far_order_gpu_matches_oracle(cuda, geo, norder, zk, iker):
    from fmm3dbie_amd import geometry as G, near
    if geo == "sphere":
        g = G.get_sphere_npat(1.0, 4, (0, 0, 0), norder)
    elif geo == "torus":
        g = G.get_wavy_torus(16, 6, norder)
    else:
        g = G.get_deformed_sphere(6, norder)
    cms, rads = near.get_centroid_rads(g)
    rfac, _ = near.get_rfacs(near.norder_avg(g))
    targs = g["srcvals"][:, :3]
    rp, ci = near.findnear(cms, rads * rfac, targs)
    ref, ref_ix = O.far_order(1e-6, g, cms, rads, targs, iker, zk, rp, ci, rfac)
    nov, ix = near.get_far_order_gpu(1e-6, g, cms, rads, targs, iker, zk, rp, ci, rfac)
    assert np.array_equal(nov, ref), (np.nonzero(nov != ref)[0][:10], nov[nov != ref][:10], ref[nov != ref][:10])
    assert np.array_equal(ix, ref_ix)
    host, _ = near.get_far_order(1e-6, g, cms, rads, targs, iker, zk, rp, ci, rfac)
    assert np.array_equal(host, ref)


@pytest.mark.parametrize("geo,norder", [("sphere", 3), ("sphere", 4), ("torus", 6), ("sphere", 9)])
def test_oversample_geom_and_qwts_gpu_match_oracle(cuda, geo, norder):
    """oversample_geom_ / get_qwts_ on the GPU against the oracle's oversample_geom (with
    the clash-node copy of get_clashing_indices for orders divisible by 3) and the host
    get_qwts, for every oversampling order 1..16 (and one equal to the patch order)."""
    from fmm3dbie_amd import geometry as G, near
    g = G.get_sphere_npat(1.0, 2, (0, 0, 0), norder) if geo == "sphere" else G.get_wavy_torus(8, 4, norder)
    npat = len(g["norders"])
    rng = np.random.default_rng(norder)
    novers = rng.choice([1, 2, 3, 4, 6, 8, 10, 12, 14, 16, norder], npat).astype(np.int64)
    so, wo, ixo = near.oversample_geom_gpu(g, novers)
    ref_so, ref_wo = O.oversample_geom(g, novers, ixo)
    assert np.abs(so - ref_so).max() < 1e-14 * np.abs(ref_so).max()
    assert np.abs(wo - ref_wo).max() < 1e-14 * np.abs(ref_wo).max()
    hs, hw, hix = near.oversample_geom(g, novers)
    assert np.array_equal(hix, ixo)
    assert np.abs(hs - so).max() < 1e-14 * np.abs(hs).max()
    w = near.get_qwts_gpu(g)
    assert np.abs(w - near.get_qwts(g)).max() < 1e-15 * np.abs(w).max()
