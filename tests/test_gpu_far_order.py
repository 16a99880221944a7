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
