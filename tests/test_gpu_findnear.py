"""GPU near-field finder (findnearmem_/findnear_) against the host finder (k-d tree,
pinned against brute force in the CPU tests) and brute force: exact set and order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _brute(cms, rads, targs):
    d2 = ((targs[:, None, :] - cms[None, :, :]) ** 2).sum(-1)
    ok = d2 <= rads[None, :] ** 2
    rp = np.concatenate([[1], 1 + np.cumsum(ok.sum(1))]).astype(np.int64)
    ci = np.concatenate([np.nonzero(row)[0] + 1 for row in ok]).astype(np.int64) if ok.any() else np.zeros(0, np.int64)
    return rp, ci


def test_findnear_gpu_matches_host_on_surfaces(cuda):
    from fmm3dbie_amd import geometry as G, near
    for g in (G.get_sphere_npat(1.0, 6, (0, 0, 0), 4), G.get_wavy_torus(24, 8, 3), G.get_deformed_sphere(8, 4)):
        cms, rads = near.get_centroid_rads(g)
        rfac, _ = near.get_rfacs(near.norder_avg(g))
        targs = g["srcvals"][:, :3]
        rp, ci = near.findnear(cms, rads * rfac, targs)
        rpg, cig = near.findnear_gpu(cms, rads * rfac, targs)
        assert np.array_equal(rp, rpg) and np.array_equal(ci, cig)


def test_findnear_gpu_matches_brute_force_edge_cases(cuda):
    from fmm3dbie_amd import near
    rng = np.random.default_rng(3)
    cms = rng.uniform(-1, 1, (300, 3))
    rads = rng.uniform(0.0, 0.3, 300)
    rads[:5] = 0.0
    targs = np.concatenate([rng.uniform(-1.5, 1.5, (500, 3)), cms[:20], [[50.0, 0, 0], [-1e6, 2, 3]]])
    # targets exactly on a sphere of radius r_j around c_j (boundary cases)
    d = rng.standard_normal((20, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    targs = np.concatenate([targs, cms[20:40] + rads[20:40, None] * d])
    rp, ci = _brute(cms, rads, targs)
    rpg, cig = near.findnear_gpu(cms, rads, targs)
    assert np.array_equal(rp, rpg) and np.array_equal(ci, cig)
    # no targets, and targets far from every source
    rpg, cig = near.findnear_gpu(cms, rads, np.zeros((0, 3)))
    assert list(rpg) == [1] and len(cig) == 0
    rpg, cig = near.findnear_gpu(cms, rads, np.full((4, 3), 10.0))
    assert list(rpg) == [1] * 5 and len(cig) == 0
