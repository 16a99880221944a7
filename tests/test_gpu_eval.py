"""Off-surface evaluation (helm_comb_dir_eval_, lap_comb_dir_eval_) and the solvers with
setup (helm_comb_dir_solver_, lap_comb_dir_solver_): the evaluation against the oracle's
pipeline at far and near-surface targets, the exterior field against the exact
point-source field (to the discretization error), and the same solution as the _guru
path."""
import numpy as np
import pytest

from fmm3dbie_amd.layer import SELF_NR, SELF_NT, SELF_TAU
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _geom(na=4, norder=4):
    from fmm3dbie_amd import geometry as G
    return G.get_sphere_npat(1.0, na, (0, 0, 0), norder)


def _args(g):
    return (len(g["norders"]), g["norders"], g["ixyzs"], g["iptype"], len(g["srcvals"]), g["srccoefs"], g["srcvals"])


def _oracle_eval(g, helm, pars, targ, sigma):
    """The same pipeline on the CPU: host near list and far orders, oracle quadrature for
    off-surface targets, oracle add/subtract apply at the targets."""
    from fmm3dbie_amd import near
    cms, rads = near.get_centroid_rads(g)
    rfac, rfac0 = near.get_rfacs(near.norder_avg(g))
    rp, ci = near.findnear(cms, rads * rfac, targ)
    iq = near.get_iquad_rsc(g, rp, ci)
    zk = float(np.real(pars[0])) if helm else 0.0
    beta = pars[2] if helm else pars[1]
    nov, _ = near.get_far_order(1e-6, g, cms, rads, targ, 0 if abs(beta) > 1e-16 else -1, zk, rp, ci, rfac)
    so, wo, ix = near.oversample_geom(g, nov)
    pid = -np.ones(len(targ), np.int64)
    w = O.getnearquad(g, targ, pid, np.zeros((len(targ), 2)), 1e-6, O.HELM_COMB if helm else O.LAP_COMB, pars, rp,
                      ci, iq, rfac0, O.SELF_POLAR, SELF_NR, SELF_NT, self_tau=SELF_TAU)
    return O.lpcomp_addsub(int(helm), g, targ, pars, rp, ci, iq, w, nov, ix, so, wo, sigma)


def test_helm_solver_and_eval(cuda):
    from fmm3dbie_amd import layer, problem
    g = _geom()
    zk = 1.1
    pars = np.array([zk, 1j * zk, 1.0])
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = problem.point_source(g["srcvals"][:, :3], x0, zk)
    niter, errs, rres, soln = layer.helm_comb_dir_solver(*_args(g), 1e-6, pars, 100, 1, rhs, 1e-8)
    assert rres < 1e-7
    # the same solution through the _guru route with a host setup
    P = problem.setup(g, 1e-6, True, pars)
    w = layer.getnearquad_helm_comb_dir(*_args(g), 1e-6, pars, 1, P["nnz"], P["row_ptr"], P["col_ind"], P["iquad"],
                                        P["rfac0"], P["nquad"])
    ng, _, _, sg = layer.helm_comb_dir_solver_guru(*_args(g), 1e-6, pars, 100, 1, rhs, P["nnz"], P["row_ptr"],
                                                   P["col_ind"], P["iquad"], P["nquad"], w, P["novers"],
                                                   len(P["whtsover"]), P["ixyzso"], P["srcover"], P["whtsover"],
                                                   1e-8)
    assert ng == niter
    assert np.abs(sg - soln).max() / np.abs(sg).max() < 1e-12
    d = problem.plane_wave_dirs(20, 7)
    errs_exact = []
    for rad in (3.0, 1.05):
        targ = rad * d
        pot = layer.helm_comb_dir_eval(*_args(g), 3, len(targ), targ, -np.ones(len(targ), np.int64),
                                       np.zeros((len(targ), 2)), 1e-6, pars, soln)
        ref = _oracle_eval(g, True, pars, targ, soln)
        assert np.abs(pot - ref).max() / np.abs(ref).max() < 1e-10, rad
        exact = problem.point_source(targ, x0, zk)
        errs_exact.append(np.abs(pot - exact).max() / np.abs(exact).max())
    print("helm eval error vs exact field: r=3 %.2e, r=1.05 %.2e" % tuple(errs_exact))
    # 4x4 cube sphere at order 4: discretization error 3e-5 at r = 3 and 5e-4 at distance
    # 0.05 from the surface (the density's interpolation error weighs more there); a smooth
    # rule without the near correction would be off by percents at that distance
    assert max(errs_exact) < 1e-3


def test_lap_solver_and_eval(cuda):
    from fmm3dbie_amd import layer, problem
    g = _geom(4, 4)
    pars = np.array([1.0, 1.0])
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = problem.point_source(g["srcvals"][:, :3], x0)
    niter, errs, rres, soln = layer.lap_comb_dir_solver(*_args(g), 1e-6, pars, 100, 1, rhs, 1e-8)
    assert rres < 1e-7
    d = problem.plane_wave_dirs(20, 8)
    for rad in (3.0, 1.05):
        targ = rad * d
        pot = layer.lap_comb_dir_eval(*_args(g), 3, len(targ), targ, -np.ones(len(targ), np.int64),
                                      np.zeros((len(targ), 2)), 1e-6, pars, soln)
        ref = _oracle_eval(g, False, pars, targ, soln)
        assert np.abs(pot - ref).max() / np.abs(ref).max() < 1e-10, rad
        exact = problem.point_source(targ, x0)
        err = np.abs(pot - exact).max() / np.abs(exact).max()
        print("lap eval r=%.2f error vs exact %.2e" % (rad, err))
        assert err < 1e-3
