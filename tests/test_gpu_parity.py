"""GPU parity: the HIP kernels (through the C-ABI) against the CPU oracle on the same
inputs, plus closed forms and solver-level checks.  Marked gpu."""
import numpy as np
import pytest

from fmm3dbie_amd.layer import SELF_NR, SELF_NT, SELF_TAU
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _case(kind, norder, na, eps=1e-6):
    """helm: zk = 1.1; helmc: complex zk = 1.1 + 0.3i (decaying kernel); lap: Laplace."""
    from fmm3dbie_amd import geometry as G, problem
    helm = kind in ("helm", "helmc")
    g = G.get_sphere_npat(1.0, na, (0, 0, 0), norder)
    zk = 1.1 + 0.3j if kind == "helmc" else 1.1
    pars = np.array([zk, 1j * zk, 1.0]) if helm else np.array([1.0, 1.0])
    P = problem.setup(g, eps, helm, pars)
    return g, helm, pars, P


# orders 8 and 9: 2 * (p + 1)(p + 2) / 2 moment components exceed the 64 lanes of a wavefront
CASES = [("helm", 4, 2), ("lap", 3, 2), ("helm", 6, 2), ("helm", 3, 3), ("helm", 8, 2), ("lap", 9, 1),
         ("helmc", 4, 2)]


@pytest.mark.parametrize("kind,norder,na", CASES)
def test_nearquad_device_matches_oracle(cuda, kind, norder, na):
    from fmm3dbie_amd import layer
    g, helm, pars, P = _case(kind, norder, na)
    ref, cls, nref = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6,
                                   O.HELM_COMB if helm else O.LAP_COMB, pars, P["row_ptr"], P["col_ind"], P["iquad"],
                                   P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT, self_tau=SELF_TAU, want_class=True)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w = nq.run(want_detail=True).cpu().numpy()
    # same pair classification and the same refinement tree (levels 0..20, 1->4)
    assert np.array_equal(nq.stats["cls"], cls)
    assert np.array_equal(nq.stats["nref"][cls == 1], nref[cls == 1])
    assert nq.stats["overflow"] == 0
    tol = {0: 1e-10, 1: 1e-12, 2: 1e-13}
    for c in (0, 1, 2):
        m = np.zeros(P["nquad"], bool)
        for j in np.nonzero(cls == c)[0]:
            m[P["iquad"][j] - 1: P["iquad"][j + 1] - 1] = True
        if m.any():
            assert _rel(w[m], ref[m]) < tol[c], (c, _rel(w[m], ref[m]))


def test_host_dropins_equal_device(cuda):
    from fmm3dbie_amd import layer
    for kind, norder, na in [("helm", 4, 2), ("lap", 3, 2)]:
        g, helm, pars, P = _case(kind, norder, na)
        nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"],
                                  device=cuda)
        wd = nq.run().cpu().numpy()
        args = (len(g["norders"]), g["norders"], g["ixyzs"], g["iptype"], len(g["srcvals"]), g["srccoefs"],
                g["srcvals"], 1e-6, pars, 1, P["nnz"], P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"],
                P["nquad"])
        wh = layer.getnearquad_helm_comb_dir(*args) if helm else layer.getnearquad_lap_comb_dir(*args)
        assert np.array_equal(wh, wd)


def test_eval_offsurface_targets_matches_oracle(cuda):
    """getnearquad_helm_comb_dir_eval with off-surface targets (ipatch_id = -1)."""
    from fmm3dbie_amd import layer, near
    g, helm, pars, P = _case("helm", 4, 2)
    rng = np.random.default_rng(5)
    d = rng.standard_normal((300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    targs = d * rng.uniform(0.85, 1.15, (300, 1))
    rp, ci = near.findnear(P["cms"], P["rads"] * P["rfac"], targs)
    iq = near.get_iquad_rsc(g, rp, ci)
    pid = -np.ones(len(targs), np.int64)
    uvs = np.zeros((len(targs), 2))
    ref = O.getnearquad(g, targs, pid, uvs, 1e-6, O.HELM_COMB, pars, rp, ci, iq, P["rfac0"], O.SELF_POLAR,
                        SELF_NR, SELF_NT, self_tau=SELF_TAU)
    w = layer.getnearquad_helm_comb_dir_eval(len(g["norders"]), g["norders"], g["ixyzs"], g["iptype"],
                                             len(g["srcvals"]), g["srccoefs"], g["srcvals"], 3, len(targs), targs,
                                             pid, uvs, 1e-6, pars, 1, len(ci), rp, ci, iq, P["rfac0"], int(iq[-1]) - 1)
    assert _rel(w, ref) < 1e-12


@pytest.mark.parametrize("kind,norder,na", [("helm", 4, 2), ("lap", 3, 2), ("helmc", 4, 2)])
def test_lpcomp_matches_oracle(cuda, kind, norder, na):
    import torch
    from fmm3dbie_amd import layer
    g, helm, pars, P = _case(kind, norder, na)
    w = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6, O.HELM_COMB if helm else O.LAP_COMB,
                      pars, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT,
                      self_tau=SELF_TAU)
    N = len(g["srcvals"])
    rng = np.random.default_rng(0)
    sig = rng.standard_normal(N) + (1j * rng.standard_normal(N) if helm else 0)
    ref = O.lpcomp_addsub(int(helm), g, g["srcvals"], pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"],
                          P["ixyzso"], P["srcover"], P["whtsover"], sig)
    args = (len(g["norders"]), g["norders"], g["ixyzs"], g["iptype"], N, g["srccoefs"], g["srcvals"], 1e-6, pars,
            P["nnz"], P["row_ptr"], P["col_ind"], P["iquad"], P["nquad"], w, P["novers"], len(P["whtsover"]),
            P["ixyzso"], P["srcover"], P["whtsover"], sig)
    pot = layer.lpcomp_helm_comb_dir_addsub(*args) if helm else layer.lpcomp_lap_comb_dir_addsub(*args)
    assert _rel(pot, ref) < 1e-10
    if helm:
        plan = layer.ApplyPlan(g, helm, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                               P["srcover"], P["whtsover"])
        s = torch.as_tensor(sig, device=cuda)
        for mode in (0, 1):
            assert _rel(plan.apply(s, 64, mode).cpu().numpy(), ref) < 1e-10
        p32 = plan.apply(s.to(torch.complex64), 32, 1).cpu().numpy()
        assert _rel(p32, ref) < 1e-5
        # batched right-hand sides give the single-RHS results
        S = torch.stack([s, 2 * s, 1j * s])
        PB = plan.apply(S, 64, 1).cpu().numpy()
        for r, c in enumerate([1, 2, 1j]):
            assert _rel(PB[r], c * ref) < 1e-10
        PB32 = plan.apply(S.to(torch.complex64), 32, 1).cpu().numpy()
        assert _rel(PB32[2], 1j * ref) < 1e-5
        plan.close()


def test_gmres_matches_oracle_and_point_source(cuda):
    import torch
    from fmm3dbie_amd import layer, problem
    g, helm, pars, P = _case("helm", 4, 3)
    w = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6, O.HELM_COMB, pars, P["row_ptr"],
                      P["col_ind"], P["iquad"], P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT, self_tau=SELF_TAU)
    zk = 1.1
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = problem.point_source(g["srcvals"][:, :3], x0, zk)
    xo, nito, _, rreso = O.gmres(1, g, P["wts"], pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"],
                                 P["ixyzso"], P["srcover"], P["whtsover"], 0.5, rhs, 100, 1e-7)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    wts = torch.as_tensor(P["wts"], device=cuda)
    x, nit, errs, rres, _ = layer.gmres(plan, torch.as_tensor(rhs, device=cuda), 0.5, wts, numit=100,
                                        eps_gmres=1e-7)
    assert int(nit[0]) == nito
    assert abs(rres[0] - rreso) < 1e-3 * rreso + 1e-12
    xs = x.cpu().numpy()
    assert _rel(xs, xo) < 1e-9
    # exterior field at 10 points on radius 3 equals the point-source field
    dirs = problem.plane_wave_dirs(10, 0)
    targ = 3.0 * dirs
    d = targ[:, None, :] - g["srcvals"][None, :, :3]
    r = np.linalg.norm(d, axis=-1)
    dn = (d * g["srcvals"][None, :, 9:12]).sum(-1)
    ker = (dn * (1 - 1j * zk * r) / r ** 3 + 1j * zk / r) * np.exp(1j * zk * r) / (4 * np.pi)
    u = ker @ (xs * P["wts"])
    exact = problem.point_source(targ, x0, zk)
    assert np.abs(u - exact).max() / np.abs(exact).max() < 1e-4
    # mixed precision iterative refinement reaches the FP64 answer
    xi, outer, inner, rr, n32, n64 = layer.gmres_ir(plan, torch.as_tensor(rhs, device=cuda), 0.5, wts, numit=100,
                                                   eps_inner=1e-4, tol=1e-7)
    assert rr[0] < 1e-7
    assert _rel(xi.cpu().numpy(), xo) < 1e-6
    plan.close()


def test_solver_guru_dropin(cuda):
    from fmm3dbie_amd import layer, problem
    g, helm, pars, P = _case("lap", 3, 2)
    w = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6, O.LAP_COMB, pars, P["row_ptr"],
                      P["col_ind"], P["iquad"], P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT, self_tau=SELF_TAU)
    rhs = problem.point_source(g["srcvals"][:, :3], [0.1, 0.2, -0.3])
    niter, errs, rres, soln = layer.lap_comb_dir_solver_guru(
        len(g["norders"]), g["norders"], g["ixyzs"], g["iptype"], len(g["srcvals"]), g["srccoefs"], g["srcvals"],
        1e-6, pars, 100, 1, rhs, P["nnz"], P["row_ptr"], P["col_ind"], P["iquad"], P["nquad"], w, P["novers"],
        len(P["whtsover"]), P["ixyzso"], P["srcover"], P["whtsover"], 1e-10)
    xo, nito, _, rreso = O.gmres(0, g, P["wts"], pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"],
                                 P["ixyzso"], P["srcover"], P["whtsover"], 0.5, rhs, 100, 1e-10)
    assert niter == nito
    assert _rel(soln, xo) < 1e-9


def test_smoke_entry(cuda):
    import __graft_entry__
    __graft_entry__.smoke()


def test_side_stream_gives_the_same_weights(cuda):
    """The fixed-rule and self kernels (and, for on-surface targets, the tree cache) on the
    library's second stream give exactly the weights of the in-order run on one stream."""
    from fmm3dbie_amd import layer
    g, helm, pars, P = _case("helm", 4, 2)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    try:
        layer.set_side_stream(False)
        w0 = nq.run().cpu().numpy()
        layer.set_side_stream(True)
        for _ in range(3):
            assert np.array_equal(nq.run().cpu().numpy(), w0)
    finally:
        layer.set_side_stream(True)
