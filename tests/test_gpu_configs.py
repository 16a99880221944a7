"""GPU parity on the benchmark configurations' own geometry and physics:

* the Laplace configuration (cube sphere 4x4 cells per face, order 3, 1920 points,
  alpha = beta = 1): every near weight, the apply and the GMRES solve (tolerance 1e-7)
  against the oracle, and the exterior field of a unit charge at (0.1, 0.2, -0.3) at 10
  points on radius 3 against the exact field;
* the wavy-torus configuration (Helmholtz zk = 3, alpha = i zk, beta = 1, order 6) on a
  16x6 grid: every near weight (plain and selector-routed) and the apply against the
  oracle, the solve against the exact field of the interior point source (2, 0, 0.05);
  on the full 48x16 grid: sampled target rows against the oracle.
Tolerances are rounding-level (the same rules on both sides, summed in different
orders) and are written per pair class."""
import numpy as np
import pytest

from fmm3dbie_amd.layer import SELF_NR, SELF_NT, SELF_TAU
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {0: 1e-10, 1: 1e-12, 2: 1e-13}  # self, adaptive, fixed-rule pairs


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _by_class(w, ref, cls, iquad):
    for c in (0, 1, 2):
        m = np.zeros(len(w), bool)
        for j in np.nonzero(cls == c)[0]:
            m[iquad[j] - 1: iquad[j + 1] - 1] = True
        if m.any():
            assert _rel(w[m], ref[m]) < TOL[c], (c, _rel(w[m], ref[m]))


def _field(g, wts, x, targ, pars, zk):
    d = targ[:, None, :] - g["srcvals"][None, :, :3]
    r = np.linalg.norm(d, axis=-1)
    dn = (d * g["srcvals"][None, :, 9:12]).sum(-1)
    if zk is None:
        ker = (pars[1] * dn / r ** 3 + pars[0] / r) / (4 * np.pi)
    else:
        ker = (pars[2] * dn * (1 - 1j * zk * r) / r ** 3 + pars[1] / r) * np.exp(1j * zk * r) / (4 * np.pi)
    return ker @ (x * wts)


def test_laplace_sphere_4x4_o3(cuda):
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    g = G.get_sphere_npat(1.0, 4, (0, 0, 0), 3)
    pars = np.array([1.0, 1.0])
    P = problem.setup(g, 1e-6, False, pars)
    print("case: laplace sphere 4x4 order 3, %d points, %d pairs" % (len(g["srcvals"]), P["nnz"]), flush=True)
    ref, cls, nref = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6, O.LAP_COMB, pars,
                                   P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT,
                                   self_tau=SELF_TAU, want_class=True)
    nq = layer.NearQuadDevice(g, False, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w = nq.run(want_detail=True).cpu().numpy()
    assert np.array_equal(nq.stats["cls"], cls)
    assert np.array_equal(nq.stats["nref"][cls == 1], nref[cls == 1])
    _by_class(w, ref, cls, P["iquad"])
    # apply of a random density (both algebras) against the oracle's exact add/subtract sum
    N = len(g["srcvals"])
    sig = np.random.default_rng(3).standard_normal(N)
    pot_ref = O.lpcomp_addsub(0, g, g["srcvals"], pars, P["row_ptr"], P["col_ind"], P["iquad"], ref, P["novers"],
                              P["ixyzso"], P["srcover"], P["whtsover"], sig)
    plan = layer.ApplyPlan(g, False, pars, P["row_ptr"], P["col_ind"], P["iquad"], nq.wnear, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    for mode in (0, 1):
        pot = plan.apply(torch.as_tensor(sig + 0j, device=cuda), 64, mode).cpu().numpy()
        assert _rel(pot.real, pot_ref) < 1e-10
    # exterior Dirichlet solve (zid = +1/2, GMRES tolerance 1e-7) against the oracle's GMRES
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = problem.point_source(g["srcvals"][:, :3], x0)
    xo, nito, _, rreso = O.gmres(0, g, P["wts"], pars, P["row_ptr"], P["col_ind"], P["iquad"], ref, P["novers"],
                                 P["ixyzso"], P["srcover"], P["whtsover"], 0.5, rhs, 100, 1e-7)
    wts = torch.as_tensor(P["wts"], device=cuda)
    x, nit, _, rres, _ = layer.gmres(plan, torch.as_tensor(rhs + 0j, device=cuda), 0.5, wts, numit=100,
                                     eps_gmres=1e-7)
    assert int(nit[0]) == nito
    xs = x.cpu().numpy().real
    assert _rel(xs, xo) < 1e-9
    targ = 3.0 * problem.plane_wave_dirs(10, 0)
    exact = problem.point_source(targ, x0)
    err = _rel(_field(g, P["wts"], xs, targ, pars, None), exact)
    print("laplace sphere 4x4 o3: %d iterations, field error %.2e" % (int(nit[0]), err), flush=True)
    assert err < 3e-4  # order-3 discretization
    plan.close()


def _torus_case(nu, nv):
    from fmm3dbie_amd import geometry as G, problem
    g = G.get_wavy_torus(nu, nv, 6)
    zk = 3.0
    pars = np.array([zk, 1j * zk, 1.0])
    return g, pars, zk, problem.setup(g, 1e-6, True, pars)


def test_wavy_torus_16x6_o6(cuda):
    import torch
    from fmm3dbie_amd import layer, problem
    g, pars, zk, P = _torus_case(16, 6)
    N = len(g["srcvals"])
    print("case: wavy torus 16x6 order 6 zk 3, %d points, %d pairs" % (N, P["nnz"]), flush=True)
    ref, cls, nref = O.getnearquad(g, g["srcvals"], P["ipatch_id"], P["uvs_targ"], 1e-6, O.HELM_COMB, pars,
                                   P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT,
                                   self_tau=SELF_TAU, want_class=True)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w = nq.run(want_detail=True).cpu().numpy()
    assert np.array_equal(nq.stats["cls"], cls)
    assert np.array_equal(nq.stats["nref"][cls == 1], nref[cls == 1])
    assert nq.stats["overflow"] == 0
    _by_class(w, ref, cls, P["iquad"])
    # apply (add/subtract and masked FP64, masked FP32) against the oracle
    sig = np.random.default_rng(4).standard_normal(N) + 1j * np.random.default_rng(5).standard_normal(N)
    pot_ref = O.lpcomp_addsub(1, g, g["srcvals"], pars, P["row_ptr"], P["col_ind"], P["iquad"], ref, P["novers"],
                              P["ixyzso"], P["srcover"], P["whtsover"], sig)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], nq.wnear, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    s = torch.as_tensor(sig, device=cuda)
    for mode in (0, 1):
        assert _rel(plan.apply(s, 64, mode).cpu().numpy(), pot_ref) < 1e-10
    assert _rel(plan.apply(s.to(torch.complex64), 32, 1).cpu().numpy(), pot_ref) < 2e-5
    # exterior solve of the interior point source: FP64 GMRES and FP32 inner + FP64
    # refinement, the field at 10 exterior points against the exact field
    x0 = np.array([2.0, 0.0, 0.05])
    rhs = torch.as_tensor(problem.point_source(g["srcvals"][:, :3], x0, zk), device=cuda)
    wts = torch.as_tensor(P["wts"], device=cuda)
    x64, nit, _, rres, _ = layer.gmres(plan, rhs, 0.5, wts, numit=200, eps_gmres=1e-8)
    xir, outer, inner, rr, _, _ = layer.gmres_ir(plan, rhs, 0.5, wts, numit=200, eps_inner=1e-4, tol=1e-8)
    targ = np.array([[4.0, 0.0, 1.0], [0.0, 4.0, -1.0], [-3.5, 1.0, 2.0], [0.5, -3.8, 0.5], [0.0, 0.0, 3.0]])
    exact = problem.point_source(targ, x0, zk)
    e64 = _rel(_field(g, P["wts"], x64.cpu().numpy(), targ, pars, zk), exact)
    eir = _rel(_field(g, P["wts"], xir.cpu().numpy(), targ, pars, zk), exact)
    dx = (xir - x64).cpu().numpy()
    rel = float(np.sqrt((np.abs(dx) ** 2 * P["wts"]).sum() / (np.abs(x64.cpu().numpy()) ** 2 * P["wts"]).sum()))
    print("wavy torus 16x6 o6: %d its, field error %.2e (FP64) %.2e (mixed), mixed vs FP64 %.2e"
          % (int(nit[0]), e64, eir, rel), flush=True)
    assert rres[0] < 1e-8 and rr[0] < 1e-8
    assert e64 < 1e-4 and eir < 1e-4
    assert rel < 1e-7
    plan.close()


def test_wavy_torus_48x16_o6_sampled_rows(cuda):
    from fmm3dbie_amd import layer
    g, pars, zk, P = _torus_case(48, 16)
    N = len(g["srcvals"])
    print("case: wavy torus 48x16 order 6 zk 3, %d points, %d pairs" % (N, P["nnz"]), flush=True)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w = nq.run(want_detail=True).cpu().numpy()
    st = nq.stats
    assert st["overflow"] == 0 and st["nself"] == N
    rows = np.sort(np.random.default_rng(13).choice(N, 60, replace=False))
    rp = P["row_ptr"]
    sub_ptr = np.concatenate([[1], 1 + np.cumsum(np.diff(rp)[rows])]).astype(np.int64)
    cols = np.concatenate([P["col_ind"][rp[r] - 1: rp[r + 1] - 1] for r in rows])
    iq = np.ones(len(cols) + 1, np.int64)
    iq[1:] += np.cumsum(np.diff(g["ixyzs"])[cols - 1])
    ref, cls, nref = O.getnearquad(g, g["srcvals"][rows], P["ipatch_id"][rows], P["uvs_targ"][rows], 1e-6,
                                   O.HELM_COMB, pars, sub_ptr, cols, iq, P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT,
                                   self_tau=SELF_TAU, want_class=True)
    ent = np.concatenate([np.arange(rp[r] - 1, rp[r + 1] - 1) for r in rows])
    assert np.array_equal(st["cls"][ent], cls)
    assert np.array_equal(st["nref"][ent][cls == 1], nref[cls == 1])
    mine = np.concatenate([w[P["iquad"][j] - 1: P["iquad"][j + 1] - 1] for j in ent])
    _by_class(mine, ref, cls, iq)
