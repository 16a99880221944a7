"""GPU tests of batched right-hand sides: the RHS-grouped far sum and fused reduction +
CSR give the single-RHS results, the autotuner keeps results unchanged, and batched
mixed-precision GMRES matches the FP64 solves and the Mie series."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _sphere(zk, na=2, norder=4):
    from fmm3dbie_amd import geometry as G, layer, problem
    g = G.get_sphere_npat(1.0, na, (0, 0, 0), norder)
    pars = np.array([zk, 1j * zk, 1.0])
    P = problem.setup(g, 1e-6, True, pars)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"])
    w = nq.run().cpu().numpy()
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    return g, pars, P, plan


def test_batched_apply_and_tuning(cuda):
    import torch
    g, pars, P, plan = _sphere(5.0)
    N = len(g["srcvals"])
    rng = np.random.default_rng(4)
    S = rng.standard_normal((64, N)) + 1j * rng.standard_normal((64, N))
    dS = torch.as_tensor(S, device=cuda)
    single = np.stack([plan.apply(dS[r], 64, 0).cpu().numpy() for r in range(64)])
    for mode in (0, 1):
        assert _rel(plan.apply(dS, 64, mode).cpu().numpy(), single) < 1e-12
    rep64 = plan.tune(64, 64)
    assert len(rep64) == 4 * 4  # groups 1, 2, 4, 8 x four workgroup targets
    assert _rel(plan.apply(dS, 64, 1).cpu().numpy(), single) < 1e-12
    S32 = dS.to(torch.complex64)
    before = plan.apply(S32, 32, 1).cpu().numpy()
    rep32 = plan.tune(32, 64)
    assert len(rep32) == 5 * 4
    after = plan.apply(S32, 32, 1).cpu().numpy()
    for r in range(64):
        assert _rel(after[r], single[r]) < 1e-5
        assert _rel(before[r], single[r]) < 1e-5
    # odd batch sizes use the group sizes that divide them
    assert _rel(plan.apply(dS[:3], 64, 1).cpu().numpy(), single[:3]) < 1e-12
    plan.close()


def test_batched_mixed_gmres_matches_fp64_and_mie(cuda):
    import torch
    from fmm3dbie_amd import layer, problem
    zk = 5.0
    g, pars, P, plan = _sphere(zk, na=3)
    dirs = problem.plane_wave_dirs(16, 2)
    rhs = -problem.plane_waves(g["srcvals"][:, :3], dirs, zk)
    wts = torch.as_tensor(P["wts"], device=cuda)
    drhs = torch.as_tensor(rhs, device=cuda)
    x64, nit, _, rres, _ = layer.gmres(plan, drhs, 0.5, wts, numit=200, eps_gmres=1e-8)
    plan.tune(32, 16)
    xi, outer, inner, rr, n32, n64 = layer.gmres_ir(plan, drhs, 0.5, wts, numit=200, eps_inner=1e-4, tol=1e-7)
    x64 = x64.cpu().numpy()
    xi = xi.cpu().numpy()
    for r in range(16):
        assert _rel(xi[r], x64[r]) < 1e-6
    # scattered field at 10 points of radius 3 against the Mie series
    targ = 3.0 * problem.plane_wave_dirs(10, 0)
    d = targ[:, None, :] - g["srcvals"][None, :, :3]
    rr_ = np.linalg.norm(d, axis=-1)
    dn = (d * g["srcvals"][None, :, 9:12]).sum(-1)
    ker = (dn * (1 - 1j * zk * rr_) / rr_ ** 3 + 1j * zk / rr_) * np.exp(1j * zk * rr_) / (4 * np.pi)
    u = (xi * P["wts"][None, :]) @ ker.T
    exact = problem.mie_sound_soft(targ, dirs, zk)
    err = np.abs(u - exact).max() / np.abs(exact).max()
    print("batched IR: outer %s inner %s; Mie field error %.2e" % (outer, inner, err))
    # discretisation error of 108 order-4 triangles at k a = 5
    assert err < 1e-3
    plan.close()


def test_fp32_far_sum_phase_beyond_hardware_sine_range(cuda):
    """|k| x box diagonal / (2 pi) = 220 revolutions, beyond what the plan lets the
    hardware sine take directly: the FP32 far sum reduces the phase first and matches the
    FP64 far sum to FP32 accuracy (near correction zero, same masked algebra)."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    g = G.get_sphere_npat(1.0, 2, (0, 0, 0), 4)
    zk = 400.0
    pars = np.array([zk, 1j * zk, 1.0])
    P = problem.setup(g, 1e-6, True, pars)
    w = np.zeros(P["nquad"], complex)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    rng = np.random.default_rng(7)
    N = len(g["srcvals"])
    dS = torch.as_tensor(rng.standard_normal(N) + 1j * rng.standard_normal(N), device=cuda)
    p64 = plan.apply(dS, 64, 1).cpu().numpy()
    p32 = plan.apply(dS.to(torch.complex64), 32, 1).cpu().numpy()
    assert _rel(p32, p64) < 2e-4, _rel(p32, p64)
    plan.close()


@pytest.mark.parametrize("zk,tol", [(5.0, 1e-5), (400.0, 2e-4), (3.0 + 1.5j, 1e-5)])
def test_matrix_core_far_sum_matches_fp64(cuda, zk, tol):
    """The FP32 far sum of 16 right-hand sides on the matrix cores (plain phases, phases
    beyond the hardware sine's range, a complex wavenumber) against the FP64 apply, one
    right-hand side at a time; 32 right-hand sides run as two groups."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    g = G.get_sphere_npat(1.0, 2, (0, 0, 0), 4)
    pars = np.array([zk, 1j * zk, 1.0])
    P = problem.setup(g, 1e-6, True, pars)
    w = np.zeros(P["nquad"], complex)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    rng = np.random.default_rng(11)
    N = len(g["srcvals"])
    dS = torch.as_tensor(rng.standard_normal((32, N)) + 1j * rng.standard_normal((32, N)), device=cuda)
    ref = np.stack([plan.apply(dS[r], 64, 1).cpu().numpy() for r in range(32)])
    p32 = plan.apply(dS.to(torch.complex64), 32, 1).cpu().numpy()
    for r in range(32):
        assert _rel(p32[r], ref[r]) < tol, (r, _rel(p32[r], ref[r]))
    plan.close()
