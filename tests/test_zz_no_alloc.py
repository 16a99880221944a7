"""The allocation-free apply (fmm3dbie_apply_plan_apply_noalloc, called through the C
ABI) against the allocating apply of the same plan: FP64 to 1e-12 (relative, max norm;
the far-sum partials may be split differently, which changes the summation order) and
FP32 to 1e-5; two calls in a row, the second on a smaller problem; work buffers beyond
the fixed scratch array are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(na, norder, zk, dev):
    from fmm3dbie_amd import geometry as G, layer, problem
    g = G.get_sphere_npat(1.0, na, (0, 0, 0), norder)
    pars = np.array([zk, 1j * zk, 1.0])
    P = problem.setup(g, 1e-6, True, pars)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
    w = nq.run()
    return layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])


def _noalloc(plan, sig, prec, mode):
    import torch
    from fmm3dbie_amd._lib import lib
    pot = torch.full_like(sig, float("nan"))
    nrhs = 1 if sig.dim() == 1 else sig.shape[0]
    rc = lib().fmm3dbie_apply_plan_apply_noalloc(C.c_void_p(plan.handle), prec, mode, C.c_void_p(sig.data_ptr()),
                                                 C.c_void_p(pot.data_ptr()), nrhs,
                                                 C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, pot


def _rel(a, b):
    return float((a.to(b.dtype) - b).abs().max() / b.abs().max())


def test_noalloc_apply_matches_apply():
    import torch
    from fmm3dbie_amd._lib import lib
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    dev = torch.device("cuda:0")
    assert lib().fmm3dbie_apply_scratch_bytes() == 64 << 20
    rng = np.random.default_rng(7)
    # first call: 192 patches, order 4, three right-hand sides; second: 48 patches, order 3
    for na, norder, nrhs in ((4, 4, 3), (2, 3, 1)):
        print("noalloc case: sphere %dx%d order %d, nrhs %d" % (na, na, norder, nrhs), flush=True)
        plan = _plan(na, norder, 1.3, dev)
        npts = plan.info()[0]
        s = rng.standard_normal((nrhs, npts)) + 1j * rng.standard_normal((nrhs, npts))
        sig = torch.as_tensor(s if nrhs > 1 else s[0], device=dev)
        for mode in (0, 1):
            ref = plan.apply(sig, 64, mode)
            rc, pot = _noalloc(plan, sig, 64, mode)
            assert rc == 0
            assert _rel(pot, ref) < 1e-12
        s32 = sig.to(torch.complex64)
        ref32 = plan.apply(s32, 32, 1)
        rc, pot32 = _noalloc(plan, s32, 32, 1)
        assert rc == 0
        assert _rel(pot32, ref32) < 1e-5
        # the Python wrapper reaches the same entry point
        out = torch.empty_like(sig)
        assert _rel(plan.apply_noalloc(sig, out, 64, 1), plan.apply(sig, 64, 1)) < 1e-12
        plan.close()


def test_noalloc_apply_refuses_oversize():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    dev = torch.device("cuda:0")
    plan = _plan(12, 4, 1.1, dev)  # 25 920 points; 64 FP64 right-hand sides need > 64 MiB
    npts, nptso = plan.info()[:2]
    assert (2 * nptso * 64 + npts * 64) * 16 > 64 << 20
    sig = torch.zeros((64, npts), dtype=torch.complex128, device=dev)
    rc, _ = _noalloc(plan, sig, 64, 1)
    assert rc == -7
    # one right-hand side fits, and agrees with the allocating apply
    one = torch.as_tensor(np.random.default_rng(1).standard_normal(npts) + 0j, device=dev)
    rc, pot = _noalloc(plan, one, 64, 1)
    assert rc == 0
    assert _rel(pot, plan.apply(one, 64, 1)) < 1e-12
    plan.close()
