"""GPU tests of the learned refinement-depth selector: device gradients against the CPU
oracle, training, exactness of the routed quadrature (any routing gives the adaptive
weights), and the gain of the trained network over the plain adaptive path."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _torus(nu, nv, norder, helm=True, zk=3.0):
    from fmm3dbie_amd import geometry as G, problem
    g = G.get_wavy_torus(nu, nv, norder)
    pars = np.array([zk, 1j * zk, 1.0]) if helm else np.array([1.0, 1.0])
    return g, helm, pars, problem.setup(g, 1e-6, helm, pars)


def _forced(cls):
    """Parameters whose argmax is `cls` for every input (all weights zero)."""
    p = np.zeros(M.NPARAM, np.float32)
    p[-6 + cls] = 1.0
    return p


def test_mlp_grad_matches_oracle(cuda):
    from fmm3dbie_amd import selector as S
    rng = np.random.default_rng(7)
    n = 1000  # not a multiple of the 64-sample workgroup
    X = rng.standard_normal((n, 8)).astype(np.float32)
    y = rng.integers(0, 6, n).astype(np.int32)
    p = S.init_params(11)
    p[-6:] = rng.standard_normal(6) * 0.1
    loss, g = S.grad(p, X, y, device=cuda)
    lref, gref = M.loss_grad(p.astype(np.float64), X.astype(np.float64), y)
    assert abs(loss - lref) < 1e-5 * max(1.0, abs(lref))
    assert np.abs(g - gref).max() < 2e-5 * np.abs(gref).max(), np.abs(g - gref).max() / np.abs(gref).max()
    # predictions agree with the oracle forward pass (ties aside)
    z, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    zs = np.sort(z, axis=1)
    clear = zs[:, -1] - zs[:, -2] > 1e-4
    assert np.array_equal(S.predict(p, X, device=cuda)[clear], np.argmax(z, axis=1)[clear])


def test_mlp_training_learns_a_separable_rule(cuda):
    from fmm3dbie_amd import selector as S
    rng = np.random.default_rng(8)
    X = rng.standard_normal((8192, 8)).astype(np.float32)
    y = np.digitize(X[:, 0] + 0.5 * X[:, 1], [-1.5, -0.5, 0.0, 0.5, 1.5]).astype(np.int32)
    params, info = S.train(X, y, n_train=6144, epochs=40, batch=256, lr=3e-3, seed=3, device=cuda)
    assert info["loss"][-1] < 0.5 * info["loss"][0]
    assert info["acc_heldout"] > 0.9, info


@pytest.mark.parametrize("helm,norder", [(True, 4), (False, 3), (True, 6)])
def test_any_routing_gives_the_adaptive_weights(cuda, helm, norder):
    """Every adaptive pair forced through the depth-1 kernel, the depth-2 kernel, or the
    adaptive kernel in predicted order: same weights and refinement counts as the plain
    path (fallback of the pairs whose tree is deeper than the kernel's)."""
    import torch
    from fmm3dbie_amd import layer
    g, helm, pars, P = _torus(12, 4, norder, helm)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w0 = nq.run(want_detail=True).cpu().numpy().copy()
    nref0, cls = nq.stats["nref"].copy(), nq.stats["cls"].copy()
    ad = cls == 1
    m = (nref0[ad] - 1) // 4
    assert (m == 0).any() and (m == 1).any() and (m >= 2).any()
    scale = np.abs(w0).max()
    for c in (0, 1, 5):
        sel = torch.as_tensor(_forced(c), device=cuda)
        w = nq.run(want_detail=True, selector=sel).cpu().numpy()
        st = nq.stats
        assert np.array_equal(st["pred"][ad], np.full(ad.sum(), c))
        assert np.array_equal(st["nref"][ad], nref0[ad]), c
        assert np.abs(w - w0).max() < 1e-13 * scale, (c, np.abs(w - w0).max() / scale)
        if c == 0:
            assert st["nshallow0"] == (m == 0).sum() and st["nfallback"] == (m > 0).sum()
        if c == 1:
            assert st["nshallow0"] == (m == 0).sum() and st["nshallow1"] == (m == 1).sum()
            assert st["nfallback"] == (m > 1).sum()


def test_trained_selector_parity_and_gain(cuda):
    """Train on the integrator's own labels (held-out accuracy reported), then compare
    the learned path with the plain adaptive path: identical weights, and the time of the
    adaptive part (selector + cached-tree kernels + adaptive kernel) against the plain
    adaptive kernel."""
    from fmm3dbie_amd import layer, selector as S
    g, helm, pars, P = _torus(24, 8, 6, True)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    X, y, idx = S.collect(nq)
    w0 = nq.wnear.cpu().numpy().copy()
    nref0 = nq.stats["nref"].copy()
    params, info = S.train(X, y, n_train=int(0.7 * len(y)), epochs=30, seed=3, device=cuda)
    assert info["acc_heldout"] > 0.6, info
    plain, learned, parts = [], [], []
    layer.set_side_stream(False)  # per-kernel times of the kernels alone
    for _ in range(3):
        nq.run()
        plain.append(nq.stats["ms_tree"] + nq.stats["ms_adap"])
        nq.run(selector=params)
        st = nq.stats
        learned.append(st["ms_select"] + st["ms_tree"] + st["ms_shallow"] + st["ms_adap"])
        parts.append((st["ms_select"], st["ms_tree"], st["ms_shallow"], st["ms_adap"]))
    layer.set_side_stream(True)
    print("selector parts (select, tree, cached-tree kernels, adaptive) ms:", parts[int(np.argmin(learned))])
    w = nq.run(want_detail=True, selector=params).cpu().numpy()
    assert np.array_equal(nq.stats["nref"][idx], nref0[idx])
    assert np.abs(w - w0).max() < 1e-13 * np.abs(w0).max()
    tp, tl = min(plain), min(learned)
    print("selector: acc %.3f heldout %.3f; adaptive part plain %.3f ms learned %.3f ms (x%.2f); "
          "accepted depth1 %d depth2 %d fallback %d of %d" % (info["acc_train"], info["acc_heldout"], tp, tl,
                                                             tp / tl, nq.stats["nshallow0"], nq.stats["nshallow1"],
                                                             nq.stats["nfallback"], len(idx)))
    # the routing must not cost more than it saves; the gain itself is measured by the
    # benchmark at full size (a few per cent here)
    assert tl < 1.15 * tp, (tp, tl)
