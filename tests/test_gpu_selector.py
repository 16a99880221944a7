"""GPU tests of learned order selection: device gradients against the CPU
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


def _pair_err(w, w0, iquad, idx, helm):
    """Max over the given CSR entries of the pair's max weight difference relative to the
    pair's largest weight."""
    worst = 0.0
    for j in idx:
        a, b = iquad[j] - 1, iquad[j + 1] - 1
        worst = max(worst, np.abs(w[a:b] - w0[a:b]).max() / np.abs(w0[a:b]).max())
    return worst


@pytest.mark.parametrize("helm,norder", [(True, 4), (False, 3), (True, 6)])
def test_every_rule_class_stays_within_tolerance(cuda, helm, norder):
    """Every adaptive pair forced into each rule class (XG degree q0 - 4 .. q0 + 12, at
    least 8; no level cap): each pair's weights within 2 eps_adap = 2e-5 (relative to the
    pair's largest weight) of its plain adaptive weights -- the integrator's own error
    test, at tolerance eps_adap = 1e-5 (near_field_routs.f:915-987), judges every
    triangle, so both results are within eps_adap of the exact integrals; the plain class
    gives the plain weights and refinement counts exactly; with a level cap of 1 every
    pair whose tree needs deeper leaves is sent back to the plain integrator and gets
    exactly its weights."""
    import torch
    from fmm3dbie_amd import layer, selector as S
    g, helm, pars, P = _torus(12, 4, norder, helm)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    w0 = nq.run(want_detail=True).cpu().numpy().copy()
    nref0, cls, lev0 = nq.stats["nref"].copy(), nq.stats["cls"].copy(), nq.stats["lev"].copy()
    ad = cls == 1
    idx = np.nonzero(ad)[0]
    for c in range(S.NCLASS):
        sel = torch.as_tensor(S.forced(c), device=cuda)
        w = nq.run(want_detail=True, selector=sel, route_lmax=[20] * 6).cpu().numpy()
        st = nq.stats
        assert np.array_equal(st["pred"][ad], np.full(ad.sum(), c))
        assert st["ncls"][c] == ad.sum() and st["nfallback"] == 0
        e = _pair_err(w, w0, P["iquad"], idx, helm)
        assert e < 2e-5, (c, e)
        if c == S.PLAIN_CLASS:
            assert np.array_equal(w, w0)
            assert np.array_equal(st["nref"][ad], nref0[ad])
        # non-adaptive pairs are untouched
        nad = np.concatenate([np.arange(P["iquad"][j] - 1, P["iquad"][j + 1] - 1) for j in np.nonzero(~ad)[0]])
        assert np.array_equal(w[nad], w0[nad])
    # level cap 1: pairs with deeper leaves fall back to the plain integrator
    sel = torch.as_tensor(S.forced(S.PLAIN_CLASS), device=cuda)
    w = nq.run(want_detail=True, selector=sel, route_lmax=[1] * 6).cpu().numpy()
    deep = lev0[ad] > 1
    assert deep.any() and nq.stats["nfallback"] == deep.sum()
    assert np.array_equal(w, w0)


def test_trained_selector_parity_and_gain(cuda):
    """Learned order selection on the wavy torus of the benchmark (48x16 cells, order 6,
    zk = 3): labels from the integrator itself, GPU training (held-out accuracy
    reported), then the routed quadrature against the plain adaptive path: every pair's
    weights within 2 eps_adap = 2e-5 of the plain ones, at most 0.9 of the plain kernel
    evaluations, and the adaptive kernels at most 0.97 of the plain adaptive kernel's
    time (measured 0.88-0.92).  The whole adaptive part (selector, the rule classes' tree caches and the
    kernels) is printed beside the plain one (tree cache and kernel)."""
    import torch
    from fmm3dbie_amd import layer, selector as S
    g, helm, pars, P = _torus(48, 16, 6, True)
    nq = layer.NearQuadDevice(g, helm, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=cuda)
    X, y, idx, info = S.collect(nq)
    w0 = nq.run(want_detail=True).cpu().numpy().copy()
    params, tinfo = S.train(X, y, n_train=200000, epochs=30, seed=3, device=cuda)
    lmax = S.level_caps(y, info)
    assert tinfo["acc_heldout"] > 0.5, tinfo
    plain, learned, parts, kplain, klearned = [], [], [], [], []
    layer.set_side_stream(False)  # per-kernel times of the kernels alone
    for _ in range(3):
        nq.run()
        plain.append(nq.stats["ms_tree"] + nq.stats["ms_adap"])
        kplain.append(nq.stats["ms_adap"])
        nq.run(selector=params, route_lmax=lmax)
        st = nq.stats
        learned.append(st["ms_select"] + st["ms_tree"] + st["ms_tree_routed"] + st["ms_adap"])
        klearned.append(st["ms_adap"])
        parts.append((st["ms_select"], st["ms_tree"], st["ms_tree_routed"], st["ms_adap"]))
    layer.set_side_stream(True)
    w = nq.run(want_detail=True, selector=params, route_lmax=lmax).cpu().numpy()
    st = nq.stats
    e = _pair_err(w, w0, P["iquad"], idx, helm)
    pred = st["pred"][idx]
    evals_plain = info["evals"][S.PLAIN_CLASS].sum()
    evals_routed = info["evals"][pred, np.arange(len(idx))].sum()
    tp, tl = min(plain), min(learned)
    for c in range(S.NCLASS):
        lv = info["lev"][c][y == c]
        if lv.size:
            print("class %d (degree %d): %d pairs, deepest level quantiles 50/90/99/99.5/100: %s" % (
                c, S.class_degree(info["q0"], c), lv.size, np.percentile(lv, [50, 90, 99, 99.5, 100]).tolist()))
    print("selector: acc %.3f heldout %.3f; classes %s; fallback %d of %d; kernel evaluations %.3f of plain; "
          "adaptive kernels plain %.3f ms learned %.3f ms (x%.2f); "
          "adaptive part plain %.3f ms learned %.3f ms (x%.2f) parts %s; max pair error %.1e; class launch ms %s; "
          "unit costs %s; classes used %s" % (tinfo["acc_train"], tinfo["acc_heldout"], st["ncls"], st["nfallback"],
                                               len(idx), evals_routed / evals_plain, min(kplain), min(klearned),
                                               min(kplain) / min(klearned), tp, tl, tp / tl,
                                               parts[int(np.argmin(learned))], e, np.round(info["ms"], 3).tolist(),
                                               info["unit"].tolist(), info["used"].tolist()), flush=True)
    assert e < 2e-5, e
    assert evals_routed < 0.9 * evals_plain
    assert min(klearned) < 0.97 * min(kplain), (min(kplain), min(klearned))
