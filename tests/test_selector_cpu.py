"""CPU checks of the learned-selector oracle and host logic (no GPU)."""
import numpy as np

from oracle import mlp_oracle as M


def _data(n=37, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.NI))
    y = rng.integers(0, M.NO, n)
    p = rng.standard_normal(M.NPARAM) * 0.3
    return p, X, y


def test_param_count_matches_library_layout():
    assert M.NPARAM == 64 * 8 + 64 + 64 * 64 + 64 + 6 * 64 + 6


def test_oracle_gradient_central_differences():
    """Every parameter block's gradient against central differences of the loss."""
    p, X, y = _data()
    _, g = M.loss_grad(p, X, y)
    rng = np.random.default_rng(1)
    o = 0
    for name, shp in M.SIZES:
        n = int(np.prod(shp))
        for k in rng.choice(n, min(n, 12), replace=False):
            i = o + k
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (M.loss_grad(pp, X, y)[0] - M.loss_grad(pm, X, y)[0]) / (2 * h)
            assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd), (name, k, fd, g[i])
        o += n


def test_oracle_closed_form_at_zero_weights():
    """Zero parameters: uniform softmax, loss log 6, d loss / d b3 = 1/6 - class frequency,
    and no gradient reaches the hidden layers (ReLU of 0 is inactive)."""
    _, X, y = _data(50)
    p = np.zeros(M.NPARAM)
    loss, g = M.loss_grad(p, X, y)
    assert abs(loss - np.log(6.0)) < 1e-14
    q = M.unpack(g)
    freq = np.bincount(y, minlength=6) / len(y)
    assert np.allclose(q["b3"], 1.0 / 6 - freq, atol=1e-15)
    assert np.all(q["W1"] == 0) and np.all(q["W2"] == 0) and np.all(q["W3"] == 0)


def test_oracle_adam_first_step_is_signed_lr():
    """Adam's first bias-corrected step is lr * sign(g) (up to eps)."""
    g = np.array([0.3, -2.0, 1e-3])
    p1, _, _ = M.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p1, -0.01 * np.sign(g), rtol=1e-4)


def test_rule_classes_levels_and_normalisation_folding():
    from fmm3dbie_amd import selector as S
    # the rule classes around the plain degree 16 (55 nodes), 12 (33 nodes) and 22 (96)
    assert [S.class_degree(16, c) for c in range(6)] == [12, 14, 16, 20, 24, 28]
    assert [S.class_degree(12, c) for c in range(6)] == [8, 10, 12, 16, 20, 24]
    assert [S.class_degree(22, c) for c in range(6)] == [18, 20, 22, 26, 30, 30]
    assert [S.class_degree(10, c) for c in range(6)] == [8, 8, 10, 14, 18, 22]
    assert S.plain_degree(55) == 16 and S.plain_degree(33) == 12
    assert S.class_degree(S.plain_degree(55), S.PLAIN_CLASS) == 16
    # forced parameters select their class for any input
    z, _ = M.forward(S.forced(2).astype(np.float64), np.random.default_rng(0).standard_normal((5, 8)))
    assert np.all(np.argmax(z, 1) == 2)
    # level caps: deepest tree of the pairs labelled with each class, 20 without pairs
    y = np.array([0, 0, 3, 3, 3])
    lev = np.zeros((6, 5), np.int32)
    lev[0] = [2, 4, 9, 9, 9]
    lev[3] = [9, 9, 1, 3, 2]
    caps = S.level_caps(y, {"lev": lev})
    assert list(caps) == [4, 20, 20, 3, 20, 20]
    rng = np.random.default_rng(2)
    p = S.init_params(3).astype(np.float64)
    X = rng.standard_normal((20, 8)) * 3 + 1
    mean, std = X.mean(0), X.std(0)
    zf, _ = M.forward(S.fold_normalisation(p, mean, std).astype(np.float64), X)
    zr, _ = M.forward(p, (X - mean) / std)
    assert np.allclose(zf, zr, atol=1e-4)


def test_open_classes_pays_for_each_tree_cache():
    """A class is used only when its pairs save more than its tree cache costs; the plain
    class is always used; labels are the least-cost used class."""
    from fmm3dbie_amd import selector as S
    P = S.PLAIN_CLASS
    n = 100
    cost = np.full((6, n), 10.0)
    # class 0 saves 1 on every pair (100 in total), class 5 saves 5 on 10 pairs (50, and
    # 40 more than class 0 there)
    cost[0] = 9.0
    cost[5, :10] = 5.0
    cache = np.full(6, 1e9)
    cache[P] = 0.0
    cache[0], cache[5] = 60.0, 30.0
    y, used = S.open_classes(cost, cache)
    assert used.tolist() == [True, False, P == 2, False, False, True]
    assert np.all(y[:10] == 5) and np.all(y[10:] == 0)
    # with a dearer cache class 5 no longer pays (its extra saving over class 0 is 40)
    cache[5] = 45.0
    y, used = S.open_classes(cost, cache)
    assert not used[5] and np.all(y == 0)
    # nothing pays: everything stays on the plain rule
    cache[0], cache[5] = 200.0, 60.0
    y, used = S.open_classes(cost, cache)
    assert used.sum() == 1 and used[P] and np.all(y == P)
    # the plain class's cache is free; classes of the plain degree too
    c = S.tree_cost(0.3, [33, 42, 55, 79, 112, 150], 16)
    assert c[P] == 0.0 and abs(c[5] - 0.3 * 150 / 55) < 1e-15


def test_chunk_counts_follow_the_kernels_chunking():
    """Chunks of the adaptive kernel's root step and refinement test (the cost model's
    unit): one per triangle at order 6 for rules of 28..55 nodes, else equal chunks of at
    most 55 (order >= 6) or 64 nodes."""
    from fmm3dbie_amd import selector as S
    assert S.chunks(6, 55) == (1, 4)
    assert S.chunks(6, 33) == (1, 4)
    assert S.chunks(6, 16) == (1, 2)
    assert S.chunks(6, 79) == (2, 6)
    assert S.chunks(6, 171) == (4, 13)
    assert S.chunks(4, 55) == (1, 4)
    assert S.chunks(4, 79) == (2, 5)
