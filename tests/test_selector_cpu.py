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


def test_labels_and_normalisation_folding():
    from fmm3dbie_amd import selector as S
    assert list(S.labels_from_nref([1, 5, 9, 13, 17, 21, 41])) == [0, 1, 2, 3, 4, 5, 5]
    rng = np.random.default_rng(2)
    p = S.init_params(3).astype(np.float64)
    X = rng.standard_normal((20, 8)) * 3 + 1
    mean, std = X.mean(0), X.std(0)
    zf, _ = M.forward(S.fold_normalisation(p, mean, std).astype(np.float64), X)
    zr, _ = M.forward(p, (X - mean) / std)
    assert np.allclose(zf, zr, atol=1e-4)
