"""CPU oracle of the learned refinement-depth selector (test infrastructure only).

Float64 numpy restatement of the 8 -> 64 -> 64 -> 6 MLP whose forward, backward and
Adam kernels the library runs on the GPU: ReLU hidden layers, softmax cross-entropy
averaged over the batch.  There is no reference counterpart (the selector is an
extension of the adaptive integrator of triaintrouts.f:1269-1464); the oracle is pinned
by central finite differences and by closed forms in tests/test_oracle_pins.py.

Parameter packing (shared with the device code): W1[64][8] b1[64] W2[64][64] b2[64]
W3[6][64] b3[6], row-major.
"""
import numpy as np

NI, NH, NO = 8, 64, 6
SIZES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NO, NH)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def forward(p, X):
    q = unpack(p)
    a1 = X @ q["W1"].T + q["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ q["W2"].T + q["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ q["W3"].T + q["b3"]
    return z, (X, a1, h1, a2, h2)


def loss_grad(p, X, y):
    """Mean cross-entropy and its gradient with respect to the packed parameters."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.int64)
    n = len(y)
    q = unpack(p)
    z, (X, a1, h1, a2, h2) = forward(p, X)
    zm = z - z.max(axis=1, keepdims=True)
    e = np.exp(zm)
    pr = e / e.sum(axis=1, keepdims=True)
    loss = float(np.mean(np.log(e.sum(axis=1)) - zm[np.arange(n), y]))
    dz = pr.copy()
    dz[np.arange(n), y] -= 1.0
    dz /= n
    g = {"W3": dz.T @ h2, "b3": dz.sum(axis=0)}
    d2 = (dz @ q["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ q["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(axis=0)
    return loss, pack(g)


def predict(p, X):
    z, _ = forward(p, np.asarray(X, np.float64))
    return np.argmax(z, axis=1)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in place; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
