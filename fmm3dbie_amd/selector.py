"""Learned refinement-depth selector for the adaptive near quadrature.

The adaptive integrator (triaadap_main, triaintrouts.f:1269-1464) refines a (target,
patch) pair's triangle tree until the children reproduce the parent's moments to epsp.
Its trees are shallow: most pairs stop after refining at most the root.  An 8 -> 64 ->
64 -> 6 MLP, trained on the GPU on the integrator's own refinement counts, predicts the
number of refined triangles m (class min(m, 5)) from eight geometric features of the
pair (distance and height over the patch, nearest node, tolerance, |k| rad, aspect).
Pairs predicted m = 0 or 1 go to the cached-tree kernels, which apply the adaptive test
to the shared depth-0..2 triangles on the matrix cores and hand failures back; the rest
run through the adaptive kernel longest-predicted first.  The weights are those of the
adaptive integrator either way; the prediction only decides where the work runs.

Workflow: ``X, y = collect(nq)`` (features and labels from a plain run), ``params, info =
train(X, y)`` (GPU training, normalisation folded into the first layer), then
``nq.run(selector=params)``.
"""
import ctypes as C

import numpy as np

from ._lib import lib
from .layer import _dptr, _torch

NCLASS = 6
NFEAT = 8
FEATURES = ("dist/rad", "height/rad", "nearest node dist/rad", "nearest node height/rad", "nearest node edge",
            "log10 epsp", "|k| rad", "rad^2/area")


def nparam():
    return int(lib().fmm3dbie_mlp_nparam())


def labels_from_nref(nref):
    """Class = number of refined triangles m = (nref - 1) / 4, capped at 5."""
    m = (np.asarray(nref, np.int64) - 1) // 4
    return np.minimum(np.maximum(m, 0), NCLASS - 1).astype(np.int32)


def init_params(seed=3):
    """He-normal weights, zero biases (float32, packed W1 b1 W2 b2 W3 b3)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((NFEAT, 64), (64, 64), (64, NCLASS)):
        parts.append((rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)).ravel())
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == nparam()
    return p


def collect(nq, stream=None):
    """Plain adaptive run of a NearQuadDevice; returns (features (nadap, 8) float32,
    labels (nadap,) int32, CSR entry index of each adaptive pair)."""
    torch = _torch()
    feat = torch.zeros((nq.nnz, NFEAT), dtype=torch.float32, device=nq.wnear.device)
    nq.run(stream=stream, want_detail=True, features=feat)
    torch.cuda.synchronize()
    idx = np.nonzero(nq.stats["cls"] == 1)[0]
    X = feat.cpu().numpy()[idx]
    y = labels_from_nref(nq.stats["nref"][idx])
    return X, y, idx


def fold_normalisation(params, mean, std):
    """First layer on raw features: W1' = W1 / std, b1' = b1 - W1' mean."""
    p = np.array(params, np.float32, copy=True)
    W1 = p[:64 * NFEAT].reshape(64, NFEAT).astype(np.float64)
    b1 = p[64 * NFEAT:64 * NFEAT + 64].astype(np.float64)
    W1n = W1 / std[None, :]
    p[:64 * NFEAT] = W1n.ravel().astype(np.float32)
    p[64 * NFEAT:64 * NFEAT + 64] = (b1 - W1n @ mean).astype(np.float32)
    return p


def grad(params, X, y, device="cuda"):
    """Device mean cross-entropy and gradient (the backward kernels), as numpy."""
    torch = _torch()
    dp = torch.as_tensor(np.asarray(params, np.float32), device=device)
    dX = torch.as_tensor(np.ascontiguousarray(X, np.float32), device=device)
    dy = torch.as_tensor(np.ascontiguousarray(y, np.int32), device=device)
    dg = torch.zeros(nparam(), dtype=torch.float32, device=device)
    loss = C.c_float(0.0)
    s = torch.cuda.current_stream()
    rc = lib().fmm3dbie_mlp_grad(_dptr(dp), _dptr(dX), _dptr(dy), len(y), _dptr(dg), C.byref(loss),
                                 C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise RuntimeError("fmm3dbie_mlp_grad failed (status %d)" % rc)
    torch.cuda.synchronize()
    return float(loss.value), dg.cpu().numpy()


def predict(params, X, device="cuda"):
    torch = _torch()
    dp = params if hasattr(params, "is_cuda") else torch.as_tensor(np.asarray(params, np.float32), device=device)
    dX = torch.as_tensor(np.ascontiguousarray(X, np.float32), device=device)
    dc = torch.zeros(len(X), dtype=torch.int32, device=device)
    s = torch.cuda.current_stream()
    rc = lib().fmm3dbie_mlp_predict(_dptr(dp), _dptr(dX), len(X), _dptr(dc), None, C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise RuntimeError("fmm3dbie_mlp_predict failed (status %d)" % rc)
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def train(X, y, n_train=200000, epochs=30, batch=256, lr=3e-3, seed=3, device="cuda"):
    """Train on (at most n_train, sampled with seed) pairs on the GPU.

    Returns (device float32 parameters acting on raw features, info dict with the
    per-epoch loss, train/held-out accuracy and training time)."""
    torch = _torch()
    X = np.asarray(X, np.float32)
    y = np.asarray(y, np.int32)
    n = len(y)
    rng = np.random.default_rng(seed)
    sel = rng.permutation(n)
    tr, ho = sel[:min(n_train, n)], sel[min(n_train, n):]
    ntr = (len(tr) // batch) * batch
    if ntr < batch:
        raise ValueError("need at least %d training pairs, got %d" % (batch, len(tr)))
    mean = X[tr].astype(np.float64).mean(axis=0)
    std = X[tr].astype(np.float64).std(axis=0)
    std[std < 1e-12] = 1.0
    Xn = ((X[tr].astype(np.float64) - mean) / std).astype(np.float32)
    dX = torch.as_tensor(np.ascontiguousarray(Xn), device=device)
    dy = torch.as_tensor(np.ascontiguousarray(y[tr]), device=device)
    dp = torch.as_tensor(init_params(seed), device=device)
    hist = np.zeros(epochs, np.float32)
    s = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(s)
    rc = lib().fmm3dbie_mlp_train(_dptr(dp), _dptr(dX), _dptr(dy), len(tr), epochs, batch, lr, seed,
                                  hist.ctypes.data_as(C.c_void_p), C.c_void_p(s.cuda_stream))
    t1.record(s)
    if rc != 0:
        raise RuntimeError("fmm3dbie_mlp_train failed (status %d)" % rc)
    torch.cuda.synchronize()
    params = fold_normalisation(dp.cpu().numpy(), mean, std)
    dpar = torch.as_tensor(params, device=device)
    info = dict(loss=hist.tolist(), ms_train=t0.elapsed_time(t1), n_train=len(tr), n_heldout=len(ho))
    info["acc_train"] = float(np.mean(predict(dpar, X[tr]) == y[tr]))
    info["acc_heldout"] = float(np.mean(predict(dpar, X[ho]) == y[ho])) if len(ho) else None
    return dpar, info
