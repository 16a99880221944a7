"""Learned order selection for the adaptive near quadrature.

The adaptive integrator (triaadap_main, triaintrouts.f:1269-1464) refines a (target,
patch) pair's triangle tree until the children reproduce the parent's moments to epsp,
with one XG rule on every triangle (degree 16 at eps = 1e-6, 55 nodes).  Which rule is
cheapest depends on the pair: a smooth integrand is done with a high degree and no
refinement, a nearly singular one with a lower degree on a deeper tree; measured on the
wavy torus, the cheapest of six degrees per pair needs 0.65 of the kernel evaluations of
the fixed degree.  An 8 -> 64 -> 64 -> 6 MLP, trained on the GPU, predicts that rule
class from eight geometric features of the pair (distance and height over the patch,
nearest node, tolerance, |k| rad, aspect).  Class c runs the same adaptive integrator
with XG degree q0 + (-4, -2, 0, +4, +8, +12)[c] (q0 the reference's degree; class 2 is
the plain rule) and a level cap learned with it; the integrator's own test (children
against parent, one level up) judges every triangle, and a pair whose tree needs a
level past its class's cap is integrated again by the plain integrator.

Labels come from the integrator itself: every adaptive pair is integrated once per class
(forced routing, no cap) and labelled with the class of least cost, the cost being the
pair's own work in 64-node chunks (on the cached tree, and below it where the geometry
is computed) weighted by unit costs fitted to the measured launch times of the six
classes; the class's level cap is the deepest tree among the training pairs labelled
with it, and its tree cache holds no depth below that cap.

Workflow: ``X, y, idx, info = collect(nq)``, ``params, tinfo = train(X, y)`` (GPU
training, normalisation folded into the first layer), ``lmax = level_caps(y, info)``,
then ``nq.run(selector=params, route_lmax=lmax)``.
"""
import ctypes as C

import numpy as np

from ._lib import lib
from . import basis
from .layer import _dptr, _torch

NCLASS = 6
NFEAT = 8
FEATURES = ("dist/rad", "height/rad", "nearest node dist/rad", "nearest node height/rad", "nearest node edge",
            "log10 epsp", "|k| rad", "rad^2/area")
DEGREE_OFFSETS = (-4, -2, 0, 4, 8, 12)
PLAIN_CLASS = 2


def nparam():
    return int(lib().fmm3dbie_mlp_nparam())


def class_degree(q0, c):
    """XG degree of rule class c around the plain degree q0 (as the library: not below
    degree 8, where the children-against-parent test stops being a reliable estimate,
    and not above 30, the largest rule in the tables)."""
    q = q0 + DEGREE_OFFSETS[c]
    return min(30, q) if q >= 8 else min(q0, 8)


def xg_size(deg):
    return len(basis.xg_rule(deg)[1])


def plain_degree(kq):
    """The XG degree with kq nodes (the plain rule of a generation)."""
    for d in range(1, 31):
        if xg_size(d) == kq:
            return d
    raise ValueError("no XG rule with %d nodes" % kq)


def forced(c):
    """Parameters whose argmax is class c for every input (all weights zero)."""
    p = np.zeros(nparam(), np.float32)
    p[-NCLASS + c] = 1.0
    return p


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


def chunk_rows(norder):
    """Nodes per chunk of the adaptive kernel: 55 at order 6 and above (the LDS of 12
    resident wavefronts per CU), else 64 (quad.hip chunk_rows)."""
    return 55 if norder >= 6 else 64


def per_triangle(norder, kq):
    """Whether the adaptive kernel takes one chunk per triangle for a rule of kq nodes
    (order 6 and above, a triangle filling more than half a chunk: quad.hip step_tris)."""
    cr = chunk_rows(norder)
    return norder >= 6 and cr < 2 * kq <= 2 * cr


def chunks(norder, kq):
    """Chunks of the adaptive kernel's root step and of a refinement test with a rule of
    kq nodes: one per triangle (per_triangle), else the nodes of the step in equal
    chunks."""
    if per_triangle(norder, kq):
        return 1, 4
    cr = chunk_rows(norder)
    return -(-kq // cr), -(-4 * kq // cr)


def collect(nq, stream=None):
    """Features of every adaptive pair (from a plain run) and its label: the rule class
    whose adaptive integration takes the least time.  Every pair is integrated once
    with each class (forced routing, no level cap, every kernel in order on one stream).
    The kernel works in chunks of at most chunk_rows(norder) nodes (one node per lane;
    chunks() counts them for the root and a refinement test), and a
    chunk costs the same whether its lanes are all busy or not, so the measured time of
    each class's launch is fitted by a cost per chunk on the cached tree (one for the
    classes with a chunk per triangle, one for the others: the two code paths), a
    cost per chunk below it (geometry evaluated) and a cost per pair, summed over the
    pairs' own counts; the label is the class of least fitted chunk cost among the
    classes worth their tree cache (open_classes).  Returns (X
    (nadap, 8) float32, y (nadap,) int32, CSR entry index of each adaptive pair, info
    with the per-class fitted cost 'cost', kernel evaluations 'evals', deepest levels
    'lev' (6, nadap), launch times 'ms', the fitted unit costs 'unit' (ms per cached
    chunk of the per-triangle classes and of the others, per computed chunk, per pair),
    node counts 'kq', the plain degree 'q0', the
    classes used 'used' and the plain rule's tree-cache time for depths 0..2 'ms_tree')."""
    from scipy.optimize import nnls
    from . import layer
    torch = _torch()
    dev = nq.wnear.device
    feat = torch.zeros((nq.nnz, NFEAT), dtype=torch.float32, device=dev)
    side = layer.side_stream_enabled()
    layer.set_side_stream(False)
    try:
        nq.run(stream=stream, want_detail=True, features=feat)
        torch.cuda.synchronize()
        ms_tree = nq.stats["ms_tree"] * 21.0 / max(nq.stats["cache_tris"], 1)  # per depth-0..2 cache
        idx = np.nonzero(nq.stats["cls"] == 1)[0]
        X = feat.cpu().numpy()[idx]
        q0 = plain_degree(nq.stats["kq"])
        kq = np.array([xg_size(class_degree(q0, c)) for c in range(NCLASS)])
        ncached = np.zeros((NCLASS, len(idx)))
        ncomp = np.zeros((NCLASS, len(idx)))
        ccached = np.zeros((NCLASS, len(idx)))
        ccomp = np.zeros((NCLASS, len(idx)))
        lev = np.zeros((NCLASS, len(idx)), np.int32)
        ms = np.zeros(NCLASS)
        for c in range(NCLASS):
            sel = torch.as_tensor(forced(c), device=dev)
            for rep in range(2):  # the second run is timed (tables and caches built)
                nq.run(stream=stream, want_detail=True, selector=sel, route_lmax=[20] * NCLASS)
            st = nq.stats
            nref = st["nref"][idx].astype(np.float64)
            unc = st["unc"][idx].astype(np.float64)
            ncached[c] = kq[c] * (1 + 4 * (nref - unc))
            ncomp[c] = kq[c] * 4 * unc
            croot, cstep = chunks(nq.norder, kq[c])
            ccached[c] = croot + cstep * (nref - unc)
            ccomp[c] = cstep * unc
            lev[c] = st["lev"][idx]
            ms[c] = st["ms_adap"]
    finally:
        layer.set_side_stream(side)
    # a cost per pair (queue, staging, weights: the same for every class) keeps the chunk
    # costs from absorbing it
    tri = np.array([per_triangle(nq.norder, int(k)) for k in kq])
    A = np.stack([np.where(tri, ccached.sum(1), 0.0), np.where(tri, 0.0, ccached.sum(1)), ccomp.sum(1),
                  np.full(NCLASS, float(len(idx)))], 1)
    unit, _ = nnls(A, ms)
    cost = np.where(tri, unit[0], unit[1])[:, None] * ccached + unit[2] * ccomp
    y, used = open_classes(cost, tree_cost(ms_tree, kq, q0))
    return X, y, idx, dict(cost=cost, evals=ncached + ncomp, lev=lev, ms=ms, unit=unit, kq=kq, q0=q0,
                           used=used, ms_tree=ms_tree)


def tree_cost(ms_tree, kq, q0):
    """Time of the tree cache of each routed rule class (depths 0..2, quad.hip), scaled
    from the plain rule's (ms_tree, for depths 0..2) by the node count; 0 for the plain
    rule, whose cache every generation builds."""
    kq = np.asarray(kq, np.float64)
    c = ms_tree * kq / kq[PLAIN_CLASS]
    c[[class_degree(q0, k) == q0 for k in range(NCLASS)]] = 0.0
    return c


def open_classes(cost, cache):
    """Which rule classes to use and each pair's label.  A class other than the plain one
    costs the build of its tree cache once per generation (cache[c], same unit as cost);
    starting from the plain class alone, the class whose pairs would save the most
    (sum over pairs of the drop in their least cost, minus its cache) is added while that
    saving is positive.  Returns (labels: the least-cost used class of every pair, used
    classes as a boolean array)."""
    ncls = cost.shape[0]
    used = np.zeros(ncls, bool)
    used[[c for c in range(ncls) if cache[c] == 0.0] or [PLAIN_CLASS]] = True
    best = cost[used].min(axis=0)
    while True:
        gain = np.array([np.maximum(best - cost[c], 0.0).sum() - cache[c] if not used[c] else -np.inf
                         for c in range(ncls)])
        c = int(np.argmax(gain))
        if not np.isfinite(gain[c]) or gain[c] <= 0.0:
            break
        used[c] = True
        best = np.minimum(best, cost[c])
    masked = np.where(used[:, None], cost, np.inf)
    return np.argmin(masked, axis=0).astype(np.int32), used


def level_caps(y, info):
    """Level cap of each class: the deepest tree among the pairs labelled with it (20,
    the integrator's own cap, for a class without pairs)."""
    caps = np.full(NCLASS, 20, np.int32)
    for c in range(NCLASS):
        m = y == c
        if m.any():
            caps[c] = max(1, int(info["lev"][c][m].max()))
    return caps


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
