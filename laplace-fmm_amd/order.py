"""The learned per-box order selector: training data made by the FMM itself,
model initialisation and the training driver.

The reference chooses one expansion size per solve from eps
(src/cpu/multipole3d.cpp:1328-1331); here every leaf of the adaptive octree
gets its own order from an 8 -> 64 -> 64 -> 13 network (order.hip).  Labels
come from solves of the same adaptive tree at every order p = 4 .. 16
against a reference solve at 18 terms: the global error budget (a rel-L2
of safety * tol) is spread over the leaves where it saves the most M2L work
(see make_labels).  Training runs on the GPU through lfmm_order_train (hand-written
forward, backward and Adam kernels).
"""
import numpy as np

MIN_TERMS = 5
CLASSES = 13
HIDDEN = 64
FEATURES = 8


def init_model(features, seed=0):
    """Parameter block for the order network: feature normalisation from the
    training rows, He-initialised weights, zero biases."""
    f = np.asarray(features, dtype=np.float64)
    mu = f.mean(axis=0)
    sd = f.std(axis=0)
    sd[sd < 1e-6] = 1.0
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / FEATURES), (HIDDEN, FEATURES))
    w2 = rng.normal(0.0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, np.sqrt(1.0 / HIDDEN), (CLASSES, HIDDEN))
    parts = [mu, sd, w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3.ravel(), np.zeros(CLASSES)]
    return np.concatenate(parts).astype(np.float32)


def m2l_flops_per_op(terms):
    """Algorithmic FP64 flops of one rotation-based M2L with `terms` terms."""
    nx2 = terms * (terms + 1) // 2
    rot = sum((n + 1) * (4 * n + 2) for n in range(terms))
    return 2 * (6 * nx2 + rot) + 4 * sum(t * t for t in range(1, terms + 1))


def allocate(err2, cost, budget):
    """Orders (classes) minimising total cost with total error within budget.

    err2[c, leaf]: squared potential error summed over the leaf's particles at
    class c; cost[c, leaf]: its M2L cost.  For a multiplier lam every leaf
    takes argmin_c cost + lam * err2 (the Lagrangian of the budget
    constraint); lam is found by bisection in log space as the smallest value
    whose total error fits.  Returns (classes, total error)."""
    def pick(lam):
        c = np.argmin(cost + lam * err2, axis=0)
        return c, float(err2[c, np.arange(err2.shape[1])].sum())
    lo, hi = -60.0, 60.0
    c, e = pick(10.0 ** hi)
    if e > budget:  # not reachable: the most accurate class everywhere
        return c, e
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        c, e = pick(10.0 ** mid)
        if e <= budget:
            hi = mid
        else:
            lo = mid
    return pick(10.0 ** hi)


def make_labels(lfmm, xyz, q, leaf_max=64, tol=1e-6, safety=0.5, mixed=False, ctx=None):
    """Features and order labels of every leaf of one particle set.

    The FMM itself makes the data: solves of the same adaptive tree at every
    order p = 4 .. 16 against a reference at 18 terms give each leaf's summed
    squared potential error per order.  The error budget
    (safety * tol)^2 * sum(potential^2) -- a global rel-L2 of safety * tol --
    is then spent where it saves the most M2L work: each leaf's order
    minimises flops(order) * (1 + its V-list size) + lam * error, with lam
    set by bisection so the errors add up to the budget.

    Returns (features (nleaf, 8), labels (nleaf,) = classes 0..12, info)."""
    feats, leaf = lfmm.order_features(xyz, q, leaf_max=leaf_max, ctx=ctx)
    nleaf = len(feats)
    ref, _, _ = lfmm.fmm3d(xyz, q, terms=18, leaf_max=leaf_max, adaptive=True, exclude_self=True, ctx=ctx)
    err2 = np.zeros((CLASSES, nleaf))
    errs = {}
    for c in range(CLASSES):
        terms = MIN_TERMS + c
        p, _, _ = lfmm.fmm3d(xyz, q, terms=terms, leaf_max=leaf_max, adaptive=True, exclude_self=True, mixed=mixed,
                             ctx=ctx)
        d = p - ref
        errs[terms] = float(np.linalg.norm(d) / np.linalg.norm(ref))
        err2[c] = np.bincount(leaf, weights=d * d, minlength=nleaf)
    nv = np.rint(feats[:, 4].astype(np.float64) * 189.0)
    cost = np.array([m2l_flops_per_op(MIN_TERMS + c) for c in range(CLASSES)], dtype=np.float64)[:, None] * (1.0 + nv)
    budget = (safety * tol) ** 2 * float(np.sum(ref * ref))
    labels, total = allocate(err2, cost, budget)
    info = {"rel_l2_by_terms": errs, "leaves": nleaf, "planned_rel_l2": float(np.sqrt(total / np.sum(ref * ref))),
            "planned_flops": float(cost[labels, np.arange(nleaf)].sum())}
    return feats, labels.astype(np.int32), info


def train(lfmm, sets, epochs=40, batch=1024, lr=3e-3, seed=1, ctx=None):
    """Trains a model on the (features, labels) pairs of several particle
    sets; returns (model, loss per epoch)."""
    f = np.concatenate([s[0] for s in sets]).astype(np.float32)
    y = np.concatenate([s[1] for s in sets]).astype(np.int32)
    model = init_model(f, seed)
    return lfmm.order_train(model, f, y, epochs=epochs, batch=batch, lr=lr, seed=seed, ctx=ctx)
