"""The learned per-box order selector: training data made by the FMM itself,
model initialisation and the training driver.

The reference chooses one expansion size per solve from eps
(src/cpu/multipole3d.cpp:1328-1331); here every leaf of the adaptive octree
gets its own order from an 8 -> 64 -> 64 -> 13 network (order.hip).  Labels
come from solves of the same adaptive tree at every order p = 4 .. 16
against a reference solve at 18 terms: a leaf's label is the smallest order
at which the RMS potential error over its particles stays below
safety * tol * RMS(potential).  Meeting that in every leaf keeps the global
rel-L2 error below safety * tol, since the squared errors add up over the
leaves.  Training runs on the GPU through lfmm_order_train (hand-written
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


def make_labels(lfmm, xyz, q, leaf_max=64, tol=1e-6, safety=0.5, mixed=False, ctx=None):
    """Features and order labels of every leaf of one particle set.

    Returns (features (nleaf, 8), labels (nleaf,), info): labels are classes
    0..12 (order p = 4 + class); a leaf that misses the bar at every order
    gets the largest class.  info holds the per-order global rel-L2 errors
    against the reference solve."""
    feats, leaf = lfmm.order_features(xyz, q, leaf_max=leaf_max, ctx=ctx)
    nleaf = len(feats)
    ref, _, _ = lfmm.fmm3d(xyz, q, terms=18, leaf_max=leaf_max, adaptive=True, exclude_self=True, ctx=ctx)
    bar = safety * tol * np.sqrt(np.mean(ref * ref))
    counts = np.maximum(np.bincount(leaf, minlength=nleaf), 1)
    labels = np.full(nleaf, CLASSES - 1, dtype=np.int32)
    done = np.zeros(nleaf, dtype=bool)
    errs = {}
    for c in range(CLASSES):
        terms = MIN_TERMS + c
        p, _, _ = lfmm.fmm3d(xyz, q, terms=terms, leaf_max=leaf_max, adaptive=True, exclude_self=True, mixed=mixed,
                             ctx=ctx)
        d = p - ref
        errs[terms] = float(np.linalg.norm(d) / np.linalg.norm(ref))
        rms = np.sqrt(np.bincount(leaf, weights=d * d, minlength=nleaf) / counts)
        ok = (rms <= bar) & ~done
        labels[ok] = c
        done |= ok
    return feats, labels, {"rel_l2_by_terms": errs, "leaves": nleaf}


def train(lfmm, sets, epochs=40, batch=1024, lr=3e-3, seed=1, ctx=None):
    """Trains a model on the (features, labels) pairs of several particle
    sets; returns (model, loss per epoch)."""
    f = np.concatenate([s[0] for s in sets]).astype(np.float32)
    y = np.concatenate([s[1] for s in sets]).astype(np.int32)
    model = init_model(f, seed)
    return lfmm.order_train(model, f, y, epochs=epochs, batch=batch, lr=lr, seed=seed, ctx=ctx)
