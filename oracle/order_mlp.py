"""FP64 numpy restatement of the order network (test infrastructure: the
checker of the GPU forward, backward and decision kernels in order.hip).

Layout of the parameter block (float32 in the library): mu[8], sd[8],
W1[64][8], b1[64], W2[64][64], b2[64], W3[13][64], b3[13].  Forward:
x = (f - mu) / sd, h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2),
z = W3 h2 + b3, p = softmax(z).  Loss: mean cross-entropy.  Decision: the
smallest class whose cumulative probability reaches the confidence level
(class c = 4 + c order, 5 + c terms).  Pinned by central finite
differences of the loss (tests/test_oracle_order.py).
"""
import numpy as np

F, H, C = 8, 64, 13


def unpack(m):
    m = np.asarray(m, dtype=np.float64)
    o = 0

    def take(k, shape):
        nonlocal o
        v = m[o:o + k].reshape(shape)
        o += k
        return v
    mu = take(F, (F,))
    sd = take(F, (F,))
    w1 = take(H * F, (H, F))
    b1 = take(H, (H,))
    w2 = take(H * H, (H, H))
    b2 = take(H, (H,))
    w3 = take(C * H, (C, H))
    b3 = take(C, (C,))
    assert o == len(m)
    return mu, sd, w1, b1, w2, b2, w3, b3


def forward(m, feats):
    mu, sd, w1, b1, w2, b2, w3, b3 = unpack(m)
    x = (np.asarray(feats, dtype=np.float64) - mu) / sd
    a1 = x @ w1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3.T + b3
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    return x, h1, h2, p


def loss_grad(m, feats, labels):
    """Mean cross-entropy and its gradient (zero for mu and sd)."""
    mu, sd, w1, b1, w2, b2, w3, b3 = unpack(m)
    x, h1, h2, p = forward(m, feats)
    n = len(labels)
    y = np.asarray(labels)
    loss = -np.mean(np.log(np.maximum(p[np.arange(n), y], 1e-300)))
    dz = p.copy()
    dz[np.arange(n), y] -= 1.0
    dz /= n
    gw3 = dz.T @ h2
    gb3 = dz.sum(axis=0)
    d2 = (dz @ w3) * (h2 > 0)
    gw2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ w2) * (h1 > 0)
    gw1 = d1.T @ x
    gb1 = d1.sum(axis=0)
    g = np.concatenate([np.zeros(2 * F), gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
    return loss, g


def decide(probs, confidence):
    """Terms chosen per row: 5 + the smallest class whose cumulative
    probability reaches the confidence level (the largest class otherwise)."""
    cum = np.cumsum(probs, axis=1)
    c = np.argmax(cum >= confidence, axis=1)
    c[~(cum >= confidence).any(axis=1)] = C - 1
    return 5 + np.minimum(c, C - 1)
