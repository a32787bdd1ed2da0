"""CPU restatement of the threshold network (TEST INFRASTRUCTURE ONLY).

32-64-64-1 network with ReLU, weights rounded to bf16, computed in float64:
  h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 W3 + b3,
  L = mean((y - t)^2).
Parameter layout (flat): W1[32][64], b1[64], W2[64][64], b2[64], W3[64], b3.
No reference counterpart (the reference has no learned component); pinned
by central finite differences of L in tests/test_learned.py.
"""
import numpy as np

IN, H = 32, 64
W1, B1 = 0, 32 * 64
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 64
P = B3 + 1


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where((u & 0x7F800000) == 0x7F800000, u.astype(np.uint32), r)
    return r.view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(IN, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, X):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ w1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w2 + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ w3 + b3
    return y, (X, a1, h1, a2, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def grad(p, X, T):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    B = X.shape[0]
    dy = 2.0 * (y - np.asarray(T, np.float64)) / B
    g = np.zeros(P)
    g[W3:B3] = h2.T @ dy
    g[B3] = dy.sum()
    d2 = np.outer(dy, w3) * (a2 > 0)
    g[W2:B2] = (h1.T @ d2).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ w2.T) * (a1 > 0)
    g[W1:B1] = (X.T @ d1).ravel()
    g[B1:W2] = d1.sum(0)
    return g


def _hash(v):
    v = np.asarray(v, np.uint64) & 0xFFFFFFFF
    v = (((v >> 16) ^ v) * 0x45D9F3B) & 0xFFFFFFFF
    v = (((v >> 16) ^ v) * 0x45D9F3B) & 0xFFFFFFFF
    return ((v >> 16) ^ v) & 0xFFFFFFFF


def features(nindex, eweight, nodes, edges):
    """The 32 graph features and the 1024 sorted weight samples."""
    i = np.arange(1024, dtype=np.uint64)
    w = np.sort(np.asarray(eweight, np.int64)[_hash(i) % edges]) if edges > 0 else np.zeros(1024, np.int64)
    v = (_hash(np.arange(4096, dtype=np.uint64) ^ 0x5BD1E995) % nodes).astype(np.int64)
    deg = np.asarray(nindex, np.int64)[v + 1] - np.asarray(nindex, np.int64)[v]
    bins = np.zeros(13, np.int64)
    for d in deg:
        if d <= 0:
            bins[0] += 1
        else:
            bins[min(12, 1 + int(np.floor(np.log2(d))))] += 1
    f = np.zeros(32)
    f[0] = np.log2(nodes) / 32.0
    f[1] = np.log2(edges + 1.0) / 32.0
    f[2] = np.log2(edges / nodes + 1.0) / 8.0
    f[3] = bins[0] / 4096.0
    f[4:16] = bins[1:13] / 4096.0
    wmax = float(w[-1]) if w[-1] > 0 else 1.0
    f[16:32] = np.maximum(w[32 * (2 * np.arange(16) + 1)] / wmax, 0.0)
    return f, w
