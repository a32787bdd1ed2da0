"""Small graph builders for tests (numpy only)."""
import numpy as np


def csr_from_edges(n, edges, symmetric=True, sort=True):
    """edges: iterable of (u, v, w).  symmetric mirrors each (u, v, w) as
    (v, u, w).  Keeps multi-edges and self-loops as given.  Returns
    (nindex, nlist, eweight) int32 arrays with lists sorted by neighbour
    (stable), which fixes each edge's CSR position j."""
    rows = []
    for (u, v, w) in edges:
        rows.append((u, v, w))
        if symmetric and u != v:
            rows.append((v, u, w))
        elif symmetric and u == v:
            rows.append((v, u, w))
    if rows:
        a = np.array(rows, dtype=np.int64)
        order = np.lexsort((a[:, 1], a[:, 0])) if sort else np.argsort(a[:, 0], kind="stable")
        a = a[order]
        src, dst, w = a[:, 0], a[:, 1], a[:, 2]
    else:
        src = dst = w = np.zeros(0, np.int64)
    deg = np.bincount(src, minlength=n) if len(src) else np.zeros(n, np.int64)
    nindex = np.zeros(n + 1, np.int64)
    nindex[1:] = np.cumsum(deg)
    return nindex.astype(np.int32), dst.astype(np.int32), w.astype(np.int32)


def random_graph(rng, n, m_undirected, wlo=1, whi=10, symmetric=True, allow_loops=False, allow_multi=True):
    edges = []
    for _ in range(m_undirected):
        u = int(rng.integers(0, n))
        v = int(rng.integers(0, n))
        if u == v and not allow_loops:
            continue
        edges.append((u, v, int(rng.integers(wlo, whi))))
    if not allow_multi:
        seen = set()
        uniq = []
        for (u, v, w) in edges:
            k = (min(u, v), max(u, v))
            if k in seen:
                continue
            seen.add(k)
            uniq.append((u, v, w))
        edges = uniq
    return csr_from_edges(n, edges, symmetric=symmetric)
