"""Edge-sharded minimum spanning forest: the single-graph multi-GPU scheme,
run shard by shard on one device.

The upper-direction CSR positions (nlist[j] > source, the only positions
gpuMST ever marks, ECL_MST.cu:97, 170, 195) are cut into k contiguous ranges
of equal size, one per shard.  Each shard is solved alone; the positions its
forest marks are gathered, and one more solve over the gathered positions
gives the global forest.

Why the result is bit-identical to one solve over the whole graph: every
solve orders edges by the key (weight, j) of ECL_MST.cu:221, and a shard's
CSR lists its positions in increasing j, so the shard's own positions order
its edges exactly as the global j does.  An edge that a shard's forest drops
is the heaviest edge of a cycle inside that shard (cycle property), so it is
not in the global forest either; the global forest is therefore inside the
gathered set, and it is that set's forest.

The shard and merge graphs keep all n vertices and hold upper-direction
positions only (one-way lists), which gpuMST accepts.  On k devices each
shard would be one rank's input and the gather an all-gather or a gather to
one rank; here the shards run in turn on the context's device.
"""
import numpy as np

import emst


def upper_positions(g):
    """Increasing CSR positions j with nlist[j] > source(j)."""
    return np.flatnonzero(g.upper_mask()).astype(np.int64)


def sub_graph(g, pos, sources=None):
    """The graph on g's vertices holding only the positions pos (increasing),
    in that order, with their weights (sources: g.sources(), if at hand)."""
    src = (g.sources() if sources is None else sources)[pos]
    counts = np.bincount(src, minlength=g.nodes)
    nindex = np.zeros(g.nodes + 1, np.int64)
    np.cumsum(counts, out=nindex[1:])
    return emst.Graph(nindex.astype(np.int32), g.nlist[pos], g.eweight[pos])


def shard_ranges(count, k):
    """k contiguous, near-equal ranges of [0, count)."""
    cuts = [count * s // k for s in range(k + 1)]
    return [(cuts[s], cuts[s + 1]) for s in range(k)]


def solve_sharded(g, k, solve):
    """Global inMST of g from k shard solves and one merge solve.

    solve(graph) -> uint8 inMST of that graph (any exact MSF solver that
    orders by (weight, position)).  Returns (inMST, info) where info holds
    the shard sizes, the gathered count and the per-shard forest sizes."""
    if k < 1:
        raise ValueError("k must be at least 1")
    up = upper_positions(g)
    src = g.sources()
    kept, sizes, forests = [], [], []
    for lo, hi in shard_ranges(len(up), k):
        pos = up[lo:hi]
        mark = solve(sub_graph(g, pos, src))
        kept.append(pos[mark.astype(bool)])
        sizes.append(int(hi - lo))
        forests.append(int(mark.sum()))
    gathered = np.concatenate(kept) if kept else np.zeros(0, np.int64)
    gathered.sort()
    mark = solve(sub_graph(g, gathered, src))
    out = np.zeros(g.edges, np.uint8)
    out[gathered[mark.astype(bool)]] = 1
    return out, {"k": k, "shard_edges": sizes, "shard_forest_edges": forests, "gathered": int(len(gathered))}
