"""The edge-sharded solve (the single-graph multi-GPU scheme, shard by shard
on one device): k shard forests, gathered, solved once more, must give the
global forest byte for byte.  On the CPU the oracle is the solver; on the
GPU every shard and the merge run through the C-ABI."""
import numpy as np
import pytest

import emst
import oracle_lib as O
import shard
from graphs import csr_from_edges, random_graph


def cpu_graphs():
    rng = np.random.default_rng(17)
    out = [emst.Graph(*csr_from_edges(n, e)) for n, e in [
        (1, []), (2, [(0, 1, 5)]), (3, [(0, 1, 1), (1, 2, 1), (0, 2, 1)]),
        (6, [(0, 1, 7), (0, 1, 3), (1, 2, 3), (2, 0, 3), (4, 5, 0)])]]
    for i in range(40):
        n = int(rng.integers(2, 120))
        out.append(emst.Graph(*random_graph(rng, n, int(rng.integers(0, 5 * n)), 0, 5, symmetric=i % 3 != 0)))
    out.append(emst.gen_grid(20, 30, 1, 1, 50))
    out.append(emst.gen_rmat(10, 8, 0.57, 0.19, 0.19, 2, 1, 40))
    return out


def test_shard_ranges_cover_the_list():
    for count in (0, 1, 7, 100):
        for k in (1, 2, 3, 8):
            r = shard.shard_ranges(count, k)
            assert len(r) == k and r[0][0] == 0 and r[-1][1] == count
            assert all(a[1] == b[0] for a, b in zip(r, r[1:]))
            assert max(hi - lo for lo, hi in r) - min(hi - lo for lo, hi in r) <= 1


def test_sub_graph_keeps_positions_in_order():
    g = emst.gen_rmat(9, 8, 0.57, 0.19, 0.19, 3, 1, 100)
    up = shard.upper_positions(g)
    pos = up[len(up) // 3: 2 * len(up) // 3]
    h = shard.sub_graph(g, pos)
    assert h.nodes == g.nodes and h.edges == len(pos)
    assert np.array_equal(h.sources(), g.sources()[pos])
    assert np.array_equal(h.nlist, g.nlist[pos]) and np.array_equal(h.eweight, g.eweight[pos])
    assert h.upper_mask().all()


@pytest.mark.parametrize("k", [1, 2, 3, 8])
def test_sharded_oracle_forest_is_global(k):
    """With the oracle's Kruskal as the shard solver (non-negative weights,
    both list shapes) and its schedule restatement (any weights)."""
    for g in cpu_graphs():
        got, info = shard.solve_sharded(g, k, O.kruskal)
        assert np.array_equal(got, O.kruskal(g))
        assert info["gathered"] <= sum(info["shard_forest_edges"])
    rng = np.random.default_rng(23)
    for i in range(20):
        n = int(rng.integers(2, 80))
        g = emst.Graph(*random_graph(rng, n, int(rng.integers(0, 4 * n)), -5, 5, symmetric=i % 2 == 0))
        sched = lambda h: O.schedule(h, False, emst.INT_MAX)[0]
        got, _ = shard.solve_sharded(g, k, sched)
        assert np.array_equal(got, sched(g))


def test_planted_mistake_is_caught():
    """A shard solver that drops one forest edge makes the sharded result
    differ from the global forest (the check is not vacuous)."""
    g = emst.gen_grid(10, 12, 2, 1, 1000)

    def lossy(h):
        m = O.kruskal(h)
        nz = np.flatnonzero(m)
        if len(nz):
            m = m.copy()
            m[nz[0]] = 0
        return m
    got, _ = shard.solve_sharded(g, 3, lossy)
    assert not np.array_equal(got, O.kruskal(g))


@pytest.fixture(scope="module")
def solver():
    s = emst.Solver(0)
    yield s
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 2, 8])
def test_sharded_gpu_forest_is_global(solver, k):
    graphs = cpu_graphs()[::4] + [
        emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24),
        emst.gen_rmat(13, 16, 0.57, 0.19, 0.19, 5, 1, 9),
        emst.gen_rgg(1 << 14, 6.0, 3, 1 << 24),
        emst.gen_grid(128, 100, 1, 1, 1 << 20),
    ]
    for i, g in enumerate(graphs):
        print(f"case {i} k={k} n={g.nodes} m={g.edges}", flush=True)
        ref, _ = solver.gpu_mst(g)
        for mode in (emst.FILTER_PLAIN, emst.FILTER_LEARNED):
            def solve(h):
                if mode == emst.FILTER_PLAIN:
                    return solver.gpu_mst(h)[0]
                return solver.gpu_mst(h, emst.FILTER_LEARNED, 0)[0]
            got, info = shard.solve_sharded(g, k, solve)
            assert np.array_equal(got, ref), (i, k, mode)
