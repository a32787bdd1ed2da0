"""The second related-work baseline on the GPU: edge-list Boruvka with label
propagation, the algorithm of Jucele_GPU/JuceleMST.cu.  Its list holds the
upper-direction CSR positions in order and ties go to the smaller position,
so its forest must equal Kruskal's byte for byte (and gpu_mst's, for
negative weights and one-way edges too)."""
import numpy as np
import pytest

import emst
import oracle_lib as O
from graphs import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = emst.Solver(0)
    yield s
    s.close()


def graphs():
    rng = np.random.default_rng(31)
    out = [("tiny", emst.Graph(*csr_from_edges(n, e))) for n, e in [
        (1, []), (2, []), (2, [(0, 1, 5)]), (3, [(0, 1, 1), (1, 2, 1), (0, 2, 1)]),
        (6, [(0, 1, 7), (0, 1, 3), (1, 2, 3), (2, 0, 3), (4, 5, 0)]),
        (5, [(0, 1, 2**31 - 1), (1, 2, 2**31 - 1), (3, 4, 0)])]]
    for i in range(8):
        n = int(rng.integers(2, 400))
        out.append((f"random{i}", emst.Graph(*random_graph(rng, n, int(rng.integers(0, 4 * n)), 0, 6))))
    out.append(("grid", emst.gen_grid(100, 80, 1, 1, 1 << 20)))
    out.append(("grid_ties", emst.gen_grid(64, 64, 2, 1, 3)))
    out.append(("rmat", emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)))
    out.append(("rmat_ties", emst.gen_rmat(13, 16, 0.57, 0.19, 0.19, 5, 1, 9)))
    out.append(("rgg", emst.gen_rgg(1 << 14, 6.0, 3, 1 << 24)))
    out.append(("er_forest", emst.gen_er(20000, 1.5, 4, 1, 1 << 20)))
    # one hub that collects most offers from the first iteration on, with
    # tied weights and parallel edges
    rng2 = np.random.default_rng(12)
    star = [(0, i, int(rng2.integers(1, 40))) for i in range(1, 40001)]
    star += [(int(a), int(b), int(rng2.integers(1, 40))) for a, b in rng2.integers(1, 40001, (30000, 2)) if a != b]
    star += [(0, i, 5) for i in range(1, 3000)]
    out.append(("star", emst.Graph(*csr_from_edges(40001, star))))
    # a path whose picks form one long chain: labels need many passes
    out.append(("path", emst.Graph(*csr_from_edges(3000, [(i, i + 1, i + 1) for i in range(2999)]))))
    return out


GRAPHS = graphs()


@pytest.mark.parametrize("name,g", GRAPHS, ids=[n for n, _ in GRAPHS])
def test_edgelist_forest_is_kruskals(solver, name, g):
    print(f"case {name} n={g.nodes} m={g.edges}", flush=True)
    got, st = solver.edgelist_mst(g)
    k = O.kruskal(g)
    assert np.array_equal(got, k)
    assert st.mst_edges == int(k.sum())
    assert st.weight == int(g.eweight[k.astype(bool)].astype(np.uint32).astype(np.int64).sum())
    assert st.iterations >= (1 if k.any() else 0)
    assert st.label_passes >= st.iterations


def test_edgelist_matches_gpu_mst_on_negative_and_one_way(solver):
    """Weights compare as unsigned 32-bit values, as in gpuMST, and only the
    upper direction is listed, so negative weights and one-way edges give
    gpu_mst's bytes (which the oracle's schedule pins)."""
    rng = np.random.default_rng(6)
    for i in range(10):
        n = int(rng.integers(2, 300))
        g = emst.Graph(*random_graph(rng, n, int(rng.integers(0, 5 * n)), -6, 6, symmetric=i % 2 == 0))
        print(f"case neg/asym {i} n={g.nodes} m={g.edges}", flush=True)
        got, _ = solver.edgelist_mst(g)
        ref, _ = O.schedule(g, False, emst.INT_MAX)
        assert np.array_equal(got, ref)
        gm, _ = solver.gpu_mst(g, False, emst.INT_MAX)
        assert np.array_equal(got, gm)


def test_edgelist_repeats_and_leaves_the_context_usable(solver):
    g = emst.gen_rmat(15, 16, 0.57, 0.19, 0.19, 3, 1, 20)
    a, sa = solver.edgelist_mst(g)
    for _ in range(2):
        b, sb = solver.edgelist_mst(g)
        assert np.array_equal(a, b) and sb.iterations == sa.iterations
    uf, T = emst.plain_threshold(g)
    c, _ = solver.gpu_mst(g, uf, T)
    assert np.array_equal(a, c)
    h = emst.gen_grid(30, 40, 1, 1, 1000)
    d, _ = solver.edgelist_mst(h)
    assert np.array_equal(d, O.kruskal(h))
