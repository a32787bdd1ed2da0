"""Batched solving: every graph of a packed batch must get exactly the
inMST it gets alone (the oracle's Kruskal), for every filter mode, across
several arenas, and across repeated hipGraph replays interleaved with other
work on the same context."""
import os

import numpy as np
import pytest

import emst
import oracle_lib as O
from graphs import random_graph

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp.f32")


def mixed_graphs(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = i % 5
        s = int(rng.integers(1, 1 << 30))
        n = int(2 ** rng.uniform(5, 13))
        if k == 0:
            r = max(1, int(np.sqrt(n)))
            out.append(emst.gen_grid(r, max(1, n // r), s, 1, 1 << 20))
        elif k == 1:
            out.append(emst.gen_rgg(n, 6.0, s, 1 << 20))
        elif k == 2:
            out.append(emst.gen_rmat(max(2, int(np.log2(n))), 8, 0.57, 0.19, 0.19, s, 1, 1 << 20))
        elif k == 3:
            out.append(emst.gen_er(n, 10.0, s, 1, 1 << 20))
        else:
            out.append(emst.Graph(*random_graph(rng, int(rng.integers(1, 40)), int(rng.integers(0, 100)), 0, 5)))
    return out


@pytest.fixture(scope="module")
def setup():
    s = emst.Solver(0)
    graphs = mixed_graphs(60, 4)
    refs = [O.kruskal(g) for g in graphs]
    yield s, graphs, refs
    s.close()


def check_all(b, graphs, refs):
    for i, g in enumerate(graphs):
        assert np.array_equal(b.fetch(i), refs[i]), i


def test_batch_modes_exact(setup):
    s, graphs, refs = setup
    b = emst.Batch(s, graphs)
    for mode in (emst.FILTER_OFF, emst.FILTER_PLAIN):
        st = b.solve(mode, use_graph=False)
        check_all(b, graphs, refs)
        assert st.graphs == len(graphs) and st.chunks == 1
    # plain mode applies the plain rule per graph
    t, f = b.thresholds()
    for i, g in enumerate(graphs):
        uf, T = emst.plain_threshold(g)
        assert f[i] == int(uf) and (not uf or t[i] == T)
    # explicit per-graph thresholds (any value keeps the result exact)
    rng = np.random.default_rng(1)
    thr = rng.integers(0, 1 << 20, len(graphs)).astype(np.int32)
    b.solve(emst.FILTER_PLAIN, thresholds=thr, use_graph=False)
    check_all(b, graphs, refs)
    if os.path.exists(MODEL):
        s.set_filter_model(np.fromfile(MODEL, np.float32))
    b.solve(emst.FILTER_LEARNED, use_graph=False)
    check_all(b, graphs, refs)
    s.set_filter_model(None)
    b.close()


def test_batch_many_arenas(setup):
    s, graphs, refs = setup
    b = emst.Batch(s, graphs, max_arena_edges=20000)
    chunks, n, m = b.info()
    assert chunks > 3 and n == sum(g.nodes for g in graphs) and m == sum(g.edges for g in graphs)
    st = b.solve(emst.FILTER_PLAIN, use_graph=False)
    assert st.chunks == chunks
    check_all(b, graphs, refs)
    b.close()


def test_batch_graph_replays(setup):
    s, graphs, refs = setup
    b = emst.Batch(s, graphs, max_arena_edges=50000)
    st = b.solve(emst.FILTER_PLAIN, use_graph=True)     # dynamic solve + capture
    assert st.replays == 0
    check_all(b, graphs, refs)
    eager = emst.gen_grid(20, 20, 7)
    eager_ref = O.kruskal(eager)
    for rep in range(4):
        st = b.solve(emst.FILTER_PLAIN, use_graph=True)
        assert st.replays == st.chunks, rep
        check_all(b, graphs, refs)
        # eager work on the same context between replays
        got, _ = s.gpu_mst(eager, emst.FILTER_OFF, emst.INT_MAX)
        assert np.array_equal(got, eager_ref)
        b.solve(emst.FILTER_OFF, use_graph=False)
        check_all(b, graphs, refs)
    b.close()


def test_batch_stale_capture_falls_back(setup):
    """A capture holds the round counts of the solve it was taken from.  A
    replay whose thresholds need more rounds than were captured leaves a
    phase unfinished; its posted count is then non-zero, and the arena is
    solved again dynamically and recaptured.  Captured with every edge light
    (threshold 2^30), replayed with every edge heavy (threshold 0)."""
    s, graphs, refs = setup
    b = emst.Batch(s, graphs, max_arena_edges=50000)
    all_light = np.full(len(graphs), 1 << 30, np.int32)
    all_heavy = np.zeros(len(graphs), np.int32)
    st = b.solve(emst.FILTER_PLAIN, thresholds=all_light, use_graph=True)
    assert st.replays == 0 and st.heavy == 0
    check_all(b, graphs, refs)
    st = b.solve(emst.FILTER_PLAIN, thresholds=all_heavy, use_graph=True)
    assert st.replays < st.chunks      # at least one arena's capture was stale
    check_all(b, graphs, refs)
    st = b.solve(emst.FILTER_PLAIN, thresholds=all_heavy, use_graph=True)
    assert st.replays == st.chunks     # recaptured: every arena replays
    check_all(b, graphs, refs)
    b.close()


def test_batch_learned_graph_replay(setup):
    """Learned per-graph thresholds with capture, then a replay."""
    s, graphs, refs = setup
    b = emst.Batch(s, graphs, max_arena_edges=50000)
    b.set_filter_model(np.fromfile(os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp_union.f32"), np.float32))
    st = b.solve(emst.FILTER_LEARNED, use_graph=True)
    assert st.replays == 0
    check_all(b, graphs, refs)
    st = b.solve(emst.FILTER_LEARNED, use_graph=True)
    assert st.replays == st.chunks
    check_all(b, graphs, refs)
    b.close()
