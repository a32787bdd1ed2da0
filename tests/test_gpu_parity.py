"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

The oracle's serial Kruskal (restating ECL_MST.cu:82-124) fixes the exact
inMST set for non-negative weights; the oracle's round-schedule emulator
(restating ECL_MST.cu:127-340) fixes it for any weights and fixes the
per-phase worklist sizes and round counts."""
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


def G(ni, nl, ew):
    return emst.Graph(ni, nl, ew)


def with_flags(solver, extra, fn):
    """fn() under the solver's current configuration plus flag bits extra."""
    saved = solver.config().as_dict()
    solver.set_config(**{**saved, "flags": saved["flags"] | extra})
    try:
        return fn()
    finally:
        solver.set_config(**saved)


def check(solver, g, use_filter=None, T=None, rounds=True, kruskal=True):
    """The solve under the current configuration against the oracle: bytes,
    L, H and round counts exactly; worklist sizes exactly with parallel
    records kept (the reference schedule), and no larger with them
    eliminated (the default)."""
    if use_filter is None:
        use_filter, T = emst.plain_threshold(g)
    print(f"case n={g.nodes} m={g.edges} filter={use_filter} T={T}", flush=True)
    got, st = solver.gpu_mst(g, use_filter, T)
    ref, sch = O.schedule(g, use_filter, T)
    assert np.array_equal(got, ref), "inMST differs from the round-schedule oracle"
    if kruskal:
        k = O.kruskal(g)
        assert np.array_equal(got, k), "inMST differs from Kruskal"
        rc, cw, gw, oc, og = O.verify(g, k, got)
        assert rc == 0 and cw == gw
    assert st.light == sch.light
    if use_filter:
        assert st.heavy == sch.heavy
    if rounds:
        assert st.light_rounds == sch.light_rounds
        assert st.heavy_rounds == (sch.heavy_rounds if use_filter else 0)
        sin = sum(sch.light_w[i] for i in range(sch.light_rounds)) + sum(sch.heavy_w[i] for i in range(sch.heavy_rounds))
        sout = sum(sch.light_s[i] for i in range(sch.light_rounds)) + sum(sch.heavy_s[i] for i in range(sch.heavy_rounds))
        got2, st2 = with_flags(solver, emst.F_NO_PAIR_DEDUP, lambda: solver.gpu_mst(g, use_filter, T))
        assert np.array_equal(got2, got)
        assert (st2.light_rounds, st2.heavy_rounds, st2.light, st2.heavy) == \
            (st.light_rounds, st.heavy_rounds, st.light, st.heavy)
        assert (st2.round_in_sum, st2.round_out_sum) == (sin, sout)
        assert st.round_in_sum <= sin and st.round_out_sum <= sout
    assert st.mst_edges == sch.mst_edges
    return got, st


def test_tiny_fixtures(solver):
    cases = [
        (1, []),
        (2, [(0, 1, 5)]),
        (3, [(0, 1, 1), (1, 2, 1), (0, 2, 1)]),          # all ties: lowest j wins
        (4, [(0, 1, 3), (2, 3, 3)]),                      # forest
        (5, [(0, 1, 2), (1, 2, 2), (2, 3, 2), (3, 4, 2), (0, 4, 1)]),
        (6, [(0, 1, 7), (0, 1, 3), (1, 2, 3), (2, 0, 3), (4, 5, 0)]),  # multi-edge, zero weight, isolated 3
    ]
    for n, edges in cases:
        g = G(*csr_from_edges(n, edges))
        for uf, T in ((False, emst.INT_MAX), (True, 2), (True, 0), (True, -1)):
            check(solver, g, uf, T)


def test_random_small(solver):
    rng = np.random.default_rng(7)
    for it in range(60):
        n = int(rng.integers(1, 40))
        g = G(*random_graph(rng, n, int(rng.integers(0, 120)), 0, int(rng.integers(1, 6))))
        for uf, T in ((False, emst.INT_MAX), (True, int(rng.integers(0, 6)))):
            check(solver, g, uf, T)


def test_grid_512(solver):
    g = emst.gen_grid(512, 512, 1, 1, 1 << 20)
    got, st = check(solver, g)
    assert got.sum() == 512 * 512 - 1


@pytest.mark.parametrize("scale", [12, 16])
def test_rmat(solver, scale):
    g = emst.gen_rmat(scale, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)
    check(solver, g)
    check(solver, g, False, emst.INT_MAX)


def test_rmat_heavy_ties(solver):
    g = emst.gen_rmat(15, 16, 0.57, 0.19, 0.19, 5, 1, 65)
    check(solver, g)


def test_rgg_and_er(solver):
    check(solver, emst.gen_rgg(1 << 16, 6.0, 3, 1 << 24))
    check(solver, emst.gen_er(20000, 10.0, 4, 1, 1 << 20))


def test_zero_degree_runs(solver):
    """Initialisation tasks over short and long vertex ranges: a dense block
    (a task spans few vertices), then members spread among long runs of
    zero-degree vertices (a task spans thousands), and trailing isolated
    vertices; both source-vertex searches of the init kernels run."""
    rng = np.random.default_rng(11)
    n = 60000
    dense = list(range(0, 300))
    sparse = sorted(rng.choice(np.arange(300, n - 5000), 700, replace=False).tolist())
    edges = []
    for _ in range(6000):
        u, v = rng.choice(dense, 2)
        if u != v:
            edges.append((int(u), int(v), int(rng.integers(1, 1000))))
    for _ in range(4000):
        u, v = rng.choice(sparse, 2)
        if u != v:
            edges.append((int(u), int(v), int(rng.integers(1, 1000))))
    for i in range(0, len(sparse), 7):   # a few links between the regions
        edges.append((int(rng.choice(dense)), int(sparse[i]), int(rng.integers(1, 1000))))
    g = G(*csr_from_edges(n, edges))
    for tile in (256, 512, 1024):
        solver.set_config(tile_edges=tile)
        check(solver, g)
        check(solver, g, False, emst.INT_MAX)
    solver.set_config()


def test_negative_weights_follow_gpu_order(solver):
    # The reference GPU orders keys as unsigned (ECL_MST.cu:221), so a
    # negative weight sorts above every non-negative one there; the
    # schedule oracle models that, Kruskal (signed) does not.
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(2, 30))
        g = G(*random_graph(rng, n, 60, -5, 6))
        for uf, T in ((False, emst.INT_MAX), (True, 0), (True, -3)):
            check(solver, g, uf, T, kruskal=False)


def test_asymmetric_input(solver):
    rng = np.random.default_rng(11)
    for _ in range(20):
        n = int(rng.integers(2, 50))
        g = G(*random_graph(rng, n, 150, 1, 9, symmetric=False))
        check(solver, g, False, emst.INT_MAX)
        check(solver, g, True, 4)


def test_modes_agree(solver):
    g = emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 9, 1, 1 << 10)
    uf, T = emst.plain_threshold(g)
    base, _ = solver.gpu_mst(g, uf, T)
    for kw in (dict(flags=emst.F_HOST_SYNC), dict(flags=emst.F_NO_GUARD), dict(flags=emst.F_NO_COMBINE),
               dict(flags=emst.F_NO_PAIR_DEDUP), dict(flags=emst.F_HOST_SYNC | emst.F_NO_PAIR_DEDUP),
               dict(flags=emst.F_NO_GUARD | emst.F_NO_COMBINE), dict(tail_max=-1),
               dict(tail_max=1 << 20), dict(stage_recs=64), dict(grid_per_cu=1)):
        print("mode", kw, flush=True)
        solver.set_config(**kw)
        try:
            got, _ = solver.gpu_mst(g, uf, T)
        finally:
            solver.set_config()
        assert np.array_equal(got, base), kw


def test_repeat_and_roots(solver):
    g = emst.gen_rmat(14, 8, 0.57, 0.19, 0.19, 12, 1, 100)
    uf, T = emst.plain_threshold(g)
    # in the caller's numbering (a skewed graph is renumbered by default)
    with_flags(solver, emst.F_NO_RELABEL, lambda: solver.load(g))
    outs = []
    for _ in range(3):
        assert solver.solve(uf, T).relabelled == 0
        outs.append(solver.fetch())
    assert all(np.array_equal(outs[0], o) for o in outs)
    parent = solver.fetch_parent(g.nodes)
    # root of every vertex is the smallest id of its tree
    roots = parent.copy()
    for _ in range(64):
        roots = roots[roots]
    src = g.sources()
    mst = outs[0].astype(bool)
    comp = np.arange(g.nodes)
    # union-find over MST edges on the host
    p = np.arange(g.nodes)

    def f(x):
        while p[x] != x:
            p[x] = p[p[x]]
            x = p[x]
        return x
    for u, v in zip(src[mst], g.nlist[mst]):
        a, b = f(u), f(v)
        if a != b:
            p[max(a, b)] = min(a, b)
    comp = np.array([f(x) for x in range(g.nodes)])
    assert np.array_equal(roots, comp)
    # renumbered by degree: the same forest; the parent array comes back in
    # the caller's ids, each tree's root is one of its members (the one of
    # largest degree) and a fixed point
    solver.load(g)
    st = solver.solve(uf, T)
    assert st.relabelled == 1
    assert np.array_equal(solver.fetch(), outs[0])
    parent = solver.fetch_parent(g.nodes)
    r2 = parent.copy()
    for _ in range(64):
        r2 = r2[r2]
    assert np.array_equal(r2[r2], r2)
    assert np.array_equal(comp[r2], comp)       # each vertex's root lies in its own tree
    assert len(np.unique(r2)) == len(np.unique(comp))  # one root per tree
    deg = np.diff(g.nindex)
    for c in np.unique(comp)[:50]:
        members = np.flatnonzero(comp == c)
        assert deg[r2[members[0]]] == deg[members].max()


def test_solve_device_pointers(solver):
    """emst_solve_device on caller-owned device memory.  The buffers come
    from the same HIP runtime libemst.so links (/opt/rocm), through ctypes;
    torch ships its own HIP runtime and cannot share the process."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    H2D, D2H = 1, 2
    g = emst.gen_grid(64, 80, 3, 1, 100)
    bufs = []

    def dev(arr):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        bufs.append(p)
        assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p.value
    try:
        out = np.full(g.edges, 7, np.uint8)
        d_out = dev(out)
        solver.solve_device(g.nodes, g.edges, dev(g.nindex), dev(g.nlist), dev(g.eweight), False, emst.INT_MAX,
                            d_out)
        got = np.zeros(g.edges, np.uint8)
        assert hip.hipMemcpy(got.ctypes.data, C.c_void_p(d_out), g.edges, D2H) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
    assert np.array_equal(got, O.kruskal(g))


def test_shapes_and_autotune_exact(solver):
    g = emst.gen_rmat(15, 16, 0.57, 0.19, 0.19, 3, 1, 1 << 12)
    uf, T = emst.plain_threshold(g)
    ref = O.kruskal(g)
    for kw in (dict(block=256), dict(block=1024), dict(tile_edges=256), dict(tile_edges=1024),
               dict(stage_recs=128, init_grid_per_cu=2), dict(block=1024, tile_edges=1024, stage_recs=1024)):
        solver.set_config(**kw)
        got, _ = solver.gpu_mst(g, uf, T)
        assert np.array_equal(got, ref), kw
    solver.set_config()
    solver.load(g)
    cfg, ms = solver.autotune(uf, T, reps=1)
    assert ms > 0 and cfg["block"] in (256, 512, 1024)
    solver.solve(uf, T)
    assert np.array_equal(solver.fetch(), ref)
    solver.set_config()


def test_pair_elimination_shrinks_worklists(solver):
    """After the first round most records of a skewed graph join the same
    few components; keeping only the lightest per root pair and tile leaves
    the bytes and round counts alone and shrinks the worklists."""
    g = emst.gen_rmat(16, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)
    uf, T = emst.plain_threshold(g)
    got, st = solver.gpu_mst(g, uf, T)
    ref, st_ref = with_flags(solver, emst.F_NO_PAIR_DEDUP, lambda: solver.gpu_mst(g, uf, T))
    assert np.array_equal(got, ref) and np.array_equal(got, O.kruskal(g))
    assert (st.light_rounds, st.heavy_rounds) == (st_ref.light_rounds, st_ref.heavy_rounds)
    print("in/out sums", (st.round_in_sum, st.round_out_sum), "reference", (st_ref.round_in_sum, st_ref.round_out_sum))
    assert st.round_in_sum < st_ref.round_in_sum and st.round_out_sum < st_ref.round_out_sum


def test_parallel_edge_multigraphs(solver):
    """Multigraphs dense in parallel edges and equal weights, where parallel
    record elimination drops the most: clusters of vertices with many edges
    between clusters, so later rounds carry many records per root pair.
    Bytes, L, H and rounds against the oracle; worklist sizes with records
    kept."""
    rng = np.random.default_rng(17)
    for it in range(6):
        n = int(rng.integers(2000, 20000))
        k = int(rng.integers(20, 200))                   # clusters of consecutive ids
        m = 4 * n
        u = rng.integers(0, n, m)
        # 80% of the edges stay inside their cluster, the rest cross to a
        # random vertex; each edge comes 1-3 times (parallel edges)
        width = max(2, n // k)
        inside = (u // width) * width + rng.integers(0, width, m)
        v = np.where(rng.random(m) < 0.8, np.minimum(inside, n - 1), rng.integers(0, n, m))
        rep = rng.integers(1, 4, m)
        edges = []
        for a, b, r in zip(u.tolist(), v.tolist(), rep.tolist()):
            if a != b:
                for _ in range(r):
                    edges.append((a, b, int(rng.integers(1, 6))))
        g = G(*csr_from_edges(n, edges))
        for uf, T in ((True, 2), (True, 4), (False, emst.INT_MAX)):
            check(solver, g, uf, T)


def _sparse_cases():
    rng = np.random.default_rng(23)
    out = [("rmat14", emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)),
           ("rmat13-ties", emst.gen_rmat(13, 16, 0.57, 0.19, 0.19, 5, 1, 9)),
           ("grid", emst.gen_grid(90, 70, 4, 1, 50)),
           ("rgg", emst.gen_rgg(20000, 6.0, 3, 1 << 20)),
           ("er", emst.gen_er(20000, 10.0, 4, 1, 1 << 20))]
    for i in range(12):
        n = int(rng.integers(2, 60))
        out.append((f"multi{i}", G(*random_graph(rng, n, int(rng.integers(0, 200)), 0, 4, allow_loops=True))))
    return out


def test_sparse_heavy_matches_full_pass(solver):
    """The sparse heavy initialisation (sorted, symmetric lists: only the
    adjacency outside the marked component is examined, crossing edges out
    of it found through reverse positions) emits the full pass's records:
    same bytes, L, H and rounds as with EMST_F_NO_SPARSE_HEAVY and as the
    oracle, for the plain rule and for thresholds that make the light phase
    small or large.  The sparse pass also makes the heavy phase's first
    offers (its select is skipped): with parallel records kept, the rounds
    and the worklist sizes entering and leaving every round are still the
    reference schedule's."""
    ran = 0
    for name, g in _sparse_cases():
        uf, T = emst.plain_threshold(g)
        w = np.sort(g.eweight) if g.edges else np.zeros(1, np.int32)
        for T2 in {T, int(w[len(w) // 50]), int(w[len(w) // 2]), int(w[-1]) - 1}:
            got, st = check(solver, g, True, T2)
            got2, st2 = with_flags(solver, emst.F_NO_SPARSE_HEAVY, lambda: solver.gpu_mst(g, True, T2))
            assert np.array_equal(got, got2), name
            assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds) == \
                (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds), name
            assert st2.sparse_heavy == 0
            ran += st.sparse_heavy
    assert ran >= 20  # the sparse pass ran wherever the marked component holds most positions


def test_light_offers_match_separate_select(solver):
    """With sorted, symmetric lists and the same weight on both copies of an
    edge, the light initialisation also makes the light phase's first offers
    (every vertex is a component of its own and sees all its edges in its own
    list) and that round's select is skipped: bytes, L, H, rounds and, with
    parallel records kept, every round's worklist sizes are the reference
    schedule's, and the same as with EMST_F_NO_LIGHT_OFFERS; filtered,
    unfiltered and learned.  The pass is chosen on the device when light
    edges are at least a tenth of the sampled positions: always unfiltered,
    and at a threshold at the third of the weights."""
    ran = 0
    cases = _sparse_cases()
    for name, g in cases:
        uf, T = emst.plain_threshold(g)
        w = np.sort(g.eweight) if g.edges else np.zeros(1, np.int32)
        for f, T2 in ((True, T), (True, int(w[len(w) // 3])), (False, emst.INT_MAX)):
            got, st = check(solver, g, f, T2)
            got2, st2 = with_flags(solver, emst.F_NO_LIGHT_OFFERS, lambda: solver.gpu_mst(g, f, T2))
            assert np.array_equal(got, got2), name
            assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds) == \
                (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds), name
            assert st2.light_offers == 0
            if T2 != T:
                assert st.light_offers == (1 if g.edges else 0), name
            ran += st.light_offers
        solver.load(g)
        st = solver.solve(emst.FILTER_LEARNED, 0)
        got = solver.fetch()
        st2 = with_flags(solver, emst.F_NO_LIGHT_OFFERS, lambda: solver.solve(emst.FILTER_LEARNED, 0))
        assert np.array_equal(got, solver.fetch()), name
        assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds) == \
            (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds), name
    assert ran >= 2 * len(cases)


def test_sparse_heavy_unmirrored_weights(solver):
    """Symmetric, sorted lists whose two copies of an edge carry different
    weights: the reference keys an edge by its upper-direction copy's weight,
    and so does the sparse pass (it reads eweight at the reverse position)."""
    rng = np.random.default_rng(31)
    for it in range(8):
        ni, nl, ew = random_graph(rng, 300, 1500, 1, 1000)
        src = np.repeat(np.arange(300), np.diff(ni))
        low = nl < src
        ew = ew.copy()
        ew[low] = rng.integers(1, 1000, int(low.sum()))
        g = G(ni, nl, ew)
        for T in (50, 300):
            got, st = check(solver, g, True, T)
            got2, _ = with_flags(solver, emst.F_NO_SPARSE_HEAVY, lambda: solver.gpu_mst(g, True, T))
            assert np.array_equal(got, got2)
            assert st.light_offers == 0  # the light offers need both copies' weights equal


def test_sparse_heavy_falls_back_on_unsorted_or_one_way_lists(solver):
    """Unsorted lists or one-way edges have no reverse positions: the full
    heavy pass runs and the result is still the oracle's."""
    rng = np.random.default_rng(37)
    for it in range(6):
        n = 200
        edges = [(int(a), int(b), int(c)) for a, b, c in zip(rng.integers(0, n, 700), rng.integers(0, n, 700),
                                                             rng.integers(1, 30, 700))]
        g1 = G(*csr_from_edges(n, edges, symmetric=True, sort=False))
        g2 = G(*csr_from_edges(n, edges, symmetric=False))
        for g in (g1, g2):
            for T in (3, 12):
                check(solver, g, True, T, rounds=False, kruskal=g is g1)


@pytest.mark.parametrize("flag,mode", [(emst.F_RELABEL, 1), (emst.F_CLUSTER, 2)], ids=["degree", "cluster"])
def test_relabelled_layout_matches_caller_numbering(solver, flag, mode):
    """emst_load renumbers a graph with sorted, symmetric lists (by
    decreasing degree when the degrees are skewed, in breadth-first cluster
    order when the ids are scattered and the diameter is high;
    EMST_F_RELABEL and EMST_F_CLUSTER force either) and solves on the rebuilt
    CSR with every position keyed by its edge's upper-direction position in
    the caller's CSR.  Bytes, L, H, rounds and,
    with parallel records kept, every round's worklist sizes are the
    oracle's and the same as in the caller's numbering (EMST_F_NO_RELABEL):
    ties, parallel edges, self-loops, unmirrored weights, filtered,
    unfiltered and learned."""
    rng = np.random.default_rng(41)
    cases = _sparse_cases()
    for i in range(4):
        ni, nl, ew = random_graph(rng, 300, 1500, 1, 1000)
        src = np.repeat(np.arange(300), np.diff(ni))
        ew = ew.copy()
        low = nl < src
        ew[low] = rng.integers(1, 1000, int(low.sum()))  # the copies disagree: the upper one counts
        cases.append((f"unmirrored{i}", G(ni, nl, ew)))
    ran = 0
    for name, g in cases:
        uf, T = emst.plain_threshold(g)
        w = np.sort(g.eweight) if g.edges else np.zeros(1, np.int32)
        for f, T2 in ((uf, T), (True, int(w[len(w) // 3])), (False, emst.INT_MAX)):
            got, st = with_flags(solver, flag, lambda: check(solver, g, f, T2))
            got2, st2 = with_flags(solver, emst.F_NO_RELABEL, lambda: solver.gpu_mst(g, f, T2))
            assert np.array_equal(got, got2), name
            assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds) == \
                (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds), name
            assert st2.relabelled == 0
            assert st.relabelled in (0, mode), name
            ran += st.relabelled == mode
        # learned threshold on the renumbered copy: the same bytes as in the
        # caller's numbering
        with_flags(solver, flag, lambda: solver.load(g))
        st = solver.solve(emst.FILTER_LEARNED, 0)
        got = solver.fetch()
        with_flags(solver, emst.F_NO_RELABEL, lambda: solver.load(g))
        st2 = solver.solve(emst.FILTER_LEARNED, 0)
        assert np.array_equal(got, solver.fetch()), name
        assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds, st.threshold_used) == \
            (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds, st2.threshold_used), name
        assert np.array_equal(got, O.kruskal(g)), name
    # every case has sorted, symmetric lists: each one with an edge was renumbered
    assert ran == 3 * sum(1 for _, g in cases if g.edges > 0 and g.nodes >= 2)


def test_relabel_default_choice(solver):
    """By default a graph whose largest degree is at least 32 times the mean
    is renumbered by degree; one of at least 65536 vertices whose edges
    mostly join distant ids is renumbered in cluster order when its
    breadth-first search is deep (a geometric graph), not when it is shallow
    (Erdos-Renyi); local ids (a grid) and small graphs keep the caller's
    numbering; lists without reverse positions never are renumbered."""
    rmat = emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)
    for g, want in ((rmat, 1), (emst.gen_grid(90, 70, 4, 1, 50), 0), (emst.gen_rgg(20000, 6.0, 3, 1 << 20), 0),
                    (emst.gen_er(20000, 10.0, 4, 1, 1 << 20), 0), (emst.gen_grid(400, 400, 4, 1, 1 << 20), 0),
                    (emst.gen_rgg(1 << 17, 6.0, 5, 1 << 20), 2), (emst.gen_er(1 << 17, 10.0, 6, 1, 1 << 20), 0)):
        solver.load(g)
        uf, T = emst.plain_threshold(g)
        st = solver.solve(uf, T)
        assert st.relabelled == want, (g.nodes, g.edges)
        if want == 2:
            assert np.array_equal(solver.fetch(), O.kruskal(g))
    rng = np.random.default_rng(43)
    edges = [(int(a), int(b), int(c)) for a, b, c in zip(rng.integers(0, 2000, 9000), rng.integers(0, 2000, 9000),
                                                         rng.integers(1, 30, 9000))]
    g = G(*csr_from_edges(2000, edges, symmetric=True, sort=False))
    got, st = with_flags(solver, emst.F_RELABEL, lambda: solver.gpu_mst(g, True, 10))
    assert st.relabelled == 0 and np.array_equal(got, O.kruskal(g))
