"""Pins the CPU oracle (oracle/emst_oracle.c) against known answers, brute
force and closed forms, so that the GPU parity tests check against
something trustworthy.  CPU only."""
import itertools
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from graphs import csr_from_edges, random_graph

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
REF_DUMP = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_egr_dump")


class G:
    def __init__(self, ni, nl, ew):
        self.nindex, self.nlist, self.eweight = ni, nl, ew


def test_hash_kats():
    for v, h in GOLD["hash"].items():
        assert O.ecl_hash(int(v)) == h


def test_synth_weight_cases():
    for i, k, n, w in GOLD["synth"]:
        # a graph with n vertices and the single edge i->k
        nindex = np.zeros(n + 1, np.int32)
        nindex[i + 1:] = 1
        got = O.synth_weights(nindex, np.array([k], np.int32))
        assert int(got[0]) == w, (i, k, n)


def test_threshold_closed_forms():
    # m/n < 4: no filter, T = INT_MAX (ECL_MST.cu:443)
    assert O.plain_threshold(10, np.arange(39, dtype=np.int32)) == (False, 2**31 - 1)
    # n=10, m=40: s=20, index = min(19, int(3*10*20/40)) = 15
    w = np.arange(40, dtype=np.int32)
    samples = sorted(int(w[O.ecl_hash(i) % 40]) for i in range(20))
    assert O.plain_threshold(10, w) == (True, samples[15])
    # n=100, m=400 (avg 4): index = int(3*100*20/400) = 15; n=100, m=6000: index = 1
    w = np.arange(6000, dtype=np.int32)[::-1].copy()
    samples = sorted(int(w[O.ecl_hash(i) % 6000]) for i in range(20))
    assert O.plain_threshold(100, w) == (True, samples[1])
    # fewer than 20 edges: n=1, m=4 -> s=4, index = int(3*1*4/4) = 3 (largest)
    w = np.array([9, 2, 7, 5], np.int32)
    samples = sorted(int(w[O.ecl_hash(i) % 4]) for i in range(4))
    assert O.plain_threshold(1, w) == (True, samples[3])


def _rank_brute_msf(n, nindex, nlist, ew):
    """Unique MSF under the strict order (w, j) by exhaustive search: among
    all maximal forests pick the one minimising sum(2**rank)."""
    src = np.repeat(np.arange(n), np.diff(nindex))
    ups = [j for j in range(len(nlist)) if nlist[j] > src[j]]
    order = sorted(ups, key=lambda j: (int(ew[j]), j))
    rank = {j: r for r, j in enumerate(order)}

    def comps(edges):
        p = list(range(n))

        def f(x):
            while p[x] != x:
                x = p[x]
            return x
        ok = True
        for j in edges:
            a, b = f(src[j]), f(nlist[j])
            if a == b:
                ok = False
            else:
                p[a] = b
        return ok, len({f(x) for x in range(n)})
    _, c_all = comps(ups) if ups else (True, n)
    best, best_cost = None, None
    target = n - c_all
    for sub in itertools.combinations(ups, target):
        ok, _ = comps(sub)
        if not ok:
            continue
        cost = sum(1 << rank[j] for j in sub)
        if best_cost is None or cost < best_cost:
            best, best_cost = sub, cost
    out = np.zeros(len(nlist), np.uint8)
    for j in best or ():
        out[j] = 1
    return out


def test_kruskal_vs_brute_force():
    rng = np.random.default_rng(1)
    done = 0
    while done < 150:
        n = int(rng.integers(1, 8))
        ni, nl, ew = random_graph(rng, n, int(rng.integers(0, 11)), 0, int(rng.integers(1, 4)))
        g = G(ni, nl, ew)
        ref = _rank_brute_msf(n, ni, nl, ew)
        assert np.array_equal(O.kruskal(g), ref)
        assert np.array_equal(O.schedule(g, False, 2**31 - 1)[0], ref)
        for T in (0, 1, 2):
            assert np.array_equal(O.schedule(g, True, T)[0], ref)
        done += 1


def test_closed_forms():
    # a tree: every upper-direction edge is in the MSF
    rng = np.random.default_rng(2)
    n = 50
    edges = [(v, int(rng.integers(0, v)), int(rng.integers(1, 100))) for v in range(1, n)]
    ni, nl, ew = csr_from_edges(n, edges)
    g = G(ni, nl, ew)
    up = nl > np.repeat(np.arange(n), np.diff(ni))
    assert np.array_equal(O.kruskal(g).astype(bool), up)
    # complete graph, equal weights: ties go to the smallest j, which are
    # vertex 0's edges, so the MSF is the star centred at 0
    n = 9
    ni, nl, ew = csr_from_edges(n, [(a, b, 5) for a in range(n) for b in range(a + 1, n)])
    k = O.kruskal(G(ni, nl, ew))
    assert k.sum() == n - 1 and np.all(k[ni[0]:ni[1]] == 1)
    # a cycle with equal weights drops exactly its largest-j upper edge
    n = 12
    ni, nl, ew = csr_from_edges(n, [(v, (v + 1) % n, 3) for v in range(n)])
    k = O.kruskal(G(ni, nl, ew))
    src = np.repeat(np.arange(n), np.diff(ni))
    upj = np.nonzero(nl > src)[0]
    assert k.sum() == n - 1 and k[upj.max()] == 0
    # isolated vertices and two components: MSF has n - c edges
    ni, nl, ew = csr_from_edges(7, [(0, 1, 4), (1, 2, 4), (4, 5, 1)])
    k, st = O.schedule(G(ni, nl, ew), False, 2**31 - 1)
    assert k.sum() == 3 and st.components == 4 and st.mst_edges == 3


def test_round_schedule_hand_trace():
    # path 0-1-2-3, weights 1, 3, 1: round 1 picks (0,1) and (2,3);
    # round 2 sees 3 records, only (1,2) crosses and is joined; round 3
    # sees 1 record and drops it.
    ni, nl, ew = csr_from_edges(4, [(0, 1, 1), (1, 2, 3), (2, 3, 1)])
    k, st = O.schedule(G(ni, nl, ew), False, 2**31 - 1)
    assert st.light == 3 and st.light_rounds == 3
    assert st.rounds(0) == [(3, 3), (3, 1), (1, 0)]
    # same graph with T=1: light phase has the two weight-1 edges (2 rounds:
    # 2 cross, then 0); heavy pass finds (1,2) crossing: 1 round + empty
    k2, st2 = O.schedule(G(ni, nl, ew), True, 1)
    assert np.array_equal(k, k2)
    assert st2.rounds(0) == [(2, 2), (2, 0)]
    assert st2.heavy_all == 1 and st2.heavy == 1 and st2.rounds(1) == [(1, 1), (1, 0)]


def test_schedule_equals_kruskal_random():
    rng = np.random.default_rng(5)
    for _ in range(100):
        n = int(rng.integers(1, 300))
        g = G(*random_graph(rng, n, int(rng.integers(0, 900)), 0, int(rng.integers(1, 50))))
        k = O.kruskal(g)
        for uf, T in ((False, 2**31 - 1), (True, int(rng.integers(0, 50)))):
            s, st = O.schedule(g, uf, T)
            assert np.array_equal(s, k)
            assert st.mst_edges + st.components == n


def test_negative_weights_differ_only_by_order():
    # A negative weight sorts first for Kruskal (signed tuple order) but
    # last for the GPU key (unsigned, ECL_MST.cu:221).
    ni, nl, ew = csr_from_edges(3, [(0, 1, -1), (1, 2, 2), (0, 2, 3)])
    g = G(ni, nl, ew)
    k = O.kruskal(g)
    s, _ = O.schedule(g, False, 2**31 - 1)
    src = np.repeat(np.arange(3), np.diff(ni))
    def edges_of(x):
        return sorted((int(src[j]), int(nl[j])) for j in np.nonzero(x)[0])
    assert edges_of(k) == [(0, 1), (1, 2)]
    assert edges_of(s) == [(0, 2), (1, 2)]


def test_verify_cases():
    ni, nl, ew = csr_from_edges(4, [(0, 1, 2), (1, 2, 3), (2, 3, 4), (0, 3, 1)])
    g = G(ni, nl, ew)
    k = O.kruskal(g)
    assert O.verify(g, k, k)[0] == 0
    bad = k.copy()
    j = int(np.nonzero(k)[0][0])
    bad[j] = 0
    rc, cw, gw, oc, og = O.verify(g, k, bad)
    assert rc == 1 and oc == 1 and og == 0 and cw - gw == ew[j]


def _write(path, ints, extra=b""):
    with open(path, "wb") as f:
        f.write(np.asarray(ints, np.int32).tobytes() + extra)


def egr_fixtures(tmp):
    """(name, path, expected oracle rc)."""
    fx = []
    ni, nl, ew = csr_from_edges(5, [(0, 1, 7), (1, 2, 8), (3, 4, 9), (0, 4, 1)])
    n, m = 5, len(nl)
    full = [n, m] + list(ni) + list(nl) + list(ew)
    def add(name, ints, rc, extra=b""):
        p = os.path.join(tmp, name + ".egr")
        _write(p, ints, extra)
        fx.append((name, p, rc))
    add("weighted", full, 0)
    add("weighted_trailing", full, 0, b"\x01\x02")
    add("unweighted", full[:2 + n + 1 + m], 0)
    add("partial_weights", full[:-2], O_E["WEIGHTS"])
    add("truncated_nlist", full[:2 + n + 1 + m - 1], O_E["NLIST"])
    add("truncated_nindex", full[:2 + 3], O_E["NINDEX"])
    add("short_header", full[:1], O_E["EDGES"])
    add("empty", [], O_E["NODES"])
    add("zero_nodes", [0, 0, 0], O_E["COUNTS"])
    add("neg_edges", [3, -1, 0, 0, 0, 0], O_E["COUNTS"])
    add("no_edges", [3, 0, 0, 0, 0, 0], 0)
    fx.append(("missing", os.path.join(tmp, "does_not_exist.egr"), O_E["OPEN"]))
    return fx


O_E = {"OPEN": -1, "NODES": -2, "EDGES": -3, "COUNTS": -4, "NINDEX": -6, "NLIST": -7, "WEIGHTS": -8}


def test_oracle_reader_fixtures(tmp_path):
    for name, path, rc in egr_fixtures(str(tmp_path)):
        got = O.read_egr(path)
        assert got[0] == rc, name


@pytest.mark.skipif(not os.path.exists(REF_DUMP), reason="reference reader not built (needs /root/reference)")
def test_oracle_reader_matches_reference_reader(tmp_path):
    """The reference's own readECLgraph (ECLgraph.h:37-64), compiled from
    its source into oracle/_ref, on the same bytes."""
    def fnv(a):
        h = 1469598103934665603
        for b in np.asarray(a, np.int32).tobytes():
            h ^= b
            h = (h * 1099511628211) % (1 << 64)
        return h
    for name, path, rc in egr_fixtures(str(tmp_path)):
        r = subprocess.run([REF_DUMP, path], capture_output=True, text=True)
        if rc != 0:
            assert r.returncode != 0, name          # the reference exits(-1)
            continue
        assert r.returncode == 0, name
        n, m, hw, h1, h2, h3 = map(int, r.stdout.split())
        _, ni, nl, ew = O.read_egr(path)
        assert (n, m, hw) == (len(ni) - 1, len(nl), int(ew is not None)), name
        assert h1 == fnv(ni) and h2 == fnv(nl), name
        if ew is not None:
            assert h3 == fnv(ew), name


def test_parallel_cpu_baselines_match_kruskal():
    """The OpenMP Boruvka and Filter-Kruskal baselines are independent MSF
    algorithms; both must reproduce the oracle's Kruskal exactly."""
    import emst
    rng = np.random.default_rng(21)
    graphs = [G(*random_graph(rng, int(rng.integers(1, 200)), int(rng.integers(0, 800)), -3, 6)) for _ in range(40)]
    graphs += [emst.gen_rmat(14, 16, seed=2), emst.gen_grid(60, 70, 1), emst.gen_rgg(30000, 6.0, 3),
               emst.gen_rmat(13, 16, seed=5, wlo=1, whi=65)]
    for g in graphs:
        k = O.kruskal(g)
        assert np.array_equal(O.boruvka_omp(g, 4), k)
        assert np.array_equal(O.filter_kruskal_omp(g, 4), k)


def test_forest_check_pins():
    import emst
    rng = np.random.default_rng(31)
    for _ in range(30):
        n = int(rng.integers(1, 120))
        g = G(*random_graph(rng, n, int(rng.integers(0, 300)), 0, 9))
        k = O.kruskal(g)
        mk, cc, ok = O.forest_check(g, k)
        assert ok and mk == n - cc
        _, st = O.schedule(g, False, 2**31 - 1)
        assert cc == st.components
        if k.sum() > 0:                      # drop an edge: no longer spanning
            bad = k.copy(); bad[np.nonzero(k)[0][0]] = 0
            mk2, cc2, ok2 = O.forest_check(g, bad)
            assert ok2 and mk2 == n - cc2 - 1
        src = np.repeat(np.arange(n), np.diff(g.nindex))
        extra = np.nonzero((k == 0) & (g.nlist > src))[0]
        if len(extra):                       # add a non-tree edge: a cycle
            bad = k.copy(); bad[extra[0]] = 1
            assert not O.forest_check(g, bad)[2]


def _renumber(g, order):
    """A renumbered copy of a sorted, symmetric CSR (numpy restatement of
    the solver's load-time copy): new id i is the caller's order[i]; every
    list is sorted by new neighbour id (ties keep the caller's positions);
    each position carries its edge's key id kid (the caller's position of
    the upper-direction copy; the k-th copy of y in x's list pairs with the
    k-th copy of x in y's list) and that copy's weight."""
    n, m = len(g.nindex) - 1, len(g.nlist)
    deg = np.diff(g.nindex)
    src = np.repeat(np.arange(n), deg)
    nl = g.nlist.astype(np.int64)
    pos = np.arange(m)
    lo, hi = np.minimum(src, nl), np.maximum(src, nl)
    up = nl > src
    kid = pos.copy()
    # rank of each copy among the copies of its pair, in position order
    pair = lo * n + hi
    o = np.lexsort((pos, up, pair))
    ps, us = pair[o], up[o]
    grp_start = np.r_[0, np.flatnonzero((ps[1:] != ps[:-1]) | (us[1:] != us[:-1])) + 1]
    rank = np.empty(m, np.int64)
    rank[o] = np.arange(m) - np.repeat(grp_start, np.diff(np.r_[grp_start, m]))
    upper_of = {}
    for p in np.flatnonzero(up):
        upper_of[(int(pair[p]), int(rank[p]))] = int(p)
    for p in np.flatnonzero((~up) & (nl != src)):
        kid[p] = upper_of[(int(pair[p]), int(rank[p]))]
    newid = np.empty(n, np.int64)
    newid[order] = np.arange(n)
    ns, nd = newid[src], newid[nl]
    o = np.lexsort((pos, nd, ns))
    ni = np.zeros(n + 1, np.int64)
    np.add.at(ni, ns + 1, 1)
    ni = np.cumsum(ni)
    w = g.eweight[kid[o]]
    return G(ni.astype(np.int32), nd[o].astype(np.int32), w.astype(np.int32)), kid[o].astype(np.int32)


def test_schedule_is_invariant_under_renumbering_with_key_ids():
    """The mathematics behind the solver's renumbered copies: keyed by
    (weight, position of the upper-direction copy) the keys are distinct, so
    every phase's forest is the unique minimum forest under that order, and
    the round structure (records per round, L, H, rounds) depends only on
    which components merge, not on vertex ids.  The schedule run on a
    renumbered copy whose positions carry the original key ids (and the
    upper copies' weights) must therefore mark exactly the original inMST
    bytes with the same statistics: random permutations and the degree order,
    on multigraphs with ties, self-loops and copies whose weights differ."""
    rng = np.random.default_rng(53)
    for it in range(40):
        n = int(rng.integers(2, 120))
        ni, nl, ew = random_graph(rng, n, int(rng.integers(0, 400)), 1, 6, allow_loops=True)
        ew = ew.copy()
        src = np.repeat(np.arange(n), np.diff(ni))
        if it % 2:
            low = nl < src  # lower copies disagree: only the upper copy's weight counts
            ew[low] = rng.integers(1, 6, int(low.sum()))
        g = G(ni, nl, ew)
        w = np.sort(g.eweight) if len(g.nlist) else np.zeros(1, np.int32)
        deg = np.diff(ni)
        for order in (np.argsort(-deg, kind="stable"), rng.permutation(n)):
            g2, kid = _renumber(g, order)
            for uf, T in ((False, 2 ** 31 - 1), (True, int(w[len(w) // 3])), (True, int(w[-1]))):
                ref, st = O.schedule(g, uf, T)
                got, st2 = O.schedule_keyed(g2, uf, T, kid)
                assert np.array_equal(got, ref), (it, uf, T)
                assert (st.light, st.heavy, st.light_rounds, st.heavy_rounds, st.components, st.mst_edges,
                        st.weight) == (st2.light, st2.heavy, st2.light_rounds, st2.heavy_rounds, st2.components,
                                       st2.mst_edges, st2.weight)
                assert st.rounds(0) == st2.rounds(0) and st.rounds(1) == st2.rounds(1)
    # a permutation that keys by the new positions instead (no key ids) is
    # not invariant on such ties: the check above is not vacuous
    g = G(*random_graph(np.random.default_rng(59), 60, 300, 1, 3))
    g2, kid = _renumber(g, np.random.default_rng(61).permutation(60))
    diff = 0
    for uf, T in ((False, 2 ** 31 - 1),):
        ref, _ = O.schedule(g, uf, T)
        got_pos, _ = O.schedule(g2, uf, T)
        mapped = np.zeros_like(ref)
        mapped[kid[got_pos.astype(bool)]] = 1
        diff += int((mapped != ref).sum())
    assert diff > 0
