"""Full-size configurations on the GPU, through the C-ABI.

C1 grid 512x512, C2 RMAT-24 and C3 RGG-2^24: the solve's counts (records,
rounds, worklist sizes entering and leaving the rounds, weight) against the
oracle's round schedule committed in tests/golden/schedules.json (made by
make_schedules.py); C2 and C3 also in every filter mode against the
oracle's serial Kruskal, bit for bit, and so are the edge-list baseline
and the 4-shard solve (the contraction baseline by forest weight).
C5 RMAT-26 with weights in [1, 64] (heavy ties, a forest, ~2e9 directed
edges): the forest is unique under the (w, j) order, so the unfiltered,
plain and learned runs must give identical bytes; the oracle's union-find
then checks that it is a spanning forest, and the weights agree with the
OpenMP Boruvka restatement.  Set EMST_SKIP_RMAT26=1 to leave C5 out."""
import json
import os
import time

import numpy as np
import pytest

import emst
import oracle_lib as O
import shard
from golden.make_schedules import digest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp.f32")
with open(os.path.join(ROOT, "tests", "golden", "schedules.json")) as _f:
    SCHEDULES = json.load(_f)


def _solver():
    s = emst.Solver(0)
    s.set_filter_model(np.fromfile(MODEL, np.float32))
    return s


def _modes(s, g):
    uf, T = emst.plain_threshold(g)
    s.load(g)
    out = {}
    for name, (mode, thr) in (("plain", (emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T)),
                              ("learned", (emst.FILTER_LEARNED, 0)), ("off", (emst.FILTER_OFF, emst.INT_MAX))):
        t0 = time.time()
        st = s.solve(mode, thr)
        out[name] = (s.fetch(), st)
        print(f"  {name}: {st.ms:.2f} ms, rounds {st.light_rounds}+{st.heavy_rounds}, light {st.light}, "
              f"heavy {st.heavy} ({time.time()-t0:.1f}s)", flush=True)
    return out


def check_schedules(s, g, cfg):
    """The solve's deterministic counts against the oracle's schedule: with
    parallel records kept (the reference schedule) every count; with them
    eliminated (the default) the same counts but worklist sizes no larger."""
    ent = SCHEDULES[cfg]
    assert (g.nodes, g.edges) == (ent["nodes"], ent["edges"]) and digest(g) == ent["sha256"]
    s.load(g)
    saved = s.config().as_dict()
    for run in ent["runs"]:
        mode = emst.FILTER_PLAIN if run["use_filter"] else emst.FILTER_OFF
        want = (run["L"], run["H"], run["light_rounds"], run["heavy_rounds"],
                sum(run["light_in"]) + sum(run["heavy_in"]), sum(run["light_out"]) + sum(run["heavy_out"]),
                run["weight"])
        s.set_config(**{**saved, "flags": saved["flags"] | emst.F_NO_PAIR_DEDUP})
        st = s.solve(mode, run["threshold"])
        got = (st.light, st.heavy, st.light_rounds, st.heavy_rounds, st.round_in_sum, st.round_out_sum,
               emst.mst_weight(g, s.fetch()))
        s.set_config(**saved)
        print(f"  schedule T={run['threshold']}: {got}", flush=True)
        assert got == want, (cfg, run["threshold"])
        st = s.solve(mode, run["threshold"])
        got = (st.light, st.heavy, st.light_rounds, st.heavy_rounds, emst.mst_weight(g, s.fetch()))
        assert got == want[:4] + want[6:], (cfg, run["threshold"], "pair elimination")
        assert st.round_in_sum <= want[4] and st.round_out_sum <= want[5]


def test_c1_grid_schedule():
    g = emst.gen_grid(512, 512, 1, 1, 1 << 20)
    s = _solver()
    check_schedules(s, g, "grid512")
    s.close()


@pytest.mark.parametrize("cfg", ["rmat24", "rgg24"])
def test_c2_c3_against_kruskal(cfg):
    t0 = time.time()
    g = (emst.gen_rmat(24, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24) if cfg == "rmat24"
         else emst.gen_rgg(1 << 24, 6.0, 3, 1 << 24))
    print(f"{cfg}: n={g.nodes} m={g.edges} generated in {time.time()-t0:.1f}s", flush=True)
    s = _solver()
    res = _modes(s, g)
    check_schedules(s, g, cfg)
    # the related-work baselines at full size: the edge-list Boruvka must
    # give the same bytes, the contraction Boruvka the same forest weight
    el, est = s.edgelist_mst(g)
    ct, cst = s.contract_mst(g)
    print(f"  edgelist {est.ms:.2f} ms ({est.iterations} iterations), contract {cst.ms:.2f} ms", flush=True)
    # the edge-sharded scheme (4 shards, learned threshold), shard by shard
    sharded, info = shard.solve_sharded(g, 4, lambda h: s.gpu_mst(h, emst.FILTER_LEARNED, 0)[0])
    print(f"  sharded: gathered {info['gathered']}", flush=True)
    s.close()
    ref = O.kruskal(g)
    for name, (got, st) in res.items():
        assert np.array_equal(got, ref), name
    rc, cw, gw, oc, og = O.verify(g, ref, res["learned"][0])
    assert rc == 0 and cw == gw and oc == og == 0
    assert np.array_equal(el, ref) and est.weight == cw
    assert cst.weight == cw and cst.mst_edges == int(ref.sum())
    assert np.array_equal(sharded, ref)


@pytest.mark.skipif(os.environ.get("EMST_SKIP_RMAT26") == "1", reason="EMST_SKIP_RMAT26=1")
def test_c5_rmat26_ties_properties():
    t0 = time.time()
    g = emst.gen_rmat(26, 16, 0.57, 0.19, 0.19, 5, 1, 65)
    assert g.edges < 2**31
    print(f"rmat26: n={g.nodes} m={g.edges} generated in {time.time()-t0:.1f}s", flush=True)
    s = _solver()
    res = _modes(s, g)
    s.close()
    base = res["off"][0]
    for name, (got, st) in res.items():
        assert np.array_equal(got, base), name
    t1 = time.time()
    mk, cc, ok = O.forest_check(g, base)
    print(f"  forest check {time.time()-t1:.1f}s: marked {mk}, components {cc}", flush=True)
    assert ok and mk == g.nodes - cc
    assert cc > 1   # a forest: isolated vertices
    w = emst.mst_weight(g, base)
    t1 = time.time()
    b = O.boruvka_omp(g, 0)
    print(f"  OpenMP Boruvka {time.time()-t1:.1f}s", flush=True)
    assert emst.mst_weight(g, b) == w
    assert np.array_equal(b, base)


def test_c4_batch_4096_every_graph():
    """C4 at its real size: the 4,096-graph batch of the bench (seed 4, n
    log-uniform in [1e3, 1e6], grid / geometric / RMAT ef 8 / ER deg 10,
    weights in [1, 2^20)), about 5.3e9 directed edges in three arenas of
    just under 2^31 edges (64-bit arena offsets).  Plain per-graph rule and
    learned thresholds (union network), each solved with capture and then
    replayed from the captured hipGraph; after every solve every graph's
    inMST must equal the oracle's Kruskal of that graph alone."""
    import sys
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, ROOT)
    import bench
    t0 = time.time()
    graphs = [g for _, g in bench.batch_graphs(emst, 4096, 4, 1000, 1_000_000)]
    tot = sum(g.edges for g in graphs)
    print(f"c4: {len(graphs)} graphs, {tot} edges, generated in {time.time()-t0:.0f}s", flush=True)
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:   # the oracle's ctypes calls run in parallel
        refs = list(ex.map(O.kruskal, graphs))
    print(f"  oracle Kruskal of every graph in {time.time()-t0:.0f}s", flush=True)
    s = _solver()
    b = emst.Batch(s, graphs)
    b.set_filter_model(np.fromfile(os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp_union.f32"), np.float32))
    chunks, nodes, edges = b.info()
    assert edges == tot and chunks == 3
    for mode in (emst.FILTER_PLAIN, emst.FILTER_LEARNED):
        for rep in range(2):
            st = b.solve(mode, use_graph=True)
            print(f"  mode {mode} solve {rep}: {st.ms:.1f} ms, arenas {st.chunks}, replays {st.replays}, "
                  f"filtered {st.filtered_graphs}", flush=True)
            assert st.graphs == 4096 and st.chunks == 3
            assert st.replays == (3 if rep else 0)   # the second solve replays every arena's capture
            bad = [i for i in range(len(graphs)) if not np.array_equal(b.fetch(i), refs[i])]
            assert not bad, bad[:10]
    b.close()
    s.close()
