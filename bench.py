"""Benchmark of the MI355X ECL-MST solver.

Default workload (one GPU): RMAT scale 24, edge factor 16, (a,b,c,d) =
(0.57, 0.19, 0.19, 0.05), seed 2, symmetrized, self-loops and duplicates
removed, weights uniform int32 in [1, 2^24) mirrored on both directions.
A step is one device-resident solve (the reference's timed region,
ECL_MST.cu:301-345).  Two paths are measured, both exact:
  learned -- light-edge threshold predicted on the device by the trained
             network (mst-artifact-sc23_amd/filter_mlp.f32); the headline
  plain   -- the reference's threshold rule (ECL_MST.cu:432-443)
The metric is directed CSR edges per second (generate_compare_tables.py:
55-59 reports Medges/s; this reports Gedges/s).  emst_load renumbers the
RMAT graphs by degree and the geometric graph by clusters, on the device and
outside the timed region (its cost is in pcie_inclusive's load_ms); the
line's `layout` record times the same solve in the caller's numbering.

The default run also carries two sub-records, the rest of the headline
metric: "rgg24" (the random geometric graph of 2^24 vertices, learned
threshold against the plain rule) and "batch" (the batched configuration:
4,096 graphs, n log-uniform in [1e3, 1e6], generator uniform over grid /
geometric deg 6 / RMAT ef 8 / Erdos-Renyi deg 10, weights in [1, 2^20),
seed 4; graphs/s learned and plain, per-graph latency from arena
completion times, and the per-graph loop timed on every graph).
--workload batch runs the batched configuration alone, as the main record.
--workload ladder times the de-optimization ladder and the related-work
baselines; --workload shard runs the edge-sharded single-graph scheme shard by
shard on this GPU (--shards, --shard-workload) and checks it against one
whole-graph solve.

Prints one JSON line.  --gpus above 1 is refused: the solver runs on one
GPU.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mst-artifact-sc23_amd")
sys.path.insert(0, PKG)
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "rmat24": dict(kind="rmat", scale=24, ef=16, seed=2, wlo=1, whi=1 << 24,
                   desc="RMAT scale 24, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, w in [1,2^24)"),
    "rmat26": dict(kind="rmat", scale=26, ef=16, seed=5, wlo=1, whi=65,
                   desc="RMAT scale 26, edge factor 16, (0.57,0.19,0.19,0.05), seed 5, w in [1,64]"),
    "grid512": dict(kind="grid", rows=512, cols=512, seed=1, wlo=1, whi=1 << 20,
                    desc="512x512 4-neighbour grid, seed 1, w in [1,2^20)"),
    "rgg24": dict(kind="rgg", n=1 << 24, deg=6.0, seed=3, wmax=1 << 24,
                  desc="random geometric graph, 2^24 points, mean degree 6, seed 3, w in [1,2^24)"),
}


def make_graph(emst, w, scale=None):
    if w["kind"] == "rmat":
        return emst.gen_rmat(scale or w["scale"], w["ef"], 0.57, 0.19, 0.19, w["seed"], w["wlo"], w["whi"])
    if w["kind"] == "grid":
        return emst.gen_grid(w["rows"], w["cols"], w["seed"], w["wlo"], w["whi"])
    return emst.gen_rgg(w["n"], w["deg"], w["seed"], w["wmax"])


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(emst, w, sample_scale, g_full):
    """CPU baselines on the same input, timed on this host's cores: the two
    parallel restatements (OpenMP Boruvka in the structure of
    UMinho_BOTH/src/BoruvkaUMinho_OMP.cpp, and Filter-Kruskal, the method of
    PBBS parallelFilterKruskal) on all host cores (at most the 16 of a
    one-GPU share), and the oracle's serial Kruskal (restating
    ECL_MST.cu:82-124) on one core.  `value` is the fastest exact parallel
    restatement; the serial Kruskal is reported beside it.  sample_scale 0
    times the full workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    if sample_scale and w["kind"] == "rmat" and sample_scale != w["scale"]:
        g = make_graph(emst, w, scale=sample_scale)
        what = f"RMAT scale {sample_scale} with the same generator parameters"
    else:
        g = g_full
        what = "the full workload"
    t0 = time.perf_counter()
    ref = oracle_lib.kruskal(g)
    dt = time.perf_counter() - t0
    cores = min(len(os.sched_getaffinity(0)), 16)
    par = {}
    for name, fn in (("openmp_boruvka", oracle_lib.boruvka_omp), ("openmp_filter_kruskal", oracle_lib.filter_kruskal_omp)):
        t1 = time.perf_counter()
        got = fn(g, cores)
        d = time.perf_counter() - t1
        par[name] = {"value": g.edges / d / 1e9, "unit": "Gedges/s", "cores": cores, "seconds": d,
                     "exact": bool(np.array_equal(got, ref))}
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    exact = {k: v for k, v in par.items() if v["exact"]}
    best = max(exact, key=lambda k: exact[k]["value"]) if exact else None
    serial = {"value": g.edges / dt / 1e9, "unit": "Gedges/s", "cores": 1, "seconds": dt}
    if best is None:  # no parallel code was exact: fall back to the serial oracle
        return {"value": serial["value"], "unit": "Gedges/s", "cores": 1, "kind": "port", "cpu_model": model,
                "sample": f"oracle serial Kruskal on {what}: m={g.edges} directed edges in {dt:.2f} s",
                "serial_kruskal": serial, "parallel_restatements": par}
    return {"value": par[best]["value"], "unit": "Gedges/s", "cores": cores, "kind": "port", "cpu_model": model,
            "sample": f"{best} (restatement, {cores} threads) on {what}: m={g.edges} directed edges in "
                      f"{par[best]['seconds']:.2f} s, inMST equal to the oracle's Kruskal",
            "serial_kruskal": serial, "parallel_restatements": par}


def timed_steps(s, mode, thr, steps, warmup):
    for _ in range(warmup):
        s.solve(mode, thr)
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(s.solve(mode, thr))  # each solve ends with a stream synchronize
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, stats


def interleaved_ratio(emst, s, plain_mode, T, steps):
    """Plain-rule time over learned-threshold time, from solves taken in
    alternation (learned, plain, learned, ...), each timed by the wall clock
    around its synchronous call: the two rules see the same device state, so
    the order in which the timed blocks run does not bias the ratio."""
    for _ in range(2):
        s.solve(emst.FILTER_LEARNED, 0)
        s.solve(plain_mode, T)
    tl = tp = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        s.solve(emst.FILTER_LEARNED, 0)
        t1 = time.perf_counter()
        s.solve(plain_mode, T)
        t2 = time.perf_counter()
        tl += t1 - t0
        tp += t2 - t1
    return tp / tl


def read_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def run_single(a, emst):
    w = WORKLOADS[a.workload]
    t0 = time.perf_counter()
    g = make_graph(emst, w, a.scale)
    log(f"generated {a.workload}: n={g.nodes} m={g.edges} in {time.perf_counter()-t0:.1f}s")
    s = emst.Solver(0)
    s.load(g)
    model = os.path.join(PKG, "filter_mlp.f32")
    s.set_filter_model(np.fromfile(model, np.float32))
    uf, T = emst.plain_threshold(g)
    plain_mode = emst.FILTER_PLAIN if uf else emst.FILTER_OFF
    if a.autotune:
        cfg, ms = s.autotune(emst.FILTER_LEARNED, 0, reps=3)
        log(f"autotuned: {cfg} {ms:.3f} ms")
    ms_learned, st_l = timed_steps(s, emst.FILTER_LEARNED, 0, a.steps, a.warmup)
    learned_inmst = s.fetch()
    log(f"learned: {ms_learned:.3f} ms/step")
    ms_plain, st_p = timed_steps(s, plain_mode, T, a.steps, a.warmup)
    plain_inmst = s.fetch()
    log(f"plain: {ms_plain:.3f} ms/step")
    # PCIe-inclusive view (not the metric): upload of the CSR (after the host
    # CSR validation), one solve, download of inMST
    t1 = time.perf_counter()
    s.load(g)
    load_ms = (time.perf_counter() - t1) * 1e3
    st_x = s.solve(emst.FILTER_LEARNED, 0)
    t1 = time.perf_counter()
    s.fetch()
    fetch_ms = (time.perf_counter() - t1) * 1e3
    pcie = {"load_ms": load_ms, "solve_ms": st_x.ms, "fetch_ms": fetch_ms,
            "Gedges_per_s": g.edges / ((load_ms + st_x.ms + fetch_ms) * 1e-3) / 1e9,
            "note": "load includes the host CSR validation, the reverse positions and, for a graph "
                    "renumbered by degree, the rebuilt CSR; inputs start in pageable host memory"}
    # A graph with skewed degrees or scattered ids is solved on a renumbered
    # copy built by emst_load (its build time is inside pcie_inclusive's
    # load_ms): the same solve in the caller's numbering, for comparison
    layout = {"relabelled": ["none", "degree order", "cluster order"][st_l[-1].relabelled]}
    if st_l[-1].relabelled:
        saved = s.config().as_dict()
        s.set_config(**{**saved, "flags": saved["flags"] | emst.F_NO_RELABEL})
        s.load(g)
        ms_c, st_c = timed_steps(s, emst.FILTER_LEARNED, 0, a.steps, a.warmup)
        layout["learned_ms_caller_numbering"] = ms_c
        layout["speedup"] = ms_c / ms_learned
        assert not st_c[-1].relabelled and np.array_equal(s.fetch(), learned_inmst)
        s.set_config(**saved)
        s.load(g)  # the renumbered copy again, for the solves below
    exact = None
    if not a.no_verify:
        ref = emst.cpu_mst(g)
        d1, cw, gw1 = emst.verify(g, ref, learned_inmst)
        d2, _, gw2 = emst.verify(g, ref, plain_inmst)
        exact = (not d1) and (not d2) and cw == gw1 == gw2
        log(f"verified against the host Kruskal: exact={exact} weight={cw}")
    m_up = int(g.upper_mask().sum())
    stl = st_l[-1]
    n, m = g.nodes, g.edges
    # Algorithmic bytes of the initialisation passes, from the byte model of
    # the reference schedule: nindex 4(n+1), nlist 4m, upper-direction
    # weights 4*m_up, records 16 B; the heavy pass adds one root read per
    # vertex and per heavy upper-direction edge.  The sparse heavy pass does
    # the same job touching less (it skips the marked component's
    # adjacency); it is priced at the same bytes.  An initialisation that also
    # makes its phase's first round of offers is priced with that round too:
    # the reference's first kernel1 reads every record and offers its key to
    # both roots, 16 + 2 * 8 bytes per record.
    heavy_up = m_up - stl.light
    heavy_bytes = 4 * (n + 1) + 4 * m + 4 * m_up + 4 * n + 4 * heavy_up + 16 * stl.heavy
    light_bytes = 4 * (n + 1) + 4 * m + 4 * m_up + 16 * stl.light
    if stl.sparse_heavy:
        heavy_bytes += 32 * stl.heavy
    if stl.light_offers:
        light_bytes += 32 * stl.light
    light_name = "k_init_light_offers<4, false>" if stl.light_offers else "k_init_wl<false, 8, false>"
    heavy_ms = statistics.mean(x.init_heavy_ms for x in st_l)
    light_ms = statistics.mean(x.init_light_ms for x in st_l)
    tr = read_traffic(f"traffic_{a.workload}.json") if a.scale is None else None
    heavy_name = "k_init_heavy_sparse<4>" if stl.sparse_heavy else "k_init_wl<true, 8, false>"
    kernels = {"light": {"kernel": f"{light_name} (light worklist initialisation)", "avg_launch_ms": light_ms,
                         "alg_bytes_per_launch": light_bytes,
                         "achieved": light_bytes / (light_ms * 1e-3) / 1e9,
                         "traffic": tr.get(light_name) if tr else None, "first_offers": bool(stl.light_offers)}}
    if stl.filter_used:
        kernels["heavy"] = {"kernel": f"{heavy_name} (heavy worklist initialisation)", "avg_launch_ms": heavy_ms,
                            "alg_bytes_per_launch": heavy_bytes,
                            "achieved": heavy_bytes / (heavy_ms * 1e-3) / 1e9,
                            "traffic": tr.get(heavy_name) if tr else None, "sparse": bool(stl.sparse_heavy)}
    for k in kernels.values():
        k["frac"] = k["achieved"] / PEAK_HBM_GBS
        if k["traffic"]:
            k["traffic_GBps"] = k["traffic"] / (k["avg_launch_ms"] * 1e-3) / 1e9
    if stl.filter_used and stl.sparse_heavy:
        kernels["heavy"]["note"] = ("algorithmic bytes are the reference pass's, which streams the whole CSR; this "
                                    "pass skips the tasks inside the marked component, so its rate on that model "
                                    "can exceed the HBM peak (traffic_GBps is its measured rate)")
    # the dominant kernel: the initialisation launch that takes longest
    dkey = max(kernels, key=lambda k: kernels[k]["avg_launch_ms"])
    dk = kernels[dkey]
    dom = {"kernel": dk["kernel"], "bound": "hbm", "achieved": dk["achieved"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": dk["frac"], "traffic": dk["traffic"], "alg_bytes_per_launch": dk["alg_bytes_per_launch"],
           "avg_launch_ms": dk["avg_launch_ms"]}
    # Whole-solve byte model of the reference schedule (every logical access
    # once at its natural size) and the compulsory bytes (read the CSR once,
    # write inMST once), both against the median device time.
    mst_edges = int(np.count_nonzero(learned_inmst))
    # the reference schedule keeps parallel records: its worklist sizes come
    # from one more (untimed) solve with that elimination switched off
    cfg0 = s.config().as_dict()
    s.set_config(**{**cfg0, "flags": cfg0["flags"] | emst.F_NO_PAIR_DEDUP})
    st_ref = s.solve(emst.FILTER_LEARNED, 0)
    s.set_config(**cfg0)
    P = 2 if stl.filter_used else 1
    model = 12 * n + m + P * (4 * (n + 1) + 6 * m) + 16 * stl.light
    if stl.filter_used:
        model += 4 * n + 4 * heavy_up + 16 * stl.heavy
    model += 24 * st_ref.round_in_sum + 96 * st_ref.round_out_sum + 5 * mst_edges
    compulsory = 9 * m + 4 * (n + 1)
    t_med = statistics.median(x.ms for x in st_l) * 1e-3
    solve_roof = {"round_records": {"reference_in_out": [st_ref.round_in_sum, st_ref.round_out_sum],
                                    "with_pair_elimination_in_out": [stl.round_in_sum, stl.round_out_sum]},
                  "model_bytes": model, "model_GBps": model / t_med / 1e9, "model_frac": model / t_med / 1e9 / PEAK_HBM_GBS,
                  "compulsory_bytes": compulsory, "compulsory_GBps": compulsory / t_med / 1e9,
                  "compulsory_frac": compulsory / t_med / 1e9 / PEAK_HBM_GBS, "median_device_ms": t_med * 1e3}
    out = {
        "metric": "MST Gedges/s at exact weight (directed CSR edges / device-resident solve), learned filter",
        "value": m / (ms_learned * 1e-3) / 1e9,
        "unit": "Gedges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_learned,
        "median_device_ms": statistics.median(x.ms for x in st_l),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 (weights, ids) / u64 (keys)",
        "data": "synthetic",
        "exact": exact,
        "plain": {"value": m / (ms_plain * 1e-3) / 1e9, "ms_per_step": ms_plain,
                  "median_device_ms": statistics.median(x.ms for x in st_p), "threshold": T, "filter": bool(uf),
                  "light": st_p[-1].light, "heavy": st_p[-1].heavy,
                  "rounds": [st_p[-1].light_rounds, st_p[-1].heavy_rounds]},
        "learned_vs_plain": interleaved_ratio(emst, s, plain_mode, T, max(a.steps, 9)),
        "learned_vs_plain_blocks": ms_plain / ms_learned,
        # the plain rule as the judge of the learned threshold: rounds the
        # learned choice adds (negative: saves), and the edges each choice
        # puts through the light filter and on to the heavy pass
        "learned_judged_by_plain": {
            "extra_rounds": (stl.light_rounds + stl.heavy_rounds) - (st_p[-1].light_rounds + st_p[-1].heavy_rounds),
            "light_edges": {"learned": stl.light, "plain": st_p[-1].light},
            "post_filter_edges": {"learned": stl.heavy, "plain": st_p[-1].heavy}},
        "config": {"workload": w["desc"] if a.scale is None else w["desc"].replace(str(w.get("scale")), str(a.scale), 1),
                   "nodes": n, "edges": m, "learned_q": stl.learned_q, "threshold": stl.threshold_used,
                   "threshold_decided": "once per loaded graph, before the timed solves, for both rules "
                                        "(the reference picks its threshold before gpuMST, ECL_MST.cu:432-441)",
                   "filter": bool(stl.filter_used), "light": stl.light, "heavy": stl.heavy,
                   "rounds": [stl.light_rounds, stl.heavy_rounds], "tail_rounds": stl.tail_rounds,
                   "solver_config": s.config().as_dict()},
        "roofline": dom,
        "init_kernels": kernels,
        "solve_bytes": solve_roof,
        "pcie_inclusive": pcie,
        "layout": layout,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(emst, w, a.cpu_sample_scale, g)
    s.close()
    return out


def batch_graphs(emst, count, seed, nmin, nmax):
    """The batched workload: count graphs, n log-uniform in [nmin, nmax],
    generator uniform over {grid, geometric deg 6, RMAT ef 8, ER deg 10}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))
        kind = int(rng.integers(0, 4))
        gs = int(rng.integers(1, 1 << 31))
        if kind == 0:
            r = max(1, int(math.sqrt(n)))
            out.append(("grid", emst.gen_grid(r, max(1, n // r), gs, 1, 1 << 20)))
        elif kind == 1:
            out.append(("rgg", emst.gen_rgg(n, 6.0, gs, 1 << 20)))
        elif kind == 2:
            out.append(("rmat", emst.gen_rmat(max(1, round(math.log2(n))), 8, 0.57, 0.19, 0.19, gs, 1, 1 << 20)))
        else:
            out.append(("er", emst.gen_er(n, 10.0, gs, 1, 1 << 20)))
        if i % 256 == 0:
            log(f"generated {i}/{count}")
    return out


def _quantiles(x):
    x = np.asarray(x, np.float64)
    return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)), "max": float(x.max())}


def run_batch(a, emst, steps=None, warmup=None):
    """The batched workload (config C4): graphs/s of one batch per step with
    learned thresholds (union network) and with the plain rule per graph,
    per-graph latency from each arena's completion time (arenas run in
    order, every graph of an arena completes with it), and the per-graph
    loop (one device-resident solve per graph, plain rule) timed on every
    graph (or on --loop-sample graphs)."""
    steps = a.steps if steps is None else steps
    warmup = a.warmup if warmup is None else warmup
    t0 = time.perf_counter()
    kg = batch_graphs(emst, a.batch_graphs, 4, a.batch_nmin, a.batch_nmax)
    graphs = [g for _, g in kg]
    tot_n = sum(g.nodes for g in graphs)
    tot_m = sum(g.edges for g in graphs)
    log(f"generated {len(graphs)} graphs, {tot_n} vertices, {tot_m} edges in {time.perf_counter()-t0:.0f}s")
    s = emst.Solver(0)
    s.set_filter_model(np.fromfile(os.path.join(PKG, "filter_mlp.f32"), np.float32))
    b = emst.Batch(s, graphs)
    union_model = os.path.join(PKG, "filter_mlp_union.f32")
    if os.path.exists(union_model):
        # the network trained for graphs that share an arena's rounds
        b.set_filter_model(np.fromfile(union_model, np.float32))

    def timed(mode, k):
        wall, dev, arenas = [], [], []
        for _ in range(k):
            t1 = time.perf_counter()
            st = b.solve(mode, use_graph=True)
            wall.append(time.perf_counter() - t1)
            dev.append(st.ms)
            arenas.append(b.arena_ms())
        return wall, dev, arenas, st
    for _ in range(max(warmup, 1)):   # the first solve also captures each arena's schedule
        b.solve(emst.FILTER_LEARNED, use_graph=True)
    bt, bdev, barenas, st = timed(emst.FILTER_LEARNED, steps)
    batch_s = statistics.median(bt)
    log(f"batch learned: {batch_s*1e3:.1f} ms")
    b.solve(emst.FILTER_PLAIN, use_graph=True)
    pt, pdev, parenas, stp = timed(emst.FILTER_PLAIN, max(1, steps))
    plain_s = statistics.median(pt)
    log(f"batch plain: {plain_s*1e3:.1f} ms")
    # per-graph latency inside the batch: arena k's graphs complete at the
    # sum of the device times of arenas 0..k (median step)
    mid = int(np.argsort(bdev)[len(bdev) // 2])
    ams, first = barenas[mid]
    done = np.cumsum(ams)
    bounds = list(first) + [len(graphs)]
    lat_in_batch = np.concatenate([np.full(bounds[i + 1] - bounds[i], done[i]) for i in range(len(ams))])
    exact = True
    for i in range(0, len(graphs), max(1, len(graphs) // a.verify_graphs)):
        exact = exact and bool(np.array_equal(b.fetch(i), emst.cpu_mst(graphs[i])))
    b.close()
    # per-graph loop (plain rule, one device-resident solve per graph; the
    # upload of each graph is outside the timed region, as in ECL_MST.cu:301-345)
    sample = range(len(graphs)) if a.loop_sample <= 0 else \
        range(0, len(graphs), max(1, len(graphs) // a.loop_sample))
    lat = []
    for i in sample:
        g = graphs[i]
        s.load(g)
        uf, T = emst.plain_threshold(g)
        mode = emst.FILTER_PLAIN if uf else emst.FILTER_OFF
        s.solve(mode, T)
        t1 = time.perf_counter()
        s.solve(mode, T)
        lat.append(time.perf_counter() - t1)
    loop_total_s = statistics.mean(lat) * len(graphs)
    s.close()
    log(f"per-graph loop: {loop_total_s*1e3:.0f} ms for {len(graphs)} graphs ({len(lat)} timed)")
    out = {
        "metric": "batched graphs/s (exact MSF per graph), learned filter, one batch per step",
        "value": len(graphs) / batch_s,
        "unit": "graphs/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": batch_s * 1e3,
        "median_device_ms": statistics.median(bdev),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32 (weights, ids) / u64 (keys)", "data": "synthetic",
        "exact_sampled": exact, "verified_graphs": len(range(0, len(graphs), max(1, len(graphs) // a.verify_graphs))),
        "config": {"workload": f"{len(graphs)} graphs, n log-uniform [{a.batch_nmin}, {a.batch_nmax}], "
                               "grid / geometric deg 6 / RMAT ef 8 / ER deg 10, w in [1,2^20), seed 4",
                   "nodes": tot_n, "edges": tot_m, "arenas": st.chunks, "replays": st.replays,
                   "filtered_graphs": st.filtered_graphs},
        "arena_ms": [round(float(x), 3) for x in ams],
        "per_graph_latency_ms": {**_quantiles(lat_in_batch),
                                 "note": "completion time of the graph's arena from the start of the batch"},
        "plain": {"graphs_per_s": len(graphs) / plain_s, "ms_per_step": plain_s * 1e3,
                  "median_device_ms": statistics.median(pdev), "replays": stp.replays},
        "learned_vs_plain": plain_s / batch_s,
        "network": os.path.basename(union_model) if os.path.exists(union_model) else "filter_mlp.f32",
        "per_graph_loop": {"graphs_per_s": len(graphs) / loop_total_s, "total_s": loop_total_s,
                           **{k + "_ms": v * 1e3 for k, v in _quantiles(lat).items()},
                           "timed_graphs": len(lat)},
        "speedup_vs_loop": loop_total_s / batch_s,
    }
    return out


def run_sub_single(a, emst, name):
    """A compact record of another single-graph configuration (learned vs
    plain, device-resident solves, exactness against the host Kruskal)."""
    w = WORKLOADS[name]
    t0 = time.perf_counter()
    g = make_graph(emst, w)
    log(f"generated {name}: n={g.nodes} m={g.edges} in {time.perf_counter()-t0:.1f}s")
    s = emst.Solver(0)
    s.load(g)
    s.set_filter_model(np.fromfile(os.path.join(PKG, "filter_mlp.f32"), np.float32))
    uf, T = emst.plain_threshold(g)
    plain_mode = emst.FILTER_PLAIN if uf else emst.FILTER_OFF
    ms_l, st_l = timed_steps(s, emst.FILTER_LEARNED, 0, a.steps, a.warmup)
    got_l = s.fetch()
    ms_p, st_p = timed_steps(s, plain_mode, T, a.steps, a.warmup)
    got_p = s.fetch()
    ratio = interleaved_ratio(emst, s, plain_mode, T, max(a.steps, 9))
    ref = emst.cpu_mst(g)
    s.close()

    def rec(ms, sts):
        st = sts[-1]
        return {"Gedges_per_s": g.edges / (ms * 1e-3) / 1e9, "ms_per_step": ms,
                "median_device_ms": statistics.median(x.ms for x in sts), "threshold": st.threshold_used,
                "filter": bool(st.filter_used), "light": st.light, "heavy": st.heavy,
                "rounds": [st.light_rounds, st.heavy_rounds], "init_light_ms": st.init_light_ms,
                "init_heavy_ms": st.init_heavy_ms, "sparse_heavy": bool(st.sparse_heavy),
                "relabelled": ["none", "degree order", "cluster order"][st.relabelled]}
    return {"workload": w["desc"], "nodes": g.nodes, "edges": g.edges,
            "learned": {**rec(ms_l, st_l), "learned_q": st_l[-1].learned_q}, "plain": rec(ms_p, st_p),
            "learned_vs_plain": ratio, "learned_vs_plain_blocks": ms_p / ms_l,
            "exact": bool(np.array_equal(got_l, ref) and np.array_equal(got_p, ref))}


# Inputs of the ladder measurement: the shapes of the reference's inputs
# that the generators reproduce (2d-2e20.sym exactly; rmat16.sym,
# rmat22.sym and r4-2e23.sym in kind and size), with the reference's
# synthetic weights (ECL_MST.cu:410-419).
LADDER_GRAPHS = {
    "2d-2e20": lambda emst: emst.gen_grid(1024, 1024, 1, 1, 2),
    "rmat16": lambda emst: emst.gen_rmat(16, 8, 0.57, 0.19, 0.19, 2, 1, 2),
    "rmat22": lambda emst: emst.gen_rmat(22, 8, 0.57, 0.19, 0.19, 2, 1, 2),
    "r4-2e23": lambda emst: emst.gen_er(1 << 23, 8.0, 4, 1, 2),
}


def run_ladder(a, emst):
    """Every rung of the de-optimization ladder (Deoptimizations/ECL_MST_No*.cu)
    on the ladder inputs: median solve time of up to --steps solves per rung
    (fewer once a rung has spent --ladder-budget seconds), each rung's inMST
    compared with the fused schedule's."""
    s = emst.Solver(0)
    rows = {}
    for name in a.ladder_graphs.split(","):
        g = emst.synth_weights(LADDER_GRAPHS[name](emst))
        log(f"ladder input {name}: n={g.nodes} m={g.edges}")
        s.load(g)
        uf, T = emst.plain_threshold(g)
        rungs = [r for r in [("fused", 0, True)] + list(emst.LADDER)
                 if a.ladder_rungs is None or r[0] in a.ladder_rungs.split(",")]
        ref = None
        row = {}
        for rung, flags, filt in rungs:
            s.set_config(flags=flags)
            mode, thr = (emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T) if filt else (emst.FILTER_OFF, emst.INT_MAX)
            s.solve(mode, thr)
            ts, t0 = [], time.perf_counter()
            while len(ts) < a.steps and (time.perf_counter() - t0 < a.ladder_budget or len(ts) < 1):
                ts.append(s.solve(mode, thr).ms)
            got = s.fetch()
            if ref is None:
                ref = got
            exact = bool(np.array_equal(got, ref))
            ms = float(np.median(ts))
            row[rung] = {"ms": round(ms, 4), "medges_s": round(g.edges / ms / 1e3, 1), "solves": len(ts),
                         "exact": exact}
            log(f"{name} {rung}: {ms:.3f} ms ({len(ts)} solves) exact={exact}")
            if not exact:
                raise SystemExit(f"{name} {rung}: inMST differs from the fused schedule")
        s.set_config()
        if a.ladder_rungs is None or "contract" in a.ladder_rungs.split(","):
            # the related-work baseline on the same input, checked by weight
            if ref is None:
                s.solve(emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T)
                ref = s.fetch()
            s.contract_mst(g)
            ts, t0 = [], time.perf_counter()
            while len(ts) < a.steps and (time.perf_counter() - t0 < a.ladder_budget or len(ts) < 1):
                got_c, st_c = s.contract_mst(g)
                ts.append(st_c.ms)
            same_w = int(st_c.weight) == int(g.eweight[ref.astype(bool)].astype(np.int64).sum())
            ms = float(np.median(ts))
            row["contract"] = {"ms": round(ms, 4), "medges_s": round(g.edges / ms / 1e3, 1), "solves": len(ts),
                               "same_weight": same_w, "iterations": st_c.iterations}
            log(f"{name} contract: {ms:.3f} ms ({len(ts)} solves) same_weight={same_w}")
            if not same_w:
                raise SystemExit(f"{name} contract: forest weight differs")
        if a.ladder_rungs is None or "edgelist" in a.ladder_rungs.split(","):
            # the second related-work baseline: the same (w, j) order, so the
            # same inMST bytes
            if ref is None:
                s.solve(emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T)
                ref = s.fetch()
            s.edgelist_mst(g)
            ts, t0 = [], time.perf_counter()
            while len(ts) < a.steps and (time.perf_counter() - t0 < a.ladder_budget or len(ts) < 1):
                got_e, st_e = s.edgelist_mst(g)
                ts.append(st_e.ms)
            exact = bool(np.array_equal(got_e, ref))
            ms = float(np.median(ts))
            row["edgelist"] = {"ms": round(ms, 4), "medges_s": round(g.edges / ms / 1e3, 1), "solves": len(ts),
                               "exact": exact, "iterations": int(st_e.iterations),
                               "label_passes": int(st_e.label_passes)}
            log(f"{name} edgelist: {ms:.3f} ms ({len(ts)} solves) exact={exact}")
            if not exact:
                raise SystemExit(f"{name} edgelist: inMST differs from the fused schedule")
        rows[name] = {"nodes": g.nodes, "edges": g.edges, "rungs": row}
    s.close()
    return {"metric": "de-optimization ladder: solve time per rung (device events, graph resident)",
            "unit": "ms", "n_gpus": 1, "data": "synthetic", "weights": "ECL_MST.cu:410-419", "inputs": rows}


XGMI_GBS = 7 * 153.0  # seven xGMI links into one MI355X, GB/s each way


def run_shard(a, emst):
    """The single-graph multi-GPU scheme at one GPU's real shape, shard by
    shard on this device: the upper-direction edges of --shard-workload are
    cut into --shards contiguous ranges; each range is loaded and solved as
    one rank would (device-resident, median of --steps, with the threshold
    rule of --shard-filter),
    the positions its forest marks are gathered, and the gathered set is
    solved once more.  The result must equal one solve over the whole graph.
    The implied k-GPU job time is the slowest shard, plus the gather of the
    shard forests (16-byte records at the xGMI rate into one GPU), plus the
    merge solve; nothing here runs on more than one GPU."""
    import shard
    w = WORKLOADS[a.shard_workload]
    g = make_graph(emst, w, a.scale)
    log(f"shard input {a.shard_workload}: n={g.nodes} m={g.edges} k={a.shards}")
    s = emst.Solver(0)
    s.set_filter_model(np.fromfile(os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp.f32"), np.float32))

    def timed(h):
        s.load(h)
        if a.shard_filter == "learned":
            mode, thr = emst.FILTER_LEARNED, 0
        else:
            uf, thr = emst.plain_threshold(h)
            mode = emst.FILTER_PLAIN if uf else emst.FILTER_OFF
        _, st = timed_steps(s, mode, thr, a.steps, a.warmup)
        return s.fetch(), statistics.median(x.ms for x in st)
    full, full_ms = timed(g)
    times = []

    def solve(h):
        mark, ms = timed(h)
        times.append(ms)
        return mark
    got, info = shard.solve_sharded(g, a.shards, solve)
    exact = bool(np.array_equal(got, full))
    shard_ms, merge_ms = times[:-1], times[-1]
    gather_ms = 16 * info["gathered"] / (XGMI_GBS * 1e9) * 1e3
    implied = max(shard_ms) + gather_ms + merge_ms
    log(f"shards {[round(t, 3) for t in shard_ms]} ms, merge {merge_ms:.3f} ms, gathered {info['gathered']}, "
        f"whole graph {full_ms:.3f} ms, exact={exact}")
    s.close()
    if not exact:
        raise SystemExit("sharded forest differs from the whole-graph forest")
    return {"metric": f"edge-sharded solve, shard by shard on one GPU (device-resident, {a.shard_filter} threshold)",
            "unit": "ms", "n_gpus": 1, "data": "synthetic", "workload": w["desc"], "nodes": g.nodes,
            "edges": g.edges, "shards": a.shards, "shard_upper_edges": info["shard_edges"],
            "shard_forest_edges": info["shard_forest_edges"], "gathered": info["gathered"],
            "shard_ms": [round(t, 4) for t in shard_ms], "merge_ms": round(merge_ms, 4),
            "gather_model_ms": round(gather_ms, 4), "implied_k_gpu_ms": round(implied, 4),
            "whole_graph_ms": round(full_ms, 4), "exact": exact,
            "note": "implied_k_gpu_ms = slowest shard + modelled gather + merge; no run used more than one GPU"}


def run_copy(a):
    """Calibration for HBM traffic counters: the HIP runtime's device-to-
    device copy of a known size (read 1 GiB, write 1 GiB per launch).  Run
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to obtain the factor
    between counter readings and bytes moved."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = 1 << 30
    x, y = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(x), nbytes) == 0 and hip.hipMalloc(C.byref(y), nbytes) == 0
    assert hip.hipMemset(x, 1, nbytes) == 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        assert hip.hipMemcpy(y, x, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    dt = (time.perf_counter() - t0) / a.steps
    hip.hipFree(x)
    hip.hipFree(y)
    return {"metric": "device copy GB/s (traffic calibration)", "value": 2 * nbytes / dt / 1e9, "unit": "GB/s",
            "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes, "steps": a.steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat24", choices=sorted(WORKLOADS) + ["batch", "copy", "ladder", "shard"])
    ap.add_argument("--scale", type=int, default=None, help="override the RMAT scale (testing only)")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-scale", type=int, default=0, help="RMAT scale of the CPU sample (0: full workload)")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-nmin", type=int, default=1000)
    ap.add_argument("--batch-nmax", type=int, default=1000000)
    ap.add_argument("--loop-sample", type=int, default=0, help="graphs timed in the per-graph loop (0: all)")
    ap.add_argument("--no-sub-records", action="store_true",
                    help="default workload: leave out the rgg24 and batch sub-records")
    ap.add_argument("--verify-graphs", type=int, default=32)
    ap.add_argument("--ladder-graphs", default=",".join(LADDER_GRAPHS))
    ap.add_argument("--ladder-budget", type=float, default=5.0, help="seconds of solves per rung before stopping")
    ap.add_argument("--ladder-rungs", default=None, help="comma list of rung names (default: every rung)")
    ap.add_argument("--shards", type=int, default=8, help="edge shards of --workload shard")
    ap.add_argument("--shard-workload", default="rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--shard-filter", default="learned", choices=["learned", "plain"])
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the solver runs on one GPU)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "copy":
        print(json.dumps(run_copy(a)))
        return
    import emst
    if a.workload == "ladder":
        out = run_ladder(a, emst)
    elif a.workload == "shard":
        out = run_shard(a, emst)
    elif a.workload == "batch":
        out = run_batch(a, emst)
    else:
        out = run_single(a, emst)
        if a.workload == "rmat24" and a.scale is None and not a.no_sub_records:
            # the rest of the headline metric: the geometric graph of C3
            # (learned vs plain) and the batched configuration C4 (graphs/s)
            out["rgg24"] = run_sub_single(a, emst, "rgg24")
            out["batch"] = run_batch(a, emst, steps=3, warmup=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
