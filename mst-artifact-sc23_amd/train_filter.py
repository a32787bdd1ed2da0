"""Trains the light-edge threshold network on graphs the solver itself
solves (needs a HIP device).

1. Generates N seeded graphs: 2D grids, random geometric graphs and RMAT,
   2^12..2^20 vertices (log-uniform), mixed degrees and weight ranges;
   --union-er Erdos-Renyi graphs of the same sizes (the batched workload
   mixes them in, and a network that never saw one chose q near 0.5 where
   about 0.2 is best); and --large grids, geometric and RMAT graphs and
   --large-er Erdos-Renyi graphs of --large-range vertices.  The network extrapolates the trend of its
   labels in log2(n) linearly: trained on 2^12..2^20 only it chose q = 0.22
   on the geometric graph of 2^24 vertices, whose measured optimum is 0.5,
   so the shipped network also sees 2^21..2^24.
2. For each graph: device features and weight samples, then one solve per
   candidate quantile q (threshold = sorted sample of rank floor(q*1023);
   the last candidate is the unfiltered path).  Each candidate's cost is
   modelled from the solve's own exact round schedule:
       cost = launch_us * rounds + init_bytes / bw_init + round_bytes / bw_round
   where init_bytes counts the setup and initialisation passes and their
   records, and round_bytes 24 B per record entering a round and 96 B per
   survivor.  The label is log2 of the cheapest q; the unfiltered path is
   labelled OFF_LABEL, so that the midpoint between it and q = 0.5 lies at
   the device's cut (q >= 0.9 selects the unfiltered path) and a prediction
   between the two modes resolves to the nearer one.  Two cost models give
   two label sets:
     single  SINGLE_LAUNCH_US per round, initialisation bytes at
             SINGLE_BW_INIT, round bytes at SINGLE_BW_ROUND: one graph per
             solve, so every round's fixed cost is the graph's own.  The
             constants come from device-timed solves of 18 graphs (grids,
             geometric, RMAT and Erdos-Renyi, 2^16..2^26 vertices) at every
             candidate: they minimise the measured time lost at the model's
             choice (geomean 0.8% over the per-graph optimum, worst 10%;
             8 us per round with 3 TB/s lost 3.3%, worst 28%, choosing
             filtered paths where one round costs more than 8 us);
     union   UNION_LAUNCH_US per round: in a batched arena about a thousand
             graphs share every round, so a graph's marginal cost is its
             bytes, taken at BW (slower rates, 0.3-2 TB/s, left the labels
             almost unchanged and did not make the batch faster).
3. Full-batch Adam on the device (emst_mlp_train), one network per cost
   model (the single network on every row; the union network on the
   first --graphs rows and the Erdos-Renyi rows), then an evaluation on
   held-out graphs against the plain rule (ECL_MST.cu:432-443) under the
   same cost model.
Writes <out>/filter_mlp.f32, <out>/filter_mlp_union.f32 (fp32 parameters),
<out>/filter_costs.npz (features and the per-candidate rounds and bytes of
every training graph) and <out>/filter_mlp.json.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import emst  # noqa: E402

QS = [2.0 ** (-k / 2.0) for k in range(14, 0, -1)] + [1.0]   # 2^-7 .. 2^-0.5, then unfiltered
SINGLE_LAUNCH_US = 32.0
SINGLE_BW_INIT = 4.0e12
SINGLE_BW_ROUND = 3.0e12
UNION_LAUNCH_US = 8.0 / 1000.0
BW = 3.0e12
OFF_LABEL = 0.7


class CostModel:
    def __init__(self, launch_us, bw_init, bw_round):
        self.launch_us, self.bw_init, self.bw_round = launch_us, bw_init, bw_round

    def cost(self, rounds, init_bytes, round_bytes):
        return self.launch_us * rounds + (init_bytes / self.bw_init + round_bytes / self.bw_round) * 1e6

    def as_dict(self):
        return {"launch_us": self.launch_us, "bw_init": self.bw_init, "bw_round": self.bw_round}


MODELS = {"single": CostModel(SINGLE_LAUNCH_US, SINGLE_BW_INIT, SINGLE_BW_ROUND),
          "union": CostModel(UNION_LAUNCH_US, BW, BW)}


def make_graph(i, seed, kind=None, logn_range=(12, 20)):
    """Graph i of the seeded mix: kind cycles grid, geometric, RMAT unless
    given (3 selects Erdos-Renyi); log2 of the vertex count is uniform in
    logn_range."""
    rng = np.random.default_rng([seed, i])
    kind = i % 3 if kind is None else kind
    logn = rng.uniform(*logn_range)
    n = int(2 ** logn)
    wbits = int(rng.choice([6, 10, 16, 20, 24]))
    gseed = int(rng.integers(1, 2**31))
    if kind == 0:
        aspect = rng.uniform(0.5, 2.0)
        r = max(2, int(math.sqrt(n * aspect)))
        c = max(2, n // r)
        return "grid", emst.gen_grid(r, c, gseed, 1, 1 << wbits)
    if kind == 1:
        deg = float(rng.choice([3.0, 4.0, 6.0, 8.0, 12.0]))
        return "rgg", emst.gen_rgg(n, deg, gseed, 1 << max(wbits, 8))
    if kind == 3:
        deg = float(rng.choice([4.0, 6.0, 10.0, 16.0]))
        return "er", emst.gen_er(n, deg, gseed, 1, 1 << wbits)
    scale = int(round(logn))
    ef = int(rng.choice([4, 8, 16]))
    a = float(rng.uniform(0.45, 0.6))
    b = c = (1 - a) * float(rng.uniform(0.3, 0.45))
    return "rmat", emst.gen_rmat(scale, ef, a, b, c, gseed, 1, 1 << wbits)


def schedule_bytes(g, st, m_up_heavy):
    """(initialisation bytes, round bytes) of a solve's schedule."""
    n, m = g.nodes, g.edges
    P = 2 if st.filter_used else 1
    b = 12 * n + m + P * (4 * (n + 1) + 6 * m) + 16 * st.light + 5 * st.mst_edges
    if st.filter_used:
        b += 4 * n + 4 * m_up_heavy + 16 * st.heavy
    return b, 24 * st.round_in_sum + 96 * st.round_out_sum


def cost_of(g, st, m_up_heavy, model):
    return model.cost(st.light_rounds + st.heavy_rounds, *schedule_bytes(g, st, m_up_heavy))


def label_graph(solver, g):
    solver.load(g)
    _, _, _, feat, samples = solver.learned_threshold()
    upw = np.sort(g.eweight[g.upper_mask()])
    rows = []
    for q in QS:
        if q >= 1.0:
            st = solver.solve(emst.FILTER_OFF, emst.INT_MAX)
            heavy_up = 0
        else:
            T = int(samples[int(math.floor(q * 1023))])
            st = solver.solve(emst.FILTER_PLAIN, T)
            heavy_up = int(len(upw) - np.searchsorted(upw, T, side="right"))
        rows.append((st.light_rounds + st.heavy_rounds, *schedule_bytes(g, st, heavy_up)))
    return feat, rows


def best_q(rows, model):
    costs = [model.cost(r, bi, br) for r, bi, br in rows]
    return QS[int(np.argmin(costs))]


def label_of(qs, name):
    """Regression targets: log2 q, with the unfiltered path at OFF_LABEL
    for the single network.  The union network's targets stay log2 q
    (unfiltered = 0), the targets its shipped file was trained on."""
    y = np.log2(qs)
    if name == "single":
        y = np.where(qs >= 1.0, OFF_LABEL, y)
    return y.astype(np.float32)


def plain_cost(solver, g, model):
    uf, T = emst.plain_threshold(g)
    st = solver.solve(emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T)
    upw = np.sort(g.eweight[g.upper_mask()])
    hu = int(len(upw) - np.searchsorted(upw, T, side="right")) if uf else 0
    return cost_of(g, st, hu, model), st


def evaluate(solver, params, model, holdout, seed, er_every=0):
    """Modelled cost of the learned and the plain thresholds on held-out graphs
    (every er_every-th one Erdos-Renyi when er_every > 0)."""
    solver.set_filter_model(params)
    ev = []
    for i in range(holdout):
        er = er_every > 0 and i % er_every == er_every - 1
        kind, g = make_graph(10_000_000 + i, seed + 1, 3 if er else None)
        solver.load(g)
        pc, pst = plain_cost(solver, g, model)
        lst = solver.solve(emst.FILTER_LEARNED, 0)
        upw = np.sort(g.eweight[g.upper_mask()])
        hu = int(len(upw) - np.searchsorted(upw, lst.threshold_used, side="right")) if lst.filter_used else 0
        lc = cost_of(g, lst, hu, model)
        ev.append({"kind": kind, "n": g.nodes, "m": g.edges, "plain_cost": pc, "learned_cost": lc,
                   "plain_rounds": pst.light_rounds + pst.heavy_rounds,
                   "learned_rounds": lst.light_rounds + lst.heavy_rounds, "plain_light": pst.light,
                   "learned_light": lst.light, "q": lst.learned_q})
    ratio = float(np.exp(np.mean([math.log(e["learned_cost"] / e["plain_cost"]) for e in ev])))
    return ratio, ev


# Graphs of the cost-model calibration (--calibrate): name, generator.
CALIB_GRAPHS = [
    ("grid 256x256", lambda: emst.gen_grid(256, 256, 24, 1, 1 << 20)),
    ("grid 512x512", lambda: emst.gen_grid(512, 512, 1, 1, 1 << 20)),
    ("grid 1024x1024", lambda: emst.gen_grid(1024, 1024, 14, 1, 1 << 20)),
    ("grid 2048x2048 w<2^16", lambda: emst.gen_grid(2048, 2048, 13, 1, 1 << 16)),
    ("grid 4096x4096", lambda: emst.gen_grid(4096, 4096, 12, 1, 1 << 20)),
    ("rgg 2^20 deg 6", lambda: emst.gen_rgg(1 << 20, 6.0, 21, 1 << 24)),
    ("rgg 2^22 deg 6 w<2^20", lambda: emst.gen_rgg(1 << 22, 6.0, 7, 1 << 20)),
    ("rgg 2^23 deg 4 w<2^16", lambda: emst.gen_rgg(1 << 23, 4.0, 8, 1 << 16)),
    ("rgg 2^23 deg 8", lambda: emst.gen_rgg(1 << 23, 8.0, 9, 1 << 24)),
    ("rgg 2^24 deg 6", lambda: emst.gen_rgg(1 << 24, 6.0, 3, 1 << 24)),
    ("er 2^20 deg 8", lambda: emst.gen_er(1 << 20, 8.0, 23, 1, 1 << 20)),
    ("er 2^22 deg 10", lambda: emst.gen_er(1 << 22, 10.0, 10, 1, 1 << 20)),
    ("rmat 20 ef 16", lambda: emst.gen_rmat(20, 16, 0.57, 0.19, 0.19, 22, 1, 1 << 24)),
    ("rmat 22 ef 4 w<2^10", lambda: emst.gen_rmat(22, 4, 0.55, 0.2, 0.2, 17, 1, 1 << 10)),
    ("rmat 22 ef 16 w<2^20", lambda: emst.gen_rmat(22, 16, 0.57, 0.19, 0.19, 15, 1, 1 << 20)),
    ("rmat 23 ef 8 (0.5,0.2,0.2)", lambda: emst.gen_rmat(23, 8, 0.5, 0.2, 0.2, 16, 1, 1 << 24)),
    ("rmat 24 ef 16", lambda: emst.gen_rmat(24, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)),
    ("rmat 26 ef 16 w<=64", lambda: emst.gen_rmat(26, 16, 0.57, 0.19, 0.19, 5, 1, 65)),
]


def calibrate(path, graphs, reps=3):
    """Device-timed solves of the calibration graphs at every candidate
    (median of reps, default configuration), with the reference schedule's
    round and record counts of each candidate (parallel records kept, the
    schedule the labels price); the plain rule and the shipped network are
    timed beside them.  Then the single cost model's constants are searched
    for the smallest measured time lost at the model's choice.  Writes the
    measurements to path (JSON) and prints a table."""
    solver = emst.Solver(0)
    params = np.fromfile(os.path.join(os.path.dirname(os.path.abspath(__file__)), "filter_mlp.f32"), np.float32)
    out = {}
    for name, mk in CALIB_GRAPHS:
        if graphs and name not in graphs:
            continue
        t0 = time.time()
        g = mk()
        solver.set_config(flags=0)
        solver.set_filter_model(params)
        solver.load(g)
        _, _, _, feat, samples = solver.learned_threshold()
        upw = np.sort(g.eweight[g.upper_mask()])

        def timed(mode, thr):
            solver.solve(mode, thr)
            sts = [solver.solve(mode, thr) for _ in range(reps)]
            return float(np.median([x.ms for x in sts])), sts[-1]

        cands = [(emst.FILTER_OFF, emst.INT_MAX) if q >= 1.0 else
                 (emst.FILTER_PLAIN, int(samples[int(math.floor(q * 1023))])) for q in QS]
        rows = [{"q": q, "T": thr, "ms": timed(mode, thr)[0]} for q, (mode, thr) in zip(QS, cands)]
        uf, pt = emst.plain_threshold(g)
        plain_ms, _ = timed(emst.FILTER_PLAIN if uf else emst.FILTER_OFF, pt)
        learned_ms, lst = timed(emst.FILTER_LEARNED, 0)
        solver.set_config(flags=emst.F_NO_PAIR_DEDUP)
        for r, (mode, thr) in zip(rows, cands):
            st = solver.solve(mode, thr)
            hu = int(len(upw) - np.searchsorted(upw, thr, side="right")) if mode == emst.FILTER_PLAIN else 0
            r["rounds"] = st.light_rounds + st.heavy_rounds
            r["init_bytes"], r["round_bytes"] = schedule_bytes(g, st, hu)
        out[name] = {"n": g.nodes, "m": g.edges, "feat": [float(x) for x in feat], "rows": rows,
                     "plain_ms": plain_ms, "learned_ms": learned_ms, "learned_q": float(lst.learned_q)}
        best = min(rows, key=lambda r: r["ms"])
        print(f"[calibrate] {name}: n={g.nodes} m={g.edges} best q={best['q']:.3f} {best['ms']:.3f} ms, "
              f"plain {plain_ms:.3f} ms, shipped network q={lst.learned_q:.3f} {learned_ms:.3f} ms "
              f"({time.time() - t0:.0f} s)", flush=True)
        print("    " + " ".join(f"{r['q']:.3f}:{r['ms']:.3f}" for r in rows), flush=True)
        with open(path, "w") as f:
            json.dump(out, f)
        del g
    solver.close()
    fit_constants(out)


def model_regret(calib, launch_us, bw_init, bw_round):
    """Measured time at the model's choice over the best candidate's, per
    graph, for the single-solve cost model with these constants."""
    m = CostModel(launch_us, bw_init, bw_round)
    reg = {}
    for name, g in calib.items():
        costs = [m.cost(r["rounds"], r["init_bytes"], r["round_bytes"]) for r in g["rows"]]
        ms = [r["ms"] for r in g["rows"]]
        reg[name] = ms[int(np.argmin(costs))] / min(ms)
    return reg


def fit_constants(calib):
    grid = [(lu, bi, br) for lu in (4, 8, 12, 16, 20, 24, 32, 40, 48, 64)
            for bi in (1e12, 2e12, 3e12, 4e12, 6e12) for br in (1e12, 2e12, 3e12, 4e12, 6e12)]
    scored = []
    for c in grid:
        reg = model_regret(calib, *c)
        scored.append((float(np.exp(np.mean(np.log(list(reg.values()))))), max(reg.values()), c))
    scored.sort(key=lambda x: (round(x[0], 6), x[1]))
    print("[calibrate] cost-model constants by measured regret (geomean, worst):")
    for gm, mx, c in scored[:8]:
        print(f"    launch {c[0]:>4} us  init {c[1] / 1e12:.0f} TB/s  rounds {c[2] / 1e12:.0f} TB/s: {gm:.4f} {mx:.3f}")
    for label, c in (("shipped", (SINGLE_LAUNCH_US, SINGLE_BW_INIT, SINGLE_BW_ROUND)), ("8 us, 3 TB/s", (8.0, 3e12, 3e12))):
        reg = model_regret(calib, *c)
        print(f"[calibrate] {label}: geomean {np.exp(np.mean(np.log(list(reg.values())))):.4f}, worst "
              f"{max(reg.values()):.3f} ({max(reg, key=reg.get)})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=2000)
    ap.add_argument("--holdout", type=int, default=60)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="filter_train", help="directory for the trained networks and labels")
    ap.add_argument("--costs", default=None, help="reuse a filter_costs.npz instead of solving the graphs again")
    ap.add_argument("--union-er", type=int, default=667, help="Erdos-Renyi graphs added for the union network")
    ap.add_argument("--large", type=int, default=240,
                    help="graphs of --large-range vertices added for the single-solve network")
    ap.add_argument("--large-range", default="21,24",
                    help="log2 vertex range of the --large graphs")
    ap.add_argument("--large-er", type=int, default=60,
                    help="Erdos-Renyi graphs of --large-range vertices added for the single-solve network")
    ap.add_argument("--calibrate", default=None, metavar="JSON",
                    help="time every candidate on the calibration graphs, write JSON, fit the cost-model constants")
    ap.add_argument("--calibrate-graphs", default="", help="comma list of calibration graph names (default: all)")
    a = ap.parse_args()
    if a.calibrate:
        calibrate(a.calibrate, [x for x in a.calibrate_graphs.split(",") if x])
        return
    os.makedirs(a.out, exist_ok=True)
    solver = emst.Solver(0)
    # Labels price the reference schedule's worklists: parallel-record
    # elimination only trims them, and switching it off keeps the labels
    # (and so the networks) independent of that optimisation.
    solver.set_config(flags=emst.F_NO_PAIR_DEDUP)
    t0 = time.time()
    X, R, BI, BR, kinds = [], [], [], [], []
    if a.costs:
        # Rows already labelled (the first graphs of the same seeded mix);
        # only the missing ones are solved.
        z = np.load(a.costs)
        X, R, BI, BR = (list(z[k]) for k in ("X", "rounds", "init_bytes", "round_bytes"))
        kinds = [str(k) for k in z["kinds"]]
    total = a.graphs + a.union_er + a.large + a.large_er
    large_range = tuple(float(x) for x in a.large_range.split(","))
    for i in range(len(X), total):
        if i >= a.graphs + a.union_er + a.large:
            kind, g = make_graph(i, a.seed, 3, large_range)
        elif i >= a.graphs + a.union_er:
            kind, g = make_graph(i, a.seed, None, large_range)
        else:
            kind, g = make_graph(i, a.seed, 3 if i >= a.graphs else None)
        feat, rows = label_graph(solver, g)
        X.append(feat)
        R.append([r for r, _, _ in rows])
        BI.append([bi for _, bi, _ in rows])
        BR.append([br for _, _, br in rows])
        kinds.append(kind)
        if i % 50 == 0 or i >= a.graphs + a.union_er:
            print(f"[train_filter] labelled {i}/{total} ({time.time()-t0:.0f}s) last {kind} n={g.nodes}", flush=True)
    X = np.array(X[:total], np.float32)
    R = np.array(R[:total], np.int64)
    BI = np.array(BI[:total], np.float64)
    BR = np.array(BR[:total], np.float64)
    kinds = kinds[:total]
    np.savez(os.path.join(a.out, "filter_costs.npz"), X=X, rounds=R, init_bytes=BI, round_bytes=BR,
             kinds=np.array(kinds), qs=np.array(QS))
    info = {"params": emst.MLP_PARAMS, "graphs": int(len(X)), "seed": a.seed, "epochs": a.epochs, "lr": a.lr,
            "candidates": QS, "models": {}}
    # The single network trains on every row; the union network on the
    # first --graphs and the Erdos-Renyi rows.
    base = list(range(a.graphs))
    er_rows = list(range(a.graphs, a.graphs + a.union_er))
    large_rows = list(range(a.graphs + a.union_er, total))
    for name, fname, rows, er_every in (("single", "filter_mlp.f32", base + er_rows + large_rows, 4 if a.union_er else 0),
                                        ("union", "filter_mlp_union.f32", base + er_rows, 4 if a.union_er else 0)):
        model = MODELS[name]
        rows_used = len(rows)
        qs = [best_q(list(zip(R[i], BI[i], BR[i])), model) for i in rows]
        Y = label_of(np.array(qs, np.float64), name)
        p0 = emst.init_params(a.seed)
        t1 = time.time()
        params, loss = emst.mlp_train(p0, X[rows], Y, a.epochs, a.lr, a.seed)
        print(f"[train_filter] {name}: trained {a.epochs} epochs in {time.time()-t1:.1f}s, final mse {loss:.4f} "
              f"(label variance {float(np.var(Y)):.4f})", flush=True)
        params.astype(np.float32).tofile(os.path.join(a.out, fname))
        ratio, ev = evaluate(solver, params, model, a.holdout, a.seed, er_every)
        print(f"[train_filter] {name}: held-out geomean learned/plain modelled cost = {ratio:.3f}", flush=True)
        info["models"][name] = {"file": fname, "graphs": rows_used,
                                "train_kinds": {k: [kinds[i] for i in rows].count(k) for k in sorted(set(kinds[i] for i in rows))},
                                **model.as_dict(), "final_mse": loss,
                                "label_variance": float(np.var(Y)), "holdout_geomean_cost_ratio": ratio,
                                "holdout": ev, "label_q_histogram": {str(q): int(sum(1 for x in qs if x == q))
                                                                     for q in QS}}
    info["seconds"] = time.time() - t0
    with open(os.path.join(a.out, "filter_mlp.json"), "w") as f:
        json.dump(info, f)


if __name__ == "__main__":
    main()
