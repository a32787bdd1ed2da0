"""Trains the light-edge threshold network on graphs the solver itself
solves (needs a HIP device).

1. Generates N seeded graphs: 2D grids, random geometric graphs and RMAT,
   2^12..2^20 vertices (log-uniform), mixed degrees and weight ranges, and
   optionally --large more of 2^21..2^22.5 vertices for the single-solve
   network (48 of them helped geometric and Erdos-Renyi graphs far above
   2^20 vertices but moved the 512x512 grid from the unfiltered path to a
   slower filtered one, so the shipped network is trained without them).
   The union network also sees Erdos-Renyi graphs (--union-er of them):
   the batched workload mixes them in, and a network that never saw one
   chose q near 0.5 where about 0.2 is best.
2. For each graph: device features and weight samples, then one solve per
   candidate quantile q (threshold = sorted sample of rank floor(q*1023);
   the last candidate is the unfiltered path).  Each candidate's cost is
   modelled from the solve's own exact round schedule:
       cost = launch_us * rounds + init_bytes / bw_init + round_bytes / bw_round
   where init_bytes counts the setup and initialisation passes and their
   records, and round_bytes 24 B per record entering a round and 96 B per
   survivor.  The label is log2 of the cheapest q.  Two cost models give
   two label sets:
     single  LAUNCH_US per round, every byte at BW: one graph per solve,
             so every round's fixed cost is the graph's own;
     union   LAUNCH_US / UNION_SHARE per round: in a batched arena about
             UNION_SHARE graphs share every round, so a graph's marginal
             cost is its bytes; round bytes are taken at UNION_ROUND_BW
             (slower rates, 0.3-2 TB/s, left the labels almost unchanged
             and did not make the batch faster).
3. Full-batch Adam on the device (emst_mlp_train), one network per cost
   model, then an evaluation on held-out graphs against the plain rule
   (ECL_MST.cu:432-443) under the same cost model.
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
LAUNCH_US = 8.0
UNION_SHARE = 1000.0
BW = 3.0e12
UNION_ROUND_BW = 3.0e12


class CostModel:
    def __init__(self, launch_us, bw_init, bw_round):
        self.launch_us, self.bw_init, self.bw_round = launch_us, bw_init, bw_round

    def cost(self, rounds, init_bytes, round_bytes):
        return self.launch_us * rounds + (init_bytes / self.bw_init + round_bytes / self.bw_round) * 1e6

    def as_dict(self):
        return {"launch_us": self.launch_us, "bw_init": self.bw_init, "bw_round": self.bw_round}


MODELS = {"single": CostModel(LAUNCH_US, BW, BW),
          "union": CostModel(LAUNCH_US / UNION_SHARE, BW, UNION_ROUND_BW)}


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
    ap.add_argument("--large", type=int, default=0,
                    help="graphs of 2^21-2^22.5 vertices added for the single-solve network")
    a = ap.parse_args()
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
    total = a.graphs + a.union_er + a.large
    for i in range(len(X), total):
        if i >= a.graphs + a.union_er:
            kind, g = make_graph(i, a.seed, None, (21, 22.5))
        else:
            kind, g = make_graph(i, a.seed, 3 if i >= a.graphs else None)
        feat, rows = label_graph(solver, g)
        X.append(feat)
        R.append([r for r, _, _ in rows])
        BI.append([bi for _, bi, _ in rows])
        BR.append([br for _, _, br in rows])
        kinds.append(kind)
        if i % 50 == 0:
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
    # The single network trains on the grid, geometric and RMAT rows (the
    # first --graphs and the --large ones); the union network on the first
    # --graphs and the Erdos-Renyi rows.
    base = list(range(a.graphs))
    er_rows = list(range(a.graphs, a.graphs + a.union_er))
    large_rows = list(range(a.graphs + a.union_er, total))
    for name, fname, rows, er_every in (("single", "filter_mlp.f32", base + large_rows, 0),
                                        ("union", "filter_mlp_union.f32", base + er_rows, 4 if a.union_er else 0)):
        model = MODELS[name]
        rows_used = len(rows)
        qs = [best_q(list(zip(R[i], BI[i], BR[i])), model) for i in rows]
        Y = np.log2(np.array(qs, np.float32))
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
