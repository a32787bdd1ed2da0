"""Writes tests/golden/schedules.json: the oracle's round schedule of the
full-size configurations at fixed thresholds.

The schedule of the edge-based Boruvka rounds is deterministic (components
after a round are the components of the previous ones joined by all their
minimum edges; representatives are minimum ids), so the counts a solve
reports are fixed by the input and the threshold:
  L        light records written by the first initialisation
  H_all    upper-direction edges with w > T
  H        heavy records (upper-direction edges crossing components after
           the light phase)
  rounds   kernel1 rounds per phase
  in/out   worklist sizes entering and leaving every kernel1 round
  components, mst_edges, weight (64-bit wrapping sum)
Each entry also holds the SHA-256 of the generated graph, so a generator
change shows up as a digest mismatch rather than a schedule mismatch.

Run (CPU only, a few minutes; about 20 GB of host memory):
    python tests/golden/make_schedules.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "mst-artifact-sc23_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import emst  # noqa: E402
import oracle_lib as O  # noqa: E402

# name -> (generator, thresholds); None means the plain rule (ECL_MST.cu:432-443)
CONFIGS = {
    "grid512": (lambda: emst.gen_grid(512, 512, 1, 1, 1 << 20), [None, 1 << 18]),
    "rmat24": (lambda: emst.gen_rmat(24, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24), [None, 515806]),
    "rgg24": (lambda: emst.gen_rgg(1 << 24, 6.0, 3, 1 << 24), [None]),
}


def digest(g):
    h = hashlib.sha256()
    for a in (g.nindex, g.nlist, g.eweight):
        h.update(a.tobytes())
    return h.hexdigest()


def schedule_entry(g, use_filter, T):
    _, st = O.schedule(g, use_filter, T)
    return {
        "use_filter": int(use_filter), "threshold": int(T),
        "L": int(st.light), "H_all": int(st.heavy_all), "H": int(st.heavy),
        "light_rounds": int(st.light_rounds), "heavy_rounds": int(st.heavy_rounds),
        "light_in": [int(st.light_w[i]) for i in range(st.light_rounds)],
        "light_out": [int(st.light_s[i]) for i in range(st.light_rounds)],
        "heavy_in": [int(st.heavy_w[i]) for i in range(st.heavy_rounds)],
        "heavy_out": [int(st.heavy_s[i]) for i in range(st.heavy_rounds)],
        "components": int(st.components), "mst_edges": int(st.mst_edges), "weight": int(st.weight),
    }


def main(names):
    path = os.path.join(HERE, "schedules.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for name in names:
        gen, thresholds = CONFIGS[name]
        t0 = time.time()
        g = gen()
        ent = {"nodes": g.nodes, "edges": g.edges, "sha256": digest(g), "runs": []}
        for T in thresholds:
            if T is None:
                uf, T = emst.plain_threshold(g)
            else:
                uf = 1
            ent["runs"].append(schedule_entry(g, uf, T))
        out[name] = ent
        print(f"{name}: {time.time() - t0:.0f}s {json.dumps(ent['runs'])[:300]}", flush=True)
        del g
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CONFIGS))
