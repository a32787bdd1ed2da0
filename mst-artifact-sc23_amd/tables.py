"""Throughput tables from runs of the ecl_mst CLI, in the shape of the
reference's generate_compare_tables.py (lines 39-63 for the ECL-MST column,
231-253 for the output files).

The CLI prints, per input, "input: <path>" and then 9 run times in seconds
(ECL_MST.cu:404, 345).  Per input the table holds the median of the 9 and
the throughput edges / median / 1e6 (Medges/s).  The edge count comes from
the .egr file's header, not from a fixed list.

  python tables.py run <out_dir> <graph.egr>...    runs bin/ecl_mst on every
      graph, once with the reference's threshold rule (ecl_mst_out.csv) and
      once with the trained network (ecl_mst_learned_out.csv), and
      the related-work baselines bin/contract_mst (UMinho_gpu_out.csv) and
      bin/edgelist_mst (Jucele_gpu_out.csv), then writes the tables
  python tables.py table <out_dir> <capture.csv>...    tables from captures
      made earlier (each capture becomes one column)
  python tables.py deopt <out_dir> <graph.egr>...    the de-optimization
      ladder (run_all_ECL-MST_deopts.sh and generate_deoptimize_tables.py):
      one capture per rung, ecl_mst_<suffixes>_out.csv, then
      deopt_runtimes.csv and deopt_throughputs.csv

Tables: <out_dir>/comparison_runtimes.csv and comparison_throughputs.csv
(deopt: deopt_runtimes.csv and deopt_throughputs.csv).
"""
import os
import statistics
import struct
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "bin", "ecl_mst")
# the related-work baseline (Boruvka with graph contraction, UMinho GPU)
CONTRACT_CLI = os.path.join(HERE, "bin", "contract_mst")
# the second one (edge-list Boruvka with label propagation, Jucele GPU)
EDGELIST_CLI = os.path.join(HERE, "bin", "edgelist_mst")
RUNS = 9
# Rungs of the ladder by the reference's binary suffixes
# (Deoptimizations/run_all_ECL-MST_deopts.sh:7-15), preceded by the fused
# schedule ("", the default CLI) and the reference's kernel sequence with no
# removal (REF).  The reference's table leaves out NAG_NW
# (generate_deoptimize_tables.py:6-13); this one keeps every rung.
DEOPT_VARIANTS = ["", "REF", "NAG", "NAG_NW", "NAG_NW_NF", "NAG_NW_NF_NIPC", "NAG_NW_NF_NIPC_NOD",
                  "NAG_NW_NF_NIPC_NOD_NT", "NAG_NW_NF_NIPC_NOD_NT_NDD", "NAG_NW_NF_NIPC_NOD_NT_NDD_NEB"]


def egr_edges(path):
    """Directed edge count from an .egr header (int32 nodes, int32 edges)."""
    with open(path, "rb") as f:
        _, edges = struct.unpack("<ii", f.read(8))
    return edges


def parse_capture(text):
    """[(input path, [9 run times])] from the CLI's stdout protocol."""
    out = []
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        if lines[i].startswith("input:"):
            name = lines[i][len("input:"):].strip()
            vals = [float(x) for x in lines[i + 1:i + 1 + RUNS]]
            if len(vals) != RUNS:
                raise ValueError(f"{name}: expected {RUNS} run times, found {len(vals)}")
            out.append((name, vals))
            i += 1 + RUNS
        else:
            i += 1
    return out


def write_tables(out_dir, columns, edges_of, prefix="comparison"):
    """columns: [(column name, [(input path, times)])]; one row per input of the first column."""
    inputs = [p for p, _ in columns[0][1]]
    runtime = {c: {p: statistics.median(t) for p, t in rows} for c, rows in columns}
    thru = {c: {p: edges_of(p) / runtime[c][p] / 1e6 for p in runtime[c]} for c, _ in columns}
    for fname, table in ((f"{prefix}_runtimes.csv", runtime), (f"{prefix}_throughputs.csv", thru)):
        with open(os.path.join(out_dir, fname), "w") as w:
            w.write("," + "".join(f"{c}, " for c, _ in columns))
            for p in inputs:
                name = os.path.basename(p)[:-4] if p.endswith(".egr") else os.path.basename(p)
                w.write("\n" + name + "".join("," + (str(table[c][p]) if p in table[c] else "NC") for c, _ in columns))
            w.write("\n")
    return runtime, thru


def capture(graphs, env, cli=CLI):
    """The CLI's stdout over every graph with extra environment env."""
    text = ""
    for g in graphs:
        r = subprocess.run([cli, g], stdout=subprocess.PIPE, text=True, env={**os.environ, **env}, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(f"{cli} {g} failed with {r.returncode}")
        if "ERROR" in r.stdout:
            raise RuntimeError(f"{g}: the CLI reports a result mismatch")
        text += r.stdout
    return text


def deopt(out_dir, graphs, variants=DEOPT_VARIANTS):
    os.makedirs(out_dir, exist_ok=True)
    columns = []
    for v in variants:
        col = "ecl_mst" + (f"_{v}" if v else "")
        text = capture(graphs, {"EMST_VARIANT": v} if v else {})
        with open(os.path.join(out_dir, f"{col}_out.csv"), "w") as f:
            f.write(text)
        columns.append((col, parse_capture(text)))
    return write_tables(out_dir, columns, egr_edges, prefix="deopt")


def run(out_dir, graphs):
    os.makedirs(out_dir, exist_ok=True)
    columns = []
    for col, env, cli in (("ecl_mst", {}, CLI), ("ecl_mst_learned", {"EMST_FILTER": "learned"}, CLI),
                          ("UMinho_gpu", {}, CONTRACT_CLI), ("Jucele_gpu", {}, EDGELIST_CLI)):
        text = capture(graphs, env, cli)
        with open(os.path.join(out_dir, f"{col}_out.csv"), "w") as f:
            f.write(text)
        columns.append((col, parse_capture(text)))
    return write_tables(out_dir, columns, egr_edges)


def main(argv):
    if len(argv) < 3 or argv[0] not in ("run", "table", "deopt"):
        sys.stderr.write(__doc__)
        return 1
    out_dir = argv[1]
    if argv[0] == "run":
        run(out_dir, argv[2:])
    elif argv[0] == "deopt":
        deopt(out_dir, argv[2:])
        print(f"wrote {out_dir}/deopt_runtimes.csv and {out_dir}/deopt_throughputs.csv")
        return 0
    else:
        os.makedirs(out_dir, exist_ok=True)
        cols = []
        for cap in argv[2:]:
            with open(cap) as f:
                cols.append((os.path.basename(cap).replace("_out.csv", ""), parse_capture(f.read())))
        write_tables(out_dir, cols, egr_edges)
    print(f"wrote {out_dir}/comparison_runtimes.csv and {out_dir}/comparison_throughputs.csv")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
