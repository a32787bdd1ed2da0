"""Throughput tables from runs of the ecl_mst CLI, in the shape of the
reference's generate_compare_tables.py (lines 39-63 for the ECL-MST column,
231-253 for the output files).

The CLI prints, per input, "input: <path>" and then 9 run times in seconds
(ECL_MST.cu:404, 345).  Per input the table holds the median of the 9 and
the throughput edges / median / 1e6 (Medges/s).  The edge count comes from
the .egr file's header, not from a fixed list.

  python tables.py run <out_dir> <graph.egr>...    runs bin/ecl_mst on every
      graph, once with the reference's threshold rule (ecl_mst_out.csv) and
      once with the trained network (ecl_mst_learned_out.csv), then writes
      the tables
  python tables.py table <out_dir> <capture.csv>...    tables from captures
      made earlier (each capture becomes one column)

Tables: <out_dir>/comparison_runtimes.csv and comparison_throughputs.csv.
"""
import os
import statistics
import struct
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "bin", "ecl_mst")
RUNS = 9


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


def write_tables(out_dir, columns, edges_of):
    """columns: [(column name, [(input path, times)])]; one row per input of the first column."""
    inputs = [p for p, _ in columns[0][1]]
    runtime = {c: {p: statistics.median(t) for p, t in rows} for c, rows in columns}
    thru = {c: {p: edges_of(p) / runtime[c][p] / 1e6 for p in runtime[c]} for c, _ in columns}
    for fname, table in (("comparison_runtimes.csv", runtime), ("comparison_throughputs.csv", thru)):
        with open(os.path.join(out_dir, fname), "w") as w:
            w.write("," + "".join(f"{c}, " for c, _ in columns))
            for p in inputs:
                name = os.path.basename(p)[:-4] if p.endswith(".egr") else os.path.basename(p)
                w.write("\n" + name + "".join("," + (str(table[c][p]) if p in table[c] else "NC") for c, _ in columns))
            w.write("\n")
    return runtime, thru


def run(out_dir, graphs):
    os.makedirs(out_dir, exist_ok=True)
    columns = []
    for col, env in (("ecl_mst", {}), ("ecl_mst_learned", {"EMST_FILTER": "learned"})):
        text = ""
        for g in graphs:
            r = subprocess.run([CLI, g], stdout=subprocess.PIPE, text=True, env={**os.environ, **env}, timeout=3600)
            if r.returncode != 0:
                raise RuntimeError(f"{CLI} {g} failed with {r.returncode}")
            if "ERROR" in r.stdout:
                raise RuntimeError(f"{g}: the CLI reports a result mismatch")
            text += r.stdout
        with open(os.path.join(out_dir, f"{col}_out.csv"), "w") as f:
            f.write(text)
        columns.append((col, parse_capture(text)))
    return write_tables(out_dir, columns, egr_edges)


def main(argv):
    if len(argv) < 3 or argv[0] not in ("run", "table"):
        sys.stderr.write(__doc__)
        return 1
    out_dir = argv[1]
    if argv[0] == "run":
        run(out_dir, argv[2:])
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
