"""Host-side tests of libemst.so (no GPU): exported C-ABI, .egr reader and
writer, weight synthesis, plain threshold, host Kruskal, validation, the
workload generators, and the CLI's pre-device error paths."""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import emst
import oracle_lib as O
from graphs import csr_from_edges, random_graph
from test_oracle import egr_fixtures

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "mst-artifact-sc23_amd", "bin", "ecl_mst")
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
GEN_HASHES = json.load(open(os.path.join(HERE, "golden", "gen_hashes.json")))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(emst_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", emst.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (emst_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(emst.lib, n)


def test_flag_constants_match_header():
    text = open(os.path.join(ROOT, "include", "emst.h")).read()
    flags = {k: int(v) for k, v in re.findall(r"#define EMST_(F_\w+)\s+(\d+)", text)}
    assert len(flags) == 16
    for k, v in flags.items():
        assert getattr(emst, k) == v, k
    assert len(set(flags.values())) == len(flags) and all(v & (v - 1) == 0 for v in flags.values())


def test_ladder_rungs_remove_one_optimization_each():
    """The rung table follows the reference's variant chain
    (Deoptimizations/ECL_MST_NoAtomicGuards_NoWarp_..._NoEdgeBased.cu):
    each rung keeps the previous rung's removals and adds one."""
    names = [n for n, _, _ in emst.LADDER]
    assert names == ["ECL-MST", "NoAtomicGuards", "NoWarp", "NoFilter", "NoImplicitPathCompression",
                     "NoOneDirection", "NoTuples", "NoDataDriven", "NoEdgeBased"]
    for (_, f0, filt0), (name, f1, filt1) in zip(emst.LADDER, emst.LADDER[1:]):
        assert f1 & f0 == f0, name
        removed = bin(f1 ^ f0).count("1") + (filt0 != filt1)
        assert removed == 1, name
    assert emst.LADDER[0][1] == emst.F_LADDER and emst.LADDER[0][2]


ORC_TO_EMST = {0: 0, -1: -10, -2: -11, -3: -12, -4: -13, -6: -14, -7: -15, -8: -16}


def test_reader_matches_oracle_on_fixtures(tmp_path):
    for name, path, rc in egr_fixtures(str(tmp_path)):
        if rc != 0:
            with pytest.raises(emst.EmstError) as e:
                emst.read_egr(path)
            assert e.value.code == ORC_TO_EMST[rc], name
            continue
        g = emst.read_egr(path)
        _, ni, nl, ew = O.read_egr(path)
        assert np.array_equal(g.nindex, ni) and np.array_equal(g.nlist, nl), name
        assert (g.eweight is None) == (ew is None), name
        if ew is not None:
            assert np.array_equal(g.eweight, ew), name


def test_write_read_roundtrip(tmp_path):
    g = emst.gen_rmat(10, 8, seed=3)
    p = str(tmp_path / "g.egr")
    emst.write_egr(p, g)
    assert os.path.getsize(p) == 8 + 4 * (g.nodes + 1) + 8 * g.edges
    h = emst.read_egr(p)
    assert np.array_equal(h.nindex, g.nindex) and np.array_equal(h.nlist, g.nlist)
    assert np.array_equal(h.eweight, g.eweight)
    # weightless write: the reader then reports no weights
    emst.write_egr(p, emst.Graph(g.nindex, g.nlist, None))
    assert emst.read_egr(p).eweight is None
    with pytest.raises(emst.EmstError):
        emst.write_egr(str(tmp_path / "nodir" / "x.egr"), g)


def test_synth_weights_match_oracle_and_golden(tmp_path):
    g = emst.gen_rmat(11, 8, seed=4)
    s = emst.synth_weights(g)
    assert np.array_equal(s.eweight, O.synth_weights(g.nindex, g.nlist))
    for i, k, n, w in GOLD["synth"]:
        nindex = np.zeros(n + 1, np.int32)
        nindex[i + 1:] = 1
        got = emst.synth_weights(emst.Graph(nindex, np.array([k], np.int32)))
        assert int(got.eweight[0]) == w
    # read_egr(synth_weights=True) on a weightless file
    p = str(tmp_path / "nw.egr")
    emst.write_egr(p, emst.Graph(g.nindex, g.nlist, None))
    assert np.array_equal(emst.read_egr(p, synth_weights=True).eweight, s.eweight)


def test_plain_threshold_matches_oracle():
    rng = np.random.default_rng(9)
    for _ in range(200):
        n = int(rng.integers(1, 60))
        g = emst.Graph(*random_graph(rng, n, int(rng.integers(0, 400)), -3, int(rng.integers(1, 1000))))
        assert emst.plain_threshold(g) == O.plain_threshold(g.nodes, g.eweight)
    for scale in (10, 14):
        g = emst.gen_rmat(scale, 16, seed=2)
        assert emst.plain_threshold(g) == O.plain_threshold(g.nodes, g.eweight)


def test_cpu_mst_matches_oracle():
    rng = np.random.default_rng(10)
    for _ in range(100):
        n = int(rng.integers(1, 200))
        g = emst.Graph(*random_graph(rng, n, int(rng.integers(0, 600)), -2, int(rng.integers(1, 20))))
        assert np.array_equal(emst.cpu_mst(g), O.kruskal(g))
    for g in (emst.gen_grid(64, 64, 1), emst.gen_rmat(14, 16, seed=2), emst.gen_rgg(20000, 6.0, 3)):
        k = O.kruskal(g)
        c = emst.cpu_mst(g)
        assert np.array_equal(c, k)
        differs, cw, gw = emst.verify(g, k, c)
        assert not differs and cw == gw == emst.mst_weight(g, c)


def test_verify_reports_difference():
    g = emst.gen_grid(8, 8, 2)
    k = emst.cpu_mst(g)
    bad = k.copy()
    bad[np.nonzero(k)[0][0]] = 0
    differs, cw, gw = emst.verify(g, k, bad)
    assert differs and cw > gw


def test_validate_rejects_bad_csr():
    ni, nl, ew = csr_from_edges(4, [(0, 1, 1), (2, 3, 1)])
    assert emst.validate(emst.Graph(ni, nl, ew)) == 0
    bad = nl.copy(); bad[0] = 4
    assert emst.validate(emst.Graph(ni, bad, ew)) == -2
    bad = nl.copy(); bad[0] = -1
    assert emst.validate(emst.Graph(ni, bad, ew)) == -2
    badi = ni.copy(); badi[2] = badi[3] + 1
    assert emst.validate(emst.Graph(badi, nl, ew)) == -2
    badi = ni.copy(); badi[-1] -= 1          # nindex[nodes] != edges
    assert emst.validate(emst.Graph(badi, nl, ew)) == -2
    badi = ni.copy(); badi[0] = 1            # nindex[0] != 0
    assert emst.validate(emst.Graph(badi, nl, ew)) == -2


def _digest(g):
    h = hashlib.sha256()
    for a in (g.nindex, g.nlist, g.eweight):
        h.update(a.tobytes())
    return h.hexdigest()


def _check_structure(g, wlo, whi):
    src = g.sources().astype(np.int64)
    dst = g.nlist.astype(np.int64)
    assert np.all(src != dst), "self-loop"
    # sorted, duplicate-free lists
    key = src * g.nodes + dst
    assert np.all(np.diff(key) > 0)
    # symmetric with mirrored weights
    rev = dst * g.nodes + src
    order = np.argsort(rev)
    assert np.array_equal(rev[order], key)
    assert np.array_equal(g.eweight[order], g.eweight)
    assert g.eweight.min() >= wlo and g.eweight.max() < whi


def test_generators_structure_and_golden_digests():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_gen_hashes.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    for name, f in mk.CASES.items():
        g = f()
        assert _digest(g) == GEN_HASHES[name], name
    _check_structure(emst.gen_grid(16, 24, 1, 1, 1 << 20), 1, 1 << 20)
    _check_structure(emst.gen_rmat(12, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24), 1, 1 << 24)
    _check_structure(emst.gen_rmat(10, 16, 0.57, 0.19, 0.19, 5, 1, 65), 1, 65)
    _check_structure(emst.gen_rgg(5000, 6.0, 3, 1 << 24), 1, 1 << 24)
    _check_structure(emst.gen_er(3000, 10.0, 4, 1, 1 << 20), 1, 1 << 20)


def test_generators_independent_of_thread_count():
    try:
        emst.set_gen_threads(1)
        a = [_digest(emst.gen_rmat(11, 16, seed=7)), _digest(emst.gen_rgg(6000, 6.0, 7))]
        emst.set_gen_threads(5)
        b = [_digest(emst.gen_rmat(11, 16, seed=7)), _digest(emst.gen_rgg(6000, 6.0, 7))]
    finally:
        emst.set_gen_threads(0)
    assert a == b


def test_grid_size_matches_reference_input_count():
    # 2d-2e20.sym has 4,190,208 directed edges (generate_compare_tables.py:16)
    g = emst.gen_grid(1024, 1024, 1)
    assert g.edges == GOLD["grid_2d_2e20_edges"] == 4 * 1024 * 1023
    assert emst.gen_grid(512, 512, 1).edges == 1046528


def test_rgg_mean_degree_and_rmat_size():
    g = emst.gen_rgg(1 << 16, 6.0, 3)
    assert 5.5 < g.edges / g.nodes < 6.1   # boundary effects lower it slightly
    g = emst.gen_rmat(12, 16, seed=2)
    assert g.nodes == 4096 and g.edges <= 2 * 16 * 4096


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_error_paths(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 255 and r.stdout.startswith("USAGE:")
    r = subprocess.run([CLI, str(tmp_path / "missing.egr")], capture_output=True, text=True)
    assert r.returncode == 255 and "could not open file" in r.stderr
    p = tmp_path / "bad.egr"
    p.write_bytes(np.array([0, 0], np.int32).tobytes())
    r = subprocess.run([CLI, str(p)], capture_output=True, text=True)
    assert r.returncode == 255 and "node or edge count too low" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_baseline_clis_and_entry_points_reject_bad_arguments(tmp_path):
    """The related-work baselines fail before touching a device: usage and
    unreadable input from their CLIs, EMST_E_ARG from null arguments."""
    import ctypes as C
    for name in ("contract_mst", "edgelist_mst"):
        cli = os.path.join(ROOT, "mst-artifact-sc23_amd", "bin", name)
        r = subprocess.run([cli], capture_output=True, text=True)
        assert r.returncode == 255 and r.stdout.startswith("USAGE:"), name
        r = subprocess.run([cli, str(tmp_path / "missing.egr")], capture_output=True, text=True)
        assert r.returncode == 255 and "could not read" in r.stderr, name
        fn = getattr(emst.lib, f"emst_{name}")
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p] * 4
        assert fn(None, None, None, None) == -1  # EMST_E_ARG


NULL_CALLS = {
    "emst_set_config": 2, "emst_get_config": 2, "emst_gpu_mst": (None, None, 0, 0, None, None),
    "emst_load": 2, "emst_solve": (None, 0, 0, None), "emst_fetch": 2, "emst_autotune": (None, 0, 0, 1, None, None),
    "emst_solve_device": (None, 1, 0, None, None, None, 0, 0, None, None, None), "emst_fetch_parent": 2,
    "emst_batch_create": (None, 1, None, 0, None), "emst_batch_info": 4, "emst_batch_solve": (None, 0, None, 0, None),
    "emst_batch_fetch": (None, 0, None), "emst_batch_thresholds": 3, "emst_batch_set_filter_model": 2,
    "emst_set_filter_model": 2, "emst_learned_threshold": 6, "emst_contract_mst": 4, "emst_edgelist_mst": 4,
}


def test_every_context_entry_point_rejects_null():
    """Each entry point that takes a context or batch returns EMST_E_ARG for
    a null one (no device is touched), destroy accepts null, and
    last_error(NULL) answers; run in a child process so a crash is a
    failure, not the end of the suite."""
    code = ["import ctypes as C, sys, json", f"sys.path.insert(0, {os.path.join(ROOT, 'mst-artifact-sc23_amd')!r})",
            "import emst", f"calls = {NULL_CALLS!r}", "out = {}",
            "for name, args in calls.items():",
            "    args = (None,) * args if isinstance(args, int) else args",
            "    f = getattr(emst.lib, name); f.restype = C.c_int",
            "    f.argtypes = [C.c_void_p if a is None else C.c_int for a in args]",
            "    out[name] = f(*args)",
            "for name in ('emst_destroy', 'emst_batch_destroy'):",
            "    f = getattr(emst.lib, name); f.restype = None; f.argtypes = [C.c_void_p]; f(None)",
            "emst.lib.emst_last_error.restype = C.c_char_p; emst.lib.emst_last_error.argtypes = [C.c_void_p]",
            "out['last_error'] = emst.lib.emst_last_error(None).decode()",
            "print(json.dumps(out))"]
    r = subprocess.run([sys.executable, "-c", "\n".join(code)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out.pop("last_error")
    assert out == {name: -1 for name in NULL_CALLS}  # EMST_E_ARG


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_schedule_fixture_grid512_recomputed():
    """The committed schedule fixture of the 512x512 grid, recomputed by the
    oracle from a freshly generated graph (the full-size fixtures are made
    the same way by tests/golden/make_schedules.py)."""
    from golden.make_schedules import CONFIGS, digest, schedule_entry
    with open(os.path.join(os.path.dirname(__file__), "golden", "schedules.json")) as f:
        ent = json.load(f)["grid512"]
    gen, thresholds = CONFIGS["grid512"]
    g = gen()
    assert digest(g) == ent["sha256"]
    for run, T in zip(ent["runs"], thresholds):
        uf, t = emst.plain_threshold(g) if T is None else (1, T)
        assert schedule_entry(g, uf, t) == run
    # closed forms: a connected grid has n - 1 forest edges and one component
    assert ent["runs"][0]["mst_edges"] == g.nodes - 1 and ent["runs"][0]["components"] == 1
    assert O.schedule(g, 0, emst.INT_MAX)[1].light == g.edges // 2


def _fmt_graph():
    # 5 vertices, a multi-edge, an isolated vertex; stored one direction
    # only, so the directed edge count is odd (the Galois pad word)
    ni, nl, ew = csr_from_edges(5, [(0, 1, 7), (1, 2, 3), (2, 0, 9), (2, 3, 3), (0, 1, 2)], symmetric=False)
    return emst.Graph(ni, nl, ew)


def test_format_writers(tmp_path):
    """The baseline formats, parsed back field by field (layouts of the
    reference's ecl2*_genweights.cpp converters)."""
    g = _fmt_graph()
    assert g.edges % 2 == 1
    src = np.repeat(np.arange(g.nodes), np.diff(g.nindex))
    # MatrixMarket: lower triangle, 1-based, one decimal
    p = tmp_path / "g.mtx"
    emst.write_format(str(p), g, "mtx")
    lines = p.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real symmetric"
    assert lines[1] == f"{g.nodes} {g.nodes} {g.edges // 2}"
    want = [f"{s + 1} {d + 1} {float(w):.1f}" for s, d, w in zip(src, g.nlist, g.eweight) if s > d]
    assert lines[2:] == want
    # PBBS weighted edge array: every directed edge, %f weights
    p = tmp_path / "g.pbbs"
    emst.write_format(str(p), g, "pbbs")
    lines = p.read_text().splitlines()
    assert lines[0] == "WeightedEdgeArray"
    assert lines[1:] == [f"{s} {d} {float(w):f}" for s, d, w in zip(src, g.nlist, g.eweight)]
    # UMinho edge list
    p = tmp_path / "g.um"
    emst.write_format(str(p), g, "uminho")
    lines = p.read_text().splitlines()
    assert lines[0] == f"{g.nodes} {g.edges}"
    assert lines[1:] == [f"{s} {d} {w}" for s, d, w in zip(src, g.nlist, g.eweight)]
    # Jucele edge list: counts on two lines, each undirected edge once
    p = tmp_path / "g.jmst"
    emst.write_format(str(p), g, "jucele")
    lines = p.read_text().splitlines()
    up = [(s, d, w) for s, d, w in zip(src, g.nlist, g.eweight) if d > s]
    assert lines[:2] == [f"{g.nodes}", f"{len(up)}"]
    assert lines[2:] == [f"{s} {d} {w}" for s, d, w in up]
    # Galois/Lonestar binary v1
    p = tmp_path / "g.gr"
    emst.write_format(str(p), g, "galois")
    b = p.read_bytes()
    head = np.frombuffer(b[:32], np.uint64)
    assert list(head) == [1, 4, g.nodes, g.edges]
    off = 32
    ends = np.frombuffer(b[off:off + 8 * g.nodes], np.uint64)
    assert np.array_equal(ends, g.nindex[1:])
    off += 8 * g.nodes
    assert np.array_equal(np.frombuffer(b[off:off + 4 * g.edges], np.int32), g.nlist)
    off += 4 * g.edges + (4 if g.edges % 2 else 0)
    assert np.array_equal(np.frombuffer(b[off:], np.int32), g.eweight)
    # no weights: refused
    with pytest.raises(emst.EmstError):
        emst.write_format(str(tmp_path / "x.mtx"), emst.Graph(g.nindex, g.nlist, None), "mtx")


REF_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
REF_CONVERTERS = {"mtx": "ecl2mtx", "pbbs": "ecl2pbbs", "galois": "ecl2lsg", "uminho": "ecl2uminho"}


@pytest.mark.skipif(not all(os.path.exists(os.path.join(REF_BIN, c)) for c in REF_CONVERTERS.values()),
                    reason="reference converters not built (needs /root/reference)")
def test_format_writers_match_reference_converters(tmp_path):
    """Each baseline-format writer against the reference's own converter
    (PBBS_CPU/inputs/ecl2pbbs_genweights.cpp, Lonestar_CPU/inputs/
    ecl2lsg_genweights.cpp, cuGraph_GPU/inputs/ecl2mtx_genweights_cugraph.cpp,
    UMinho_BOTH/inputs/ecl2UMinho_genweights.cpp), compiled from its source
    into oracle/_ref, on the same .egr bytes: the output files must be
    identical byte for byte.  Inputs: a one-way multigraph with an odd edge
    count (the .lsg pad word), an RMAT graph, weights above 2^24 and negative
    ones (float rounding and sign in the text formats), and a weightless file
    (both sides synthesize the weights of ECL_MST.cu:410-419)."""
    rng = np.random.default_rng(41)
    rm = emst.gen_rmat(10, 8, 0.57, 0.19, 0.19, 3, 1, 1 << 24)
    big = emst.Graph(rm.nindex, rm.nlist, rng.integers(1 << 24, (1 << 31) - 1, rm.edges).astype(np.int32))
    neg = emst.Graph(rm.nindex, rm.nlist, rng.integers(-(1 << 20), 1 << 20, rm.edges).astype(np.int32))
    grid = emst.gen_grid(30, 17, 2)
    cases = {"oneway": _fmt_graph(), "rmat": rm, "big": big, "neg": neg,
             "weightless": emst.Graph(grid.nindex, grid.nlist, None)}
    for name, g in cases.items():
        egr = tmp_path / f"{name}.egr"
        emst.write_egr(str(egr), g)
        ours_g = emst.read_egr(str(egr), synth_weights=True)
        for fmt, conv in REF_CONVERTERS.items():
            ref_out, our_out = tmp_path / f"{name}.{fmt}.ref", tmp_path / f"{name}.{fmt}.ours"
            r = subprocess.run([os.path.join(REF_BIN, conv), str(egr), str(ref_out)], capture_output=True, text=True)
            assert r.returncode == 0, (name, fmt, r.stderr)
            emst.write_format(str(our_out), ours_g, fmt)
            assert our_out.read_bytes() == ref_out.read_bytes(), (name, fmt)


def test_gen_and_convert_tools(tmp_path):
    """bin/ecl_gen writes the generators' graphs; bin/ecl_convert adds the
    reference's synthetic weights to a weightless input and converts."""
    bindir = os.path.join(os.path.dirname(os.path.dirname(__file__)), "mst-artifact-sc23_amd", "bin")
    out = tmp_path / "grid.egr"
    r = subprocess.run([os.path.join(bindir, "ecl_gen"), "grid", "16", "24", "1", "1", str(1 << 20), str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = emst.read_egr(str(out))
    ref = emst.gen_grid(16, 24, 1, 1, 1 << 20)
    assert np.array_equal(g.nindex, ref.nindex) and np.array_equal(g.nlist, ref.nlist)
    assert np.array_equal(g.eweight, ref.eweight)
    bare = tmp_path / "bare.egr"
    emst.write_egr(str(bare), emst.Graph(ref.nindex, ref.nlist, None))
    conv = tmp_path / "bare.um"
    r = subprocess.run([os.path.join(bindir, "ecl_convert"), str(bare), "uminho", str(conv)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    assert "added missing weights using ECL-MST method" in r.stdout
    w = np.array([int(x.split()[2]) for x in conv.read_text().splitlines()[1:]], np.int32)
    assert np.array_equal(w, emst.synth_weights(emst.Graph(ref.nindex, ref.nlist, None)).eweight)
    r = subprocess.run([os.path.join(bindir, "ecl_convert"), str(bare), "nosuch", str(conv)], capture_output=True,
                       text=True)
    assert r.returncode != 0


def test_tables_from_capture(tmp_path):
    """The table generator's arithmetic (median of 9, edges / median / 1e6)
    on a capture in the CLI's protocol."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "mst-artifact-sc23_amd"))
    import tables
    g = emst.gen_grid(8, 8, 1, 1, 100)
    path = tmp_path / "grid.egr"
    emst.write_egr(str(path), g)
    times = [0.5, 0.1, 0.9, 0.3, 0.2, 0.4, 0.6, 0.8, 0.7]
    cap = f"input: {path}\n" + "".join(f"{t:12.9f}\n" for t in times)
    rows = tables.parse_capture(cap)
    assert rows == [(str(path), times)]
    runtime, thru = tables.write_tables(str(tmp_path), [("ecl_mst", rows)], tables.egr_edges)
    assert runtime["ecl_mst"][str(path)] == 0.5
    assert abs(thru["ecl_mst"][str(path)] - g.edges / 0.5 / 1e6) < 1e-12
    text = (tmp_path / "comparison_throughputs.csv").read_text().splitlines()
    assert text[0] == ",ecl_mst, " and text[1].startswith("grid,")
    with pytest.raises(ValueError):
        tables.parse_capture(f"input: {path}\n 0.1\n")


def test_deopt_tables_from_captures(tmp_path):
    """The deopt tables (generate_deoptimize_tables.py:64-84): one column per
    rung capture, in the variant order, median of 9 and Medges/s."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "mst-artifact-sc23_amd"))
    import tables
    g = emst.gen_grid(8, 8, 1, 1, 100)
    path = tmp_path / "grid.egr"
    emst.write_egr(str(path), g)
    cols = []
    for k, v in enumerate(tables.DEOPT_VARIANTS):
        times = [0.1 * (k + 1)] * 9
        cols.append(("ecl_mst" + (f"_{v}" if v else ""), tables.parse_capture(
            f"input: {path}\n" + "".join(f"{t:12.9f}\n" for t in times))))
    runtime, thru = tables.write_tables(str(tmp_path), cols, tables.egr_edges, prefix="deopt")
    head = (tmp_path / "deopt_runtimes.csv").read_text().splitlines()[0]
    assert head == "," + "".join(f"{c}, " for c, _ in cols)
    assert head.startswith(",ecl_mst, ecl_mst_REF, ecl_mst_NAG, ecl_mst_NAG_NW, ")
    row = (tmp_path / "deopt_throughputs.csv").read_text().splitlines()[1].split(",")
    assert row[0] == "grid" and abs(float(row[2]) - g.edges / 0.2 / 1e6) < 1e-9


def test_cli_rejects_bad_variant_before_the_device(tmp_path):
    g = emst.gen_grid(4, 4, 1, 1, 100)
    path = tmp_path / "g.egr"
    emst.write_egr(str(path), g)
    for bad, extra in (("NAG_XYZ", {}), ("", {}), ("NAG_NW", {"EMST_FILTER": "learned"})):
        env = {**os.environ, "EMST_VARIANT": bad, **extra}
        r = subprocess.run([CLI, str(path)], capture_output=True, text=True, env=env)
        if bad == "":
            continue  # an empty name selects the default schedule (needs the device)
        assert r.returncode == 255 and "bad EMST_VARIANT" in r.stderr, (bad, r.stderr)


def test_threshold_cost_model_calibration_on_synthetic_times():
    """The cost-model search of train_filter.py --calibrate: on measurements
    generated from a known cost model (rounds at 30 us, bytes at 4 and 3
    TB/s), a setting equal to the generating one loses nothing, and a model
    that undercharges rounds picks a wrong candidate and loses time."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "mst-artifact-sc23_amd"))
    import train_filter as TF
    rng = np.random.default_rng(5)
    truth = TF.CostModel(30.0, 4e12, 3e12)
    calib = {}
    for g in range(12):
        rows = []
        for q in TF.QS:
            rounds = int(rng.integers(4, 30))
            bi, br = float(rng.uniform(1e8, 4e9)), float(rng.uniform(1e8, 6e9))
            rows.append({"q": q, "rounds": rounds, "init_bytes": bi, "round_bytes": br,
                         "ms": truth.cost(rounds, bi, br) / 1e3})
        calib[f"g{g}"] = {"rows": rows}
    reg = TF.model_regret(calib, 30.0, 4e12, 3e12)
    assert all(abs(v - 1.0) < 1e-12 for v in reg.values())
    # rounds nearly free: the model follows bytes alone and picks worse candidates
    reg_bad = TF.model_regret(calib, 0.01, 4e12, 3e12)
    assert max(reg_bad.values()) > 1.0
    # the label rule picks the cheapest candidate under the model
    rows = calib["g0"]["rows"]
    q = TF.best_q([(r["rounds"], r["init_bytes"], r["round_bytes"]) for r in rows], truth)
    assert q == min(rows, key=lambda r: r["ms"])["q"]
    # the unfiltered label sits above log2 0.5 by as much as the cut q >= 0.9 sits below 0
    y = TF.label_of(np.array([0.5, 1.0, 0.25]), "single")
    assert y[1] == np.float32(TF.OFF_LABEL) and y[0] == -1.0 and y[2] == -2.0
    assert abs((y[0] + y[1]) / 2 - np.log2(0.9)) < 0.01
    assert TF.label_of(np.array([1.0]), "union")[0] == 0.0


def test_binding_structs_match_header_layouts():
    """The ctypes mirrors of emst_stats, emst_batch_stats and emst_config
    list the header's fields in order, with C sizes that match (a field
    added on one side only would shift every field after it)."""
    text = open(os.path.join(ROOT, "include", "emst.h")).read()
    ctype = {"double": C.c_double, "int32_t": C.c_int32, "int64_t": C.c_int64, "uint64_t": C.c_uint64,
             "float": C.c_float}
    for struct, mirror in (("emst_stats", emst.Stats), ("emst_batch_stats", emst.BatchStats),
                           ("emst_config", emst.Config)):
        body = re.search(r"typedef struct \{([^}]*)\} " + struct + ";", text).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = [(t, n.strip()) for t, names in re.findall(r"^\s*(\w+)\s+([\w\s,]+);", body, flags=re.M)
                  for n in names.split(",")]
        assert [n for _, n in fields] == [n for n, _ in mirror._fields_], struct
        for (t, n), (_, ct) in zip(fields, mirror._fields_):
            assert C.sizeof(ctype[t]) == C.sizeof(ct), (struct, n)
