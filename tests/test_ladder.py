"""The de-optimization ladder on the GPU: the reference's kernel sequence
(ECL_MST.cu:304-340) with each optimization switchable
(Deoptimizations/ECL_MST_No*.cu).

Every rung and every combination of the switches must give the same inMST
bytes as the fused schedule.  The first rung runs the reference's own
sequence, so its worklist sizes and round counts must also equal the round
schedule oracle's.  Both-direction worklists assume a symmetric CSR, as the
reference's inputs are; asymmetric inputs are checked without that switch."""
import numpy as np
import pytest

import emst
import oracle_lib as O
from graphs import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu

SINGLE = [emst.F_NO_GUARD, emst.F_NO_WARP, emst.F_NO_IMPLICIT_PC, emst.F_NO_ONE_DIRECTION, emst.F_NO_TUPLES,
          emst.F_NO_DATA_DRIVEN, emst.F_NO_EDGE_BASED]
COMBOS = [
    emst.F_NO_DATA_DRIVEN,                                             # 16-byte records revisited in place
    emst.F_NO_DATA_DRIVEN | emst.F_NO_IMPLICIT_PC,
    emst.F_NO_TUPLES | emst.F_NO_DATA_DRIVEN,                          # every CSR position every round
    emst.F_NO_TUPLES | emst.F_NO_ONE_DIRECTION,
    emst.F_NO_EDGE_BASED | emst.F_NO_ONE_DIRECTION | emst.F_NO_IMPLICIT_PC,
    emst.F_NO_WARP | emst.F_NO_TUPLES,
]


@pytest.fixture(scope="module")
def solver():
    s = emst.Solver(0)
    yield s
    s.close()


def solve(solver, g, flags, uf, T):
    solver.set_config(flags=flags)
    try:
        return solver.gpu_mst(g, uf, T)
    finally:
        solver.set_config()


def graphs():
    rng = np.random.default_rng(21)
    out = []
    for n, edges in [(1, []), (2, [(0, 1, 5)]), (3, [(0, 1, 1), (1, 2, 1), (0, 2, 1)]),
                     (6, [(0, 1, 7), (0, 1, 3), (1, 2, 3), (2, 0, 3), (4, 5, 0)])]:
        out.append(("tiny", emst.Graph(*csr_from_edges(n, edges))))
    for i in range(6):
        n = int(rng.integers(2, 300))
        out.append((f"random{i}", emst.Graph(*random_graph(rng, n, int(rng.integers(0, 4 * n)), 0, 8))))
    out.append(("grid", emst.gen_grid(64, 96, 1, 1, 1 << 20)))
    out.append(("rmat", emst.gen_rmat(12, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)))
    out.append(("rmat_ties", emst.gen_rmat(12, 16, 0.57, 0.19, 0.19, 5, 1, 9)))
    out.append(("rgg", emst.gen_rgg(1 << 13, 6.0, 3, 1 << 24)))
    out.append(("er", emst.gen_er(5000, 10.0, 4, 1, 1 << 20)))
    return out


GRAPHS = graphs()


@pytest.mark.parametrize("name,g", GRAPHS, ids=[n for n, _ in GRAPHS])
def test_first_rung_is_the_reference_schedule(solver, name, g):
    """F_LADDER alone runs initWL, kernel1-3 and the host-read sizes of
    ECL_MST.cu:304-340: same bytes, worklist sizes and round counts as the
    oracle's restatement of that schedule."""
    for uf, T in ((True, None), (False, emst.INT_MAX)):
        if T is None:
            uf, T = emst.plain_threshold(g)
        print(f"case {name} n={g.nodes} m={g.edges} filter={uf} T={T}", flush=True)
        got, st = solve(solver, g, emst.F_LADDER, uf, T)
        ref, sch = O.schedule(g, uf, T)
        assert np.array_equal(got, ref)
        assert (st.light, st.light_rounds) == (sch.light, sch.light_rounds)
        if uf:
            assert (st.heavy, st.heavy_rounds) == (sch.heavy, sch.heavy_rounds)
        sin = sum(sch.light_w[:sch.light_rounds]) + sum(sch.heavy_w[:sch.heavy_rounds])
        sout = sum(sch.light_s[:sch.light_rounds]) + sum(sch.heavy_s[:sch.heavy_rounds])
        assert (st.round_in_sum, st.round_out_sum) == (sin, sout)


@pytest.mark.parametrize("name,g", GRAPHS, ids=[n for n, _ in GRAPHS])
def test_every_rung_is_exact(solver, name, g):
    k = O.kruskal(g)
    uf0, T0 = emst.plain_threshold(g)
    for rung, flags, filt in emst.LADDER:
        uf, T = (uf0, T0) if filt else (False, emst.INT_MAX)
        print(f"case {name} rung {rung} flags={flags} filter={uf}", flush=True)
        got, st = solve(solver, g, flags, uf, T)
        assert np.array_equal(got, k), rung
        assert st.light_rounds + st.heavy_rounds >= (1 if g.edges and k.any() else 0)


@pytest.mark.parametrize("name,g", GRAPHS[-5:], ids=[n for n, _ in GRAPHS[-5:]])
def test_switches_alone_and_combined(solver, name, g):
    uf, T = emst.plain_threshold(g)
    base, _ = solver.gpu_mst(g, True, T)        # fused schedule, two phases
    off, _ = solver.gpu_mst(g, False, emst.INT_MAX)
    for flags in [emst.F_LADDER] + [emst.F_LADDER | f for f in SINGLE + COMBOS]:
        for filt, want in ((True, base), (False, off)):
            print(f"case {name} flags={flags} filter={filt}", flush=True)
            got, _ = solve(solver, g, flags, filt, T if filt else emst.INT_MAX)
            assert np.array_equal(got, want), (flags, filt)


def test_guard_flag_alone_stays_on_the_fused_schedule(solver):
    g = emst.gen_rmat(12, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)
    uf, T = emst.plain_threshold(g)
    got, st = solve(solver, g, emst.F_NO_GUARD, uf, T)
    _, sch = O.schedule(g, uf, T)
    assert np.array_equal(got, O.kruskal(g))
    assert st.light_rounds == sch.light_rounds


def test_negative_weights_and_asymmetric(solver):
    rng = np.random.default_rng(5)
    for i in range(12):
        n = int(rng.integers(2, 60))
        sym = i % 2 == 0
        g = emst.Graph(*random_graph(rng, n, 150, -4, 6, symmetric=sym))
        flags_set = [emst.F_LADDER] + [emst.F_LADDER | f for f in SINGLE + COMBOS]
        if not sym:
            flags_set = [f for f in flags_set if not f & emst.F_NO_ONE_DIRECTION]
        for uf, T in ((False, emst.INT_MAX), (True, 0), (True, -2)):
            ref, _ = O.schedule(g, uf, T)
            for flags in flags_set:
                print(f"case neg/asym {i} flags={flags} filter={uf} T={T}", flush=True)
                got, _ = solve(solver, g, flags, uf, T)
                assert np.array_equal(got, ref), (i, flags, uf, T)


def test_ladder_argument_errors(solver):
    g = emst.gen_grid(8, 8, 1, 1, 100)
    solver.set_config(flags=emst.F_LADDER)
    try:
        with pytest.raises(emst.EmstError) as e:
            solver.gpu_mst(g, emst.FILTER_LEARNED, 0)
        assert e.value.code == -1  # EMST_E_ARG
        b = emst.Batch(solver, [g, g])
        with pytest.raises(emst.EmstError):
            b.solve(emst.FILTER_PLAIN)
        b.close()
    finally:
        solver.set_config()
    with pytest.raises(emst.EmstError) as e:
        solver.set_config(flags=1 << 20)
    assert e.value.code == -1  # EMST_E_ARG
