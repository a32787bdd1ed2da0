"""Learned light-edge threshold: the network's CPU oracle (pinned by finite
differences), and on the GPU the device forward/backward, the device
features, exactness of the solve for any predicted threshold, the plain
fallback, and the gain of the committed trained network over the plain
rule."""
import math
import os
import sys

import numpy as np
import pytest

import emst
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

MODEL = os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp.f32")
MODEL_UNION = os.path.join(ROOT, "mst-artifact-sc23_amd", "filter_mlp_union.f32")


def _params(seed=0):
    return M.bf16(emst.init_params(seed)).astype(np.float64)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(1)
    X = rng.uniform(0, 1, (7, 32))
    T = rng.normal(-3, 1, 7)
    g = M.grad(p, X, T)
    idx = list(rng.integers(0, M.P, 60)) + [M.B3, M.W3, M.B2, M.B1]
    for k in idx:
        e = np.zeros(M.P)
        e[k] = 1e-6
        fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / 2e-6
        assert abs(fd - g[k]) <= 1e-5 + 1e-4 * abs(fd), k


def test_bf16_rounding_matches_library():
    vals = np.array([0.0, 1.0, 1.00390625, 1.01171875, -2.5e-3, 3.14159, 65504.0, 1e-30, -7.7], np.float32)
    lib = np.array([emst._bf16_bits(float(v)) for v in vals], np.uint32) << 16
    assert np.array_equal(lib.view(np.float32), M.bf16(vals))
    assert np.array_equal(emst.bf16_round(vals), M.bf16(vals))


@pytest.mark.parametrize("path", [MODEL, MODEL_UNION])
def test_committed_model_file_shape(path):
    if not os.path.exists(path):
        pytest.skip("no trained model committed")
    p = np.fromfile(path, np.float32)
    assert p.shape == (emst.MLP_PARAMS,) and np.all(np.isfinite(p))


gpu = pytest.mark.gpu


@gpu
def test_device_forward_and_grad_match_oracle():
    rng = np.random.default_rng(4)
    p32 = emst.init_params(3)
    pb = M.bf16(p32).astype(np.float64)
    X = rng.uniform(0, 1, (300, 32)).astype(np.float32)
    T = rng.normal(-3, 1.5, 300).astype(np.float32)
    y = emst.mlp_forward(p32, X)
    yr, _ = M.forward(pb, X)
    assert np.allclose(y, yr, rtol=1e-4, atol=1e-4)
    g, loss = emst.mlp_grad(p32, X, T)
    gr = M.grad(pb, X, T)
    assert abs(loss - M.loss(pb, X, T)) <= 1e-4 * max(1.0, M.loss(pb, X, T))
    assert np.allclose(g, gr, rtol=1e-3, atol=1e-5 * np.abs(gr).max())


@gpu
def test_device_training_reduces_loss():
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (512, 32)).astype(np.float32)
    T = (-1 - 4 * X[:, 2] + X[:, 5]).astype(np.float32)     # a learnable target
    p0 = emst.init_params(0)
    l0 = M.loss(M.bf16(p0), X, T)
    p, l1 = emst.mlp_train(p0, X, T, 400, 3e-3)
    assert l1 < 0.2 * l0
    assert abs(l1 - M.loss(M.bf16(p), X, T)) <= 1e-3 * max(1.0, l1)


@gpu
def test_device_features_match_oracle():
    s = emst.Solver(0)
    for g in (emst.gen_rmat(14, 16, seed=2), emst.gen_grid(100, 120, 1), emst.gen_rgg(30000, 6.0, 3)):
        s.load(g)
        uf, T, q, feat, samp = s.learned_threshold()
        f, w = M.features(g.nindex, g.eweight, g.nodes, g.edges)
        assert np.array_equal(samp, w)
        assert np.allclose(feat, f, rtol=1e-6, atol=1e-7)
        # no model loaded: the device falls back to the plain rule
        assert (uf, T) == O.plain_threshold(g.nodes, g.eweight) and q < 0
    s.close()


@gpu
def test_learned_mode_is_exact_for_any_model():
    s = emst.Solver(0)
    graphs = [emst.gen_rmat(13, 16, seed=2), emst.gen_grid(64, 90, 4), emst.gen_rgg(20000, 6.0, 5),
              emst.gen_er(8000, 10.0, 6), emst.gen_rmat(12, 8, seed=7, wlo=1, whi=9)]
    for seed in range(4):
        p = emst.init_params(seed) * (0.5 + seed)
        p[-1] = -1.5 * seed          # bias toward different quantiles
        s.set_filter_model(p)
        for g in graphs:
            got, st = s.gpu_mst(g, emst.FILTER_LEARNED, 0)
            assert np.array_equal(got, O.kruskal(g)), (seed, g.nodes)
            uf, T, q, _, _ = s.learned_threshold()
            assert st.threshold_used == T and st.filter_used == int(uf)
    s.close()


@gpu
@pytest.mark.parametrize("which", ["single", "union"])
def test_trained_model_beats_plain_rule(which):
    """The committed networks against the plain rule, each under the cost
    model it was trained for: a solve of its own (per-round cost counts) or
    a graph inside a batched arena (rounds shared, bytes count)."""
    path = MODEL if which == "single" else MODEL_UNION
    if not os.path.exists(path):
        pytest.skip("no trained model committed")
    sys.path.insert(0, os.path.join(ROOT, "mst-artifact-sc23_amd"))
    import train_filter as TF
    model = TF.MODELS[which]
    s = emst.Solver(0)
    s.set_config(flags=emst.F_NO_PAIR_DEDUP)  # the schedule the cost model was trained on
    s.set_filter_model(np.fromfile(path, np.float32))
    ratios = []
    for i in range(24):
        # Both networks also trained on Erdos-Renyi graphs: every fourth
        # held-out graph is one.
        er = i % 4 == 3
        kind, g = TF.make_graph(20_000_000 + i, 99, 3 if er else None)
        s.load(g)
        pc, pst = TF.plain_cost(s, g, model)
        got, lst = s.gpu_mst(g, emst.FILTER_LEARNED, 0)
        assert np.array_equal(got, O.kruskal(g))
        upw = np.sort(g.eweight[g.upper_mask()])
        hu = int(len(upw) - np.searchsorted(upw, lst.threshold_used, side="right")) if lst.filter_used else 0
        lc = TF.cost_of(g, lst, hu, model)
        print(f"{kind} n={g.nodes} plain {pc:.1f} learned {lc:.1f} q={lst.learned_q:.4f} "
              f"light {pst.light}->{lst.light} rounds {pst.light_rounds + pst.heavy_rounds}->"
              f"{lst.light_rounds + lst.heavy_rounds}", flush=True)
        ratios.append(lc / pc)
    s.close()
    gm = math.exp(np.mean(np.log(ratios)))
    print(which, "geomean learned/plain modelled cost", gm)
    assert gm < 0.95


@gpu
def test_learned_threshold_decided_once_per_loaded_graph():
    """emst_solve decides a loaded graph's learned threshold once, before its
    timed region, and reuses it until the next load or model change; the
    device-pointer entry decides it inside the solve.  Both give the same
    threshold, and a model change or a new graph is seen at once."""
    import ctypes as C
    s = emst.Solver(0)
    s.set_filter_model(np.fromfile(MODEL, np.float32))
    g1, g2 = emst.gen_rmat(14, 16, seed=3), emst.gen_rgg(30000, 6.0, 4)
    s.load(g1)
    a = s.solve(emst.FILTER_LEARNED, 0)
    got_a = s.fetch()
    b = s.solve(emst.FILTER_LEARNED, 0)
    assert (a.threshold_used, a.filter_used, a.learned_q) == (b.threshold_used, b.filter_used, b.learned_q)
    assert np.array_equal(s.fetch(), got_a) and np.array_equal(got_a, O.kruskal(g1))
    uf, T, q, _, _ = s.learned_threshold()
    assert (T, int(uf)) == (a.threshold_used, a.filter_used)
    # the device-pointer entry on the same graph decides inside the solve
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = []

    def dev(arr):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        bufs.append(p)
        assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p.value
    try:
        out = np.zeros(g1.edges, np.uint8)
        d_out = dev(out)
        c = s.solve_device(g1.nodes, g1.edges, dev(g1.nindex), dev(g1.nlist), dev(g1.eweight), emst.FILTER_LEARNED,
                           0, d_out)
        assert hip.hipMemcpy(out.ctypes.data, C.c_void_p(d_out), g1.edges, 2) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
    assert (c.threshold_used, c.filter_used) == (a.threshold_used, a.filter_used)
    assert np.array_equal(out, got_a)
    # a new graph, then no model (the plain rule): both seen by the next solve
    s.load(g2)
    d = s.solve(emst.FILTER_LEARNED, 0)
    uf2, T2, _, _, _ = s.learned_threshold()
    assert (d.threshold_used, d.filter_used) == (T2, int(uf2))
    s.set_filter_model(None)
    e = s.solve(emst.FILTER_LEARNED, 0)
    assert (bool(e.filter_used), e.threshold_used) == emst.plain_threshold(g2) and e.learned_q < 0
    assert np.array_equal(s.fetch(), O.kruskal(g2))
    s.close()
