"""The allocation-free solve (emst_solve_device_noalloc) through the C-ABI,
against the allocating device-pointer entry point (emst_solve_device) on the
same device buffers: byte-identical inMST and equal counters, on two calls in
a row whose second input is smaller (the static scratch array keeps the first
call's contents), in the plain, unfiltered and learned modes; and the size
limit is refused before any device work."""
import ctypes as C
import os

import numpy as np
import pytest

import emst
import oracle_lib as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceGraph:
    """A graph and an inMST buffer in device memory, from the same HIP
    runtime libemst.so links (/opt/rocm), through ctypes."""

    def __init__(self, hip, g):
        self.hip, self.g, self.bufs = hip, g, []
        self.nindex = self._dev(g.nindex)
        self.nlist = self._dev(g.nlist)
        self.eweight = self._dev(g.eweight)
        self.out = self._dev(np.full(max(g.edges, 1), 7, np.uint8))

    def _dev(self, arr):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p.value

    def inmst(self):
        got = np.zeros(self.g.edges, np.uint8)
        assert self.hip.hipMemcpy(got.ctypes.data, C.c_void_p(self.out), self.g.edges, D2H) == 0
        return got

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def solve_both(s, d, mode, thr):
    g = d.g
    print(f"noalloc: n={g.nodes} m={g.edges} mode={mode}", flush=True)
    st_a = s.solve_device(g.nodes, g.edges, d.nindex, d.nlist, d.eweight, mode, thr, d.out)
    ref = d.inmst()
    st_b = s.solve_device_noalloc(g.nodes, g.edges, d.nindex, d.nlist, d.eweight, mode, thr, d.out)
    got = d.inmst()
    return ref, st_a, got, st_b


def test_noalloc_matches_allocating_entry_point(hip):
    big = emst.gen_rmat(14, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24)    # 16,384 vertices, ~426k edges
    small = emst.gen_grid(64, 80, 3, 1, 100)                         # smaller, with ties
    s = emst.Solver(0)
    s.set_filter_model(np.fromfile(os.path.join(os.path.dirname(emst.__file__), "filter_mlp.f32"), np.float32))
    devs = [DeviceGraph(hip, big), DeviceGraph(hip, small)]
    try:
        for d in devs:   # the second call of each mode runs on the smaller graph
            uf, T = emst.plain_threshold(d.g)
            kr = O.kruskal(d.g)
            for mode, thr in ((emst.FILTER_PLAIN, T), (emst.FILTER_OFF, emst.INT_MAX), (emst.FILTER_LEARNED, 0)):
                ref, st_a, got, st_b = solve_both(s, d, mode, thr)
                assert np.array_equal(got, ref)
                assert np.array_equal(got, kr)
                for f in ("light", "heavy", "light_rounds", "heavy_rounds", "threshold_used", "filter_used"):
                    assert getattr(st_a, f) == getattr(st_b, f), f
            # with every parallel record kept, the worklist sizes are the
            # reference schedule's and do not depend on record order
            s.set_config(flags=emst.F_NO_PAIR_DEDUP)
            _, st_a, got, st_b = solve_both(s, d, emst.FILTER_PLAIN if uf else emst.FILTER_OFF, T)
            s.set_config()
            assert np.array_equal(got, kr)
            assert (st_a.round_in_sum, st_a.round_out_sum) == (st_b.round_in_sum, st_b.round_out_sum)
        # twice in a row on the same graph: nothing carried over
        d = devs[1]
        got1 = s.solve_device_noalloc(d.g.nodes, d.g.edges, d.nindex, d.nlist, d.eweight, emst.FILTER_OFF,
                                      emst.INT_MAX, d.out)
        assert np.array_equal(d.inmst(), O.kruskal(d.g)) and got1.light > 0
        # the parent array of the last solve is the scratch array's
        roots = s.fetch_parent(d.g.nodes)
        assert all(roots[v] <= v for v in range(d.g.nodes))
        # a graph beyond the array is refused before any device work
        with pytest.raises(emst.EmstError) as ei:
            s.solve_device_noalloc(1 << 22, 1 << 24, d.nindex, d.nlist, d.eweight, emst.FILTER_OFF, emst.INT_MAX,
                                   d.out)
        assert ei.value.code == -5  # EMST_E_LIMIT
        assert "scratch" in str(ei.value)
        # the ladder allocates its own lists: refused
        s.set_config(flags=emst.F_LADDER)
        with pytest.raises(emst.EmstError) as ei:
            s.solve_device_noalloc(d.g.nodes, d.g.edges, d.nindex, d.nlist, d.eweight, emst.FILTER_OFF,
                                   emst.INT_MAX, d.out)
        assert ei.value.code == -1  # EMST_E_ARG
        s.set_config()
        # still usable after the refusals
        s.solve_device_noalloc(d.g.nodes, d.g.edges, d.nindex, d.nlist, d.eweight, emst.FILTER_OFF, emst.INT_MAX,
                               d.out)
        assert np.array_equal(d.inmst(), O.kruskal(d.g))
    finally:
        for d in devs:
            d.free()
        s.close()
