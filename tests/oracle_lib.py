"""ctypes access to the CPU oracle (oracle/emst_oracle.c).  Test
infrastructure: used only as the checker."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
MAXR = 256

if not os.path.exists(SO):
    raise ImportError(f"{SO} not built; run oracle/build.py")
lib = C.CDLL(SO)
_p32 = C.POINTER(C.c_int32)
_pu8 = C.POINTER(C.c_uint8)


class _G(C.Structure):
    _fields_ = [("nodes", C.c_int32), ("edges", C.c_int32), ("nindex", _p32), ("nlist", _p32),
                ("eweight", _p32)]


class _GB(C.Structure):
    _fields_ = _G._fields_


class Sched(C.Structure):
    _fields_ = [("light", C.c_int64), ("heavy_all", C.c_int64), ("heavy", C.c_int64),
                ("light_rounds", C.c_int), ("heavy_rounds", C.c_int),
                ("light_w", C.c_int64 * MAXR), ("light_s", C.c_int64 * MAXR),
                ("heavy_w", C.c_int64 * MAXR), ("heavy_s", C.c_int64 * MAXR),
                ("components", C.c_int64), ("mst_edges", C.c_int64), ("weight", C.c_uint64)]

    def rounds(self, phase):
        if phase == 0:
            r = self.light_rounds
            return [(self.light_w[i], self.light_s[i]) for i in range(r)]
        r = self.heavy_rounds
        return [(self.heavy_w[i], self.heavy_s[i]) for i in range(r)]


lib.orc_read_egr.argtypes = [C.c_char_p, C.POINTER(_GB)]
lib.orc_read_egr.restype = C.c_int
lib.orc_free_graph.argtypes = [C.POINTER(_GB)]
lib.orc_synth_weights.argtypes = [C.c_int32, _p32, _p32, _p32]
lib.orc_hash.argtypes = [C.c_uint32]
lib.orc_hash.restype = C.c_uint32
lib.orc_plain_threshold.argtypes = [C.c_int32, C.c_int32, _p32, C.POINTER(C.c_int), C.POINTER(C.c_int32)]
lib.orc_kruskal.argtypes = [C.POINTER(_G), _pu8]
lib.orc_ecl_schedule.argtypes = [C.POINTER(_G), C.c_int, C.c_int32, _pu8, C.POINTER(Sched)]
lib.orc_ecl_schedule_keyed.argtypes = [C.POINTER(_G), C.c_int, C.c_int32, _p32, _pu8, C.POINTER(Sched)]
lib.orc_verify.argtypes = [C.POINTER(_G), _pu8, _pu8, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]


def _view(nindex, nlist, eweight):
    g = _G()
    g.nodes = len(nindex) - 1
    g.edges = len(nlist)
    g.nindex = nindex.ctypes.data_as(_p32)
    g.nlist = nlist.ctypes.data_as(_p32)
    g.eweight = eweight.ctypes.data_as(_p32) if eweight is not None else _p32()
    return g


def _arrs(g):
    ni = np.ascontiguousarray(g.nindex, np.int32)
    nl = np.ascontiguousarray(g.nlist, np.int32)
    ew = None if g.eweight is None else np.ascontiguousarray(g.eweight, np.int32)
    return ni, nl, ew


def read_egr(path):
    """Returns (rc, nindex, nlist, eweight|None)."""
    gb = _GB()
    rc = lib.orc_read_egr(os.fsencode(path), C.byref(gb))
    if rc != 0:
        return rc, None, None, None
    n, m = gb.nodes, gb.edges
    ni = np.ctypeslib.as_array(gb.nindex, shape=(n + 1,)).copy()
    nl = np.ctypeslib.as_array(gb.nlist, shape=(max(m, 1),))[:m].copy()
    ew = np.ctypeslib.as_array(gb.eweight, shape=(max(m, 1),))[:m].copy() if gb.eweight else None
    lib.orc_free_graph(C.byref(gb))
    return 0, ni, nl, ew


def synth_weights(nindex, nlist):
    ni = np.ascontiguousarray(nindex, np.int32)
    nl = np.ascontiguousarray(nlist, np.int32)
    out = np.zeros(max(len(nl), 1), np.int32)
    lib.orc_synth_weights(len(ni) - 1, ni.ctypes.data_as(_p32), nl.ctypes.data_as(_p32), out.ctypes.data_as(_p32))
    return out[:len(nl)]


def ecl_hash(v):
    return int(lib.orc_hash(v))


def plain_threshold(nodes, eweight):
    ew = np.ascontiguousarray(eweight, np.int32)
    uf, t = C.c_int(), C.c_int32()
    lib.orc_plain_threshold(nodes, len(ew), ew.ctypes.data_as(_p32), C.byref(uf), C.byref(t))
    return bool(uf.value), int(t.value)


def kruskal(g):
    ni, nl, ew = _arrs(g)
    out = np.zeros(max(len(nl), 1), np.uint8)
    rc = lib.orc_kruskal(C.byref(_view(ni, nl, ew)), out.ctypes.data_as(_pu8))
    assert rc == 0
    return out[:len(nl)]


def schedule(g, use_filter, threshold):
    ni, nl, ew = _arrs(g)
    out = np.zeros(max(len(nl), 1), np.uint8)
    st = Sched()
    rc = lib.orc_ecl_schedule(C.byref(_view(ni, nl, ew)), int(use_filter), int(threshold),
                              out.ctypes.data_as(_pu8), C.byref(st))
    assert rc == 0
    return out[:len(nl)], st


def schedule_keyed(g, use_filter, threshold, kid):
    """orc_ecl_schedule with position j keyed by kid[j]; inMST is indexed
    by the key ids."""
    ni, nl, ew = _arrs(g)
    k = np.ascontiguousarray(kid, np.int32)
    out = np.zeros(max(len(nl), 1), np.uint8)
    st = Sched()
    rc = lib.orc_ecl_schedule_keyed(C.byref(_view(ni, nl, ew)), int(use_filter), int(threshold),
                                    k.ctypes.data_as(_p32), out.ctypes.data_as(_pu8), C.byref(st))
    assert rc == 0
    return out[:len(nl)], st


lib.orc_boruvka_omp.argtypes = [C.POINTER(_G), _pu8, C.c_int]
lib.orc_filter_kruskal_omp.argtypes = [C.POINTER(_G), _pu8, C.c_int]


def boruvka_omp(g, threads=0):
    ni, nl, ew = _arrs(g)
    out = np.zeros(max(len(nl), 1), np.uint8)
    assert lib.orc_boruvka_omp(C.byref(_view(ni, nl, ew)), out.ctypes.data_as(_pu8), threads) == 0
    return out[:len(nl)]


def filter_kruskal_omp(g, threads=0):
    ni, nl, ew = _arrs(g)
    out = np.zeros(max(len(nl), 1), np.uint8)
    assert lib.orc_filter_kruskal_omp(C.byref(_view(ni, nl, ew)), out.ctypes.data_as(_pu8), threads) == 0
    return out[:len(nl)]


lib.orc_forest_check.argtypes = [C.POINTER(_G), _pu8, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int)]


def forest_check(g, inmst):
    """(marked edges, graph components, is_forest) of a candidate forest."""
    ni, nl, ew = _arrs(g)
    x = np.ascontiguousarray(inmst, np.uint8)
    mk, cc, ok = C.c_int64(), C.c_int64(), C.c_int()
    assert lib.orc_forest_check(C.byref(_view(ni, nl, ew)), x.ctypes.data_as(_pu8), C.byref(mk), C.byref(cc),
                                C.byref(ok)) == 0
    return mk.value, cc.value, bool(ok.value)


def verify(g, cpu, gpu):
    ni, nl, ew = _arrs(g)
    cpu = np.ascontiguousarray(cpu, np.uint8)
    gpu = np.ascontiguousarray(gpu, np.uint8)
    cw, gw, oc, og = C.c_uint64(), C.c_uint64(), C.c_int64(), C.c_int64()
    rc = lib.orc_verify(C.byref(_view(ni, nl, ew)), cpu.ctypes.data_as(_pu8), gpu.ctypes.data_as(_pu8),
                        C.byref(cw), C.byref(gw), C.byref(oc), C.byref(og))
    return rc, cw.value, gw.value, oc.value, og.value
