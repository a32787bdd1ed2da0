"""Python binding of libemst.so (ctypes), mirroring ECL-MST's host interface.

The reference (ECL-MST/ECL_MST.cu) is a C++ program; its host interface is
readECLgraph/writeECLgraph (ECLgraph.h:37-80), the threshold selection in
main (ECL_MST.cu:432-443), gpuMST<filter>(g, threshold) (ECL_MST.cu:265-362),
cpuMST (ECL_MST.cu:82-124) and verify (ECL_MST.cu:365-386).  This module
exposes the same operations over the C-ABI in include/emst.h so tests and
the benchmark can drive the HIP path.  There is no fallback: if the shared
library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libemst.so")

EMST_OK = 0
ERRORS = {
    -1: "EMST_E_ARG", -2: "EMST_E_GRAPH", -3: "EMST_E_HIP", -4: "EMST_E_NOMEM", -5: "EMST_E_LIMIT",
    -6: "EMST_E_STATE", -7: "EMST_E_NOWEIGHT", -10: "EMST_E_OPEN", -11: "EMST_E_NODES",
    -12: "EMST_E_EDGES", -13: "EMST_E_COUNTS", -14: "EMST_E_NINDEX", -15: "EMST_E_NLIST",
    -16: "EMST_E_WEIGHTS", -17: "EMST_E_WRITE",
}
F_HOST_SYNC = 1
F_NO_GUARD = 2
F_NO_COMBINE = 4
F_NO_PAIR_DEDUP = 1024
F_NO_SPARSE_HEAVY = 2048
F_NO_LIGHT_OFFERS = 4096
F_NO_RELABEL = 8192
F_RELABEL = 16384
F_CLUSTER = 32768
# the de-optimization ladder (Deoptimizations/ECL_MST_No*.cu)
F_LADDER = 8
F_NO_WARP = 16
F_NO_IMPLICIT_PC = 32
F_NO_ONE_DIRECTION = 64
F_NO_TUPLES = 128
F_NO_DATA_DRIVEN = 256
F_NO_EDGE_BASED = 512
# Rungs in the order the reference removes them (generate_deoptimize_tables.py
# names the variants by their suffix chain); the filter rung is the filter mode.
LADDER = [
    ("ECL-MST", F_LADDER, True),
    ("NoAtomicGuards", F_LADDER | F_NO_GUARD, True),
    ("NoWarp", F_LADDER | F_NO_GUARD | F_NO_WARP, True),
    ("NoFilter", F_LADDER | F_NO_GUARD | F_NO_WARP, False),
    ("NoImplicitPathCompression", F_LADDER | F_NO_GUARD | F_NO_WARP | F_NO_IMPLICIT_PC, False),
    ("NoOneDirection", F_LADDER | F_NO_GUARD | F_NO_WARP | F_NO_IMPLICIT_PC | F_NO_ONE_DIRECTION, False),
    ("NoTuples", F_LADDER | F_NO_GUARD | F_NO_WARP | F_NO_IMPLICIT_PC | F_NO_ONE_DIRECTION | F_NO_TUPLES, False),
    ("NoDataDriven", F_LADDER | F_NO_GUARD | F_NO_WARP | F_NO_IMPLICIT_PC | F_NO_ONE_DIRECTION | F_NO_TUPLES
     | F_NO_DATA_DRIVEN, False),
    ("NoEdgeBased", F_LADDER | F_NO_GUARD | F_NO_WARP | F_NO_IMPLICIT_PC | F_NO_ONE_DIRECTION | F_NO_TUPLES
     | F_NO_DATA_DRIVEN | F_NO_EDGE_BASED, False),
]
FILTER_OFF, FILTER_PLAIN, FILTER_LEARNED = 0, 1, 2
MLP_PARAMS, MLP_FEATURES, WEIGHT_SAMPLES = 6337, 32, 1024
INT_MAX = 2**31 - 1


class EmstError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run mst-artifact-sc23_amd/build.py")
lib = C.CDLL(LIB_PATH)

_libc_free = C.CDLL(None).free
_libc_free.argtypes = [C.c_void_p]
_libc_free.restype = None
_p32 = C.POINTER(C.c_int32)
_pu8 = C.POINTER(C.c_uint8)


class _Graph(C.Structure):
    _fields_ = [("nodes", C.c_int32), ("edges", C.c_int32), ("nindex", _p32), ("nlist", _p32),
                ("eweight", _p32)]


class _GraphBuf(C.Structure):
    _fields_ = [("nodes", C.c_int32), ("edges", C.c_int32), ("nindex", _p32), ("nlist", _p32),
                ("eweight", _p32)]


class Stats(C.Structure):
    _fields_ = [("ms", C.c_double), ("init_ms", C.c_double), ("init_light_ms", C.c_double),
                ("init_heavy_ms", C.c_double), ("init_launches", C.c_int32),
                ("light_rounds", C.c_int32), ("heavy_rounds", C.c_int32), ("tail_rounds", C.c_int32),
                ("light", C.c_int64), ("heavy", C.c_int64), ("mst_edges", C.c_int64),
                ("weight", C.c_uint64), ("round_in_sum", C.c_int64), ("round_out_sum", C.c_int64),
                ("threshold_used", C.c_int32), ("filter_used", C.c_int32), ("learned_q", C.c_float),
                ("sparse_heavy", C.c_int32), ("light_offers", C.c_int32), ("relabelled", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Config(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid_per_cu", C.c_int32), ("tile_edges", C.c_int32),
                ("stage_recs", C.c_int32), ("tail_max", C.c_int32), ("flags", C.c_int32),
                ("init_grid_per_cu", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_create = _sig("emst_create", C.c_int, C.c_int, C.POINTER(C.c_void_p))
_destroy = _sig("emst_destroy", None, C.c_void_p)
_last_error = _sig("emst_last_error", C.c_char_p, C.c_void_p)
_set_config = _sig("emst_set_config", C.c_int, C.c_void_p, C.POINTER(Config))
_get_config = _sig("emst_get_config", C.c_int, C.c_void_p, C.POINTER(Config))
_validate = _sig("emst_validate", C.c_int, C.POINTER(_Graph))
_plain_threshold = _sig("emst_plain_threshold", C.c_int, C.POINTER(_Graph), C.POINTER(C.c_int),
                        C.POINTER(C.c_int32))
_gpu_mst = _sig("emst_gpu_mst", C.c_int, C.c_void_p, C.POINTER(_Graph), C.c_int, C.c_int32, _pu8,
                C.POINTER(Stats))
_load = _sig("emst_load", C.c_int, C.c_void_p, C.POINTER(_Graph))
_solve = _sig("emst_solve", C.c_int, C.c_void_p, C.c_int, C.c_int32, C.POINTER(Stats))
_fetch = _sig("emst_fetch", C.c_int, C.c_void_p, _pu8)
_fetch_parent = _sig("emst_fetch_parent", C.c_int, C.c_void_p, _p32)
_solve_device = _sig("emst_solve_device", C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                     C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(Stats))
_solve_device_noalloc = _sig("emst_solve_device_noalloc", C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(Stats))
SCRATCH_BYTES = 64 << 20  # EMST_SCRATCH_BYTES
_read_egr = _sig("emst_read_egr", C.c_int, C.c_char_p, C.POINTER(_GraphBuf))
_write_egr = _sig("emst_write_egr", C.c_int, C.c_char_p, C.POINTER(_Graph))
_writers = {fmt: _sig(f"emst_write_{fmt}", C.c_int, C.c_char_p, C.POINTER(_Graph))
            for fmt in ("mtx", "pbbs", "galois", "uminho", "jucele")}
_free_graph = _sig("emst_free_graph", None, C.POINTER(_GraphBuf))
_synth_weights = _sig("emst_synth_weights", C.c_int, C.POINTER(_GraphBuf))
_mst_weight = _sig("emst_mst_weight", C.c_uint64, C.POINTER(_Graph), _pu8)
_cpu_mst = _sig("emst_cpu_mst", C.c_int, C.POINTER(_Graph), _pu8)
_verify = _sig("emst_verify", C.c_int, C.POINTER(_Graph), _pu8, _pu8, C.POINTER(C.c_uint64),
               C.POINTER(C.c_uint64))
_version = _sig("emst_version", C.c_char_p)
_gen_grid = _sig("emst_gen_grid", C.c_int, C.c_int32, C.c_int32, C.c_uint64, C.c_int32, C.c_int32,
                 C.POINTER(_GraphBuf))
_gen_rmat = _sig("emst_gen_rmat", C.c_int, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                 C.c_uint64, C.c_int32, C.c_int32, C.POINTER(_GraphBuf))
_gen_rgg = _sig("emst_gen_rgg", C.c_int, C.c_int32, C.c_double, C.c_uint64, C.c_int32, C.POINTER(_GraphBuf))
_gen_er = _sig("emst_gen_er", C.c_int, C.c_int32, C.c_double, C.c_uint64, C.c_int32, C.c_int32,
               C.POINTER(_GraphBuf))
_gen_threads = _sig("emst_gen_set_threads", None, C.c_int32)
_pf32 = C.POINTER(C.c_float)
_set_model = _sig("emst_set_filter_model", C.c_int, C.c_void_p, _pf32)
_learned_thr = _sig("emst_learned_threshold", C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int32),
                    C.POINTER(C.c_float), _pf32, _p32)
_mlp_fwd = _sig("emst_mlp_forward_batch", C.c_int, C.c_int, C.c_int32, _pf32, _pf32, _pf32)
_mlp_grad = _sig("emst_mlp_grad_batch", C.c_int, C.c_int, C.c_int32, _pf32, _pf32, _pf32, _pf32, C.POINTER(C.c_float))
_mlp_train = _sig("emst_mlp_train", C.c_int, C.c_int, C.c_int32, _pf32, _pf32, C.c_int32, C.c_float, C.c_uint64,
                  _pf32, C.POINTER(C.c_float))
_bf16_bits = _sig("emst_bf16_bits", C.c_uint16, C.c_float)
class BatchStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("chunks", C.c_int32), ("graphs", C.c_int32), ("nodes", C.c_int64),
                ("edges", C.c_int64), ("light", C.c_int64), ("heavy", C.c_int64), ("light_rounds", C.c_int32),
                ("heavy_rounds", C.c_int32), ("replays", C.c_int32), ("filtered_graphs", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_batch_create = _sig("emst_batch_create", C.c_int, C.c_void_p, C.c_int32, C.POINTER(_Graph), C.c_int64,
                     C.POINTER(C.c_void_p))
_batch_destroy = _sig("emst_batch_destroy", None, C.c_void_p)
_batch_solve = _sig("emst_batch_solve", C.c_int, C.c_void_p, C.c_int, _p32, C.c_int, C.POINTER(BatchStats))
_batch_fetch = _sig("emst_batch_fetch", C.c_int, C.c_void_p, C.c_int32, _pu8)
_batch_thresholds = _sig("emst_batch_thresholds", C.c_int, C.c_void_p, _p32, _p32)
_batch_set_model = _sig("emst_batch_set_filter_model", C.c_int, C.c_void_p, _pf32)
_batch_arena_ms = _sig("emst_batch_arena_ms", C.c_int, C.c_void_p, C.POINTER(C.c_double), _p32, C.c_int32,
                       C.POINTER(C.c_int32))
_batch_info = _sig("emst_batch_info", C.c_int, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                   C.POINTER(C.c_int64))
class ContractStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("iterations", C.c_int32), ("color_passes", C.c_int32),
                ("mst_edges", C.c_int64), ("weight", C.c_uint64)]


_contract = _sig("emst_contract_mst", C.c_int, C.c_void_p, C.POINTER(_Graph), _pu8, C.POINTER(ContractStats))


class EdgelistStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("iterations", C.c_int64), ("label_passes", C.c_int64),
                ("mst_edges", C.c_int64), ("weight", C.c_uint64)]


_edgelist = _sig("emst_edgelist_mst", C.c_int, C.c_void_p, C.POINTER(_Graph), _pu8, C.POINTER(EdgelistStats))
_autotune = _sig("emst_autotune", C.c_int, C.c_void_p, C.c_int, C.c_int32, C.c_int32, C.POINTER(Config),
                 C.POINTER(C.c_double))


def _check(rc, msg=""):
    if rc != EMST_OK:
        raise EmstError(rc, msg)
    return rc


class Graph:
    """A graph in the ECLgraph CSR layout held in numpy int32 arrays
    (ECLgraph.h:29-35).  eweight may be None for a weightless file."""

    def __init__(self, nindex, nlist, eweight=None):
        self.nindex = np.ascontiguousarray(nindex, dtype=np.int32)
        self.nlist = np.ascontiguousarray(nlist, dtype=np.int32)
        self.eweight = None if eweight is None else np.ascontiguousarray(eweight, dtype=np.int32)
        self.nodes = int(self.nindex.shape[0] - 1)
        self.edges = int(self.nlist.shape[0])

    def view(self):
        g = _Graph()
        g.nodes, g.edges = self.nodes, self.edges
        g.nindex = self.nindex.ctypes.data_as(_p32)
        g.nlist = self.nlist.ctypes.data_as(_p32)
        g.eweight = self.eweight.ctypes.data_as(_p32) if self.eweight is not None else _p32()
        return g

    def upper_mask(self):
        src = np.repeat(np.arange(self.nodes, dtype=np.int64), np.diff(self.nindex))
        return self.nlist > src

    def sources(self):
        return np.repeat(np.arange(self.nodes, dtype=np.int32), np.diff(self.nindex))


def _take(buf):
    n, m = buf.nodes, buf.edges
    nindex = np.ctypeslib.as_array(buf.nindex, shape=(n + 1,)).copy()
    nlist = np.ctypeslib.as_array(buf.nlist, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
    ew = None
    if buf.eweight:
        ew = np.ctypeslib.as_array(buf.eweight, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
    _free_graph(C.byref(buf))
    return Graph(nindex, nlist, ew)


def read_egr(path, synth_weights=False):
    """readECLgraph (ECLgraph.h:37-64); optionally the weight synthesis of
    ECL_MST.cu:410-419 when the file has none."""
    buf = _GraphBuf()
    rc = _read_egr(os.fsencode(path), C.byref(buf))
    _check(rc, path)
    if synth_weights and not buf.eweight:
        rc = _synth_weights(C.byref(buf))
        if rc:
            _free_graph(C.byref(buf))
            _check(rc)
    return _take(buf)


def write_egr(path, g):
    _check(_write_egr(os.fsencode(path), C.byref(g.view())), path)


def write_format(path, g, fmt):
    """One of the reference's baseline formats: "mtx", "pbbs", "galois", "uminho" or "jucele"."""
    _check(_writers[fmt](os.fsencode(path), C.byref(g.view())), path)


def synth_weights(g):
    """Weights of ECL_MST.cu:410-419 for g's structure (returns a new Graph)."""
    buf = _GraphBuf()
    buf.nodes, buf.edges = g.nodes, g.edges
    nindex = np.ascontiguousarray(g.nindex)
    nlist = np.ascontiguousarray(g.nlist)
    buf.nindex = nindex.ctypes.data_as(_p32)
    buf.nlist = nlist.ctypes.data_as(_p32)
    buf.eweight = _p32()
    _check(_synth_weights(C.byref(buf)))
    w = np.ctypeslib.as_array(buf.eweight, shape=(max(g.edges, 1),))[:g.edges].copy()
    _libc_free(C.cast(buf.eweight, C.c_void_p))
    return Graph(g.nindex, g.nlist, w)


def validate(g):
    return _validate(C.byref(g.view()))


def plain_threshold(g):
    """(use_filter, threshold) as chosen at ECL_MST.cu:432-443."""
    uf, t = C.c_int(), C.c_int32()
    _check(_plain_threshold(C.byref(g.view()), C.byref(uf), C.byref(t)))
    return bool(uf.value), int(t.value)


def cpu_mst(g):
    out = np.zeros(max(g.edges, 1), np.uint8)
    _check(_cpu_mst(C.byref(g.view()), out.ctypes.data_as(_pu8)))
    return out[:g.edges]


def verify(g, cpu, gpu):
    """verify (ECL_MST.cu:365-386): (differs, cpu_weight, gpu_weight)."""
    cw, gw = C.c_uint64(), C.c_uint64()
    cpu = np.ascontiguousarray(cpu, np.uint8)
    gpu = np.ascontiguousarray(gpu, np.uint8)
    rc = _verify(C.byref(g.view()), cpu.ctypes.data_as(_pu8), gpu.ctypes.data_as(_pu8), C.byref(cw), C.byref(gw))
    if rc < 0:
        _check(rc)
    return bool(rc), cw.value, gw.value


def mst_weight(g, inmst):
    inmst = np.ascontiguousarray(inmst, np.uint8)
    return int(_mst_weight(C.byref(g.view()), inmst.ctypes.data_as(_pu8)))


def set_gen_threads(t):
    _gen_threads(int(t))


def gen_grid(rows, cols, seed, wlo=1, whi=1 << 20):
    buf = _GraphBuf()
    _check(_gen_grid(rows, cols, seed, wlo, whi, C.byref(buf)))
    return _take(buf)


def gen_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=2, wlo=1, whi=1 << 24):
    buf = _GraphBuf()
    _check(_gen_rmat(scale, edge_factor, a, b, c, seed, wlo, whi, C.byref(buf)))
    return _take(buf)


def gen_rgg(n, mean_degree, seed, wmax=1 << 24):
    buf = _GraphBuf()
    _check(_gen_rgg(n, mean_degree, seed, wmax, C.byref(buf)))
    return _take(buf)


def gen_er(n, mean_degree, seed, wlo=1, whi=1 << 20):
    buf = _GraphBuf()
    _check(_gen_er(n, mean_degree, seed, wlo, whi, C.byref(buf)))
    return _take(buf)


def version():
    return _version().decode()


class Solver:
    """One device context: gpuMST (ECL_MST.cu:265-362) with the device
    buffers cached across calls."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _create(device, C.byref(h))
        if rc != EMST_OK:
            raise EmstError(rc, f"emst_create(device={device}) failed (no HIP device?)")
        self._h = h
        self.graph = None

    def close(self):
        if self._h:
            _destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc):
        if rc != EMST_OK:
            raise EmstError(rc, _last_error(self._h).decode())

    def config(self):
        c = Config()
        self._err(_get_config(self._h, C.byref(c)))
        return c

    def set_config(self, **kw):
        c = Config()
        for k, v in kw.items():
            setattr(c, k, v)
        self._err(_set_config(self._h, C.byref(c)))

    def load(self, g):
        self._err(_load(self._h, C.byref(g.view())))
        self.graph = g

    def solve(self, use_filter, threshold):
        st = Stats()
        self._err(_solve(self._h, int(use_filter), int(threshold), C.byref(st)))
        return st

    def fetch(self):
        m = self.graph.edges
        out = np.zeros(max(m, 1), np.uint8)
        self._err(_fetch(self._h, out.ctypes.data_as(_pu8)))
        return out[:m]

    def fetch_parent(self, nodes):
        out = np.zeros(nodes, np.int32)
        self._err(_fetch_parent(self._h, out.ctypes.data_as(_p32)))
        return out

    def gpu_mst(self, g, use_filter=None, threshold=None):
        """Whole gpuMST call; threshold defaults to the plain heuristic."""
        if use_filter is None:
            use_filter, threshold = plain_threshold(g)
        out = np.zeros(max(g.edges, 1), np.uint8)
        st = Stats()
        self._err(_gpu_mst(self._h, C.byref(g.view()), int(use_filter), int(threshold),
                           out.ctypes.data_as(_pu8), C.byref(st)))
        self.graph = g
        return out[:g.edges], st

    def solve_device(self, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inmst,
                     stream=None):
        st = Stats()
        self._err(_solve_device(self._h, nodes, edges, C.c_void_p(d_nindex), C.c_void_p(d_nlist),
                                C.c_void_p(d_eweight), int(use_filter), int(threshold),
                                C.c_void_p(d_inmst), C.c_void_p(stream) if stream else None, C.byref(st)))
        return st

    def solve_device_noalloc(self, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inmst,
                             stream=None):
        """emst_solve_device_noalloc: the same solve with every work buffer
        in the library's static scratch array (no device allocation)."""
        st = Stats()
        self._err(_solve_device_noalloc(self._h, nodes, edges, C.c_void_p(d_nindex), C.c_void_p(d_nlist),
                                        C.c_void_p(d_eweight), int(use_filter), int(threshold),
                                        C.c_void_p(d_inmst), C.c_void_p(stream) if stream else None, C.byref(st)))
        return st


    def contract_mst(self, g):
        """The related-work baseline (Boruvka with graph contraction, the
        UMinho GPU algorithm): (inMST, ContractStats)."""
        out = np.zeros(max(g.edges, 1), np.uint8)
        st = ContractStats()
        self._err(_contract(self._h, C.byref(g.view()), out.ctypes.data_as(_pu8), C.byref(st)))
        self.graph = g
        return out[:g.edges], st

    def edgelist_mst(self, g):
        """The second related-work baseline (edge-list Boruvka with label
        propagation, the Jucele GPU algorithm): (inMST, EdgelistStats).
        inMST equals gpu_mst's bytes for the same graph."""
        out = np.zeros(max(g.edges, 1), np.uint8)
        st = EdgelistStats()
        self._err(_edgelist(self._h, C.byref(g.view()), out.ctypes.data_as(_pu8), C.byref(st)))
        self.graph = g
        return out[:g.edges], st

    def autotune(self, use_filter, threshold, reps=3):
        """Runs the autotuner on the loaded graph; returns (config dict, ms)."""
        c, ms = Config(), C.c_double()
        self._err(_autotune(self._h, int(use_filter), int(threshold), int(reps), C.byref(c), C.byref(ms)))
        return c.as_dict(), ms.value

    def set_filter_model(self, params):
        """Loads the threshold network (MLP_PARAMS fp32 values); None removes it."""
        if params is None:
            self._err(_set_model(self._h, None))
            return
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._err(_set_model(self._h, p.ctypes.data_as(_pf32)))

    def learned_threshold(self):
        """(use_filter, threshold, q, features[32], samples[1024]) for the loaded graph."""
        uf, t, q = C.c_int(), C.c_int32(), C.c_float()
        feat = np.zeros(MLP_FEATURES, np.float32)
        samp = np.zeros(WEIGHT_SAMPLES, np.int32)
        self._err(_learned_thr(self._h, C.byref(uf), C.byref(t), C.byref(q), feat.ctypes.data_as(_pf32),
                               samp.ctypes.data_as(_p32)))
        return bool(uf.value), int(t.value), float(q.value), feat, samp


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def mlp_forward(params, X, device=0):
    X = _f32(X)
    p = _f32(params)
    Y = np.zeros(X.shape[0], np.float32)
    _check(_mlp_fwd(device, X.shape[0], p.ctypes.data_as(_pf32), X.ctypes.data_as(_pf32), Y.ctypes.data_as(_pf32)))
    return Y


def mlp_grad(params, X, T, device=0):
    X, T, p = _f32(X), _f32(T), _f32(params)
    G = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_float()
    _check(_mlp_grad(device, X.shape[0], p.ctypes.data_as(_pf32), X.ctypes.data_as(_pf32), T.ctypes.data_as(_pf32),
                     G.ctypes.data_as(_pf32), C.byref(loss)))
    return G, float(loss.value)


def mlp_train(params, X, T, epochs, lr, seed=0, device=0):
    X, T = _f32(X), _f32(T)
    p = _f32(params).copy()
    loss = C.c_float()
    _check(_mlp_train(device, X.shape[0], X.ctypes.data_as(_pf32), T.ctypes.data_as(_pf32), int(epochs), float(lr),
                      int(seed), p.ctypes.data_as(_pf32), C.byref(loss)))
    return p, float(loss.value)


def bf16_round(a):
    """Round-to-nearest-even bf16 of fp32 values (same rule as the device)."""
    a = _f32(a)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    special = (u & 0x7F800000) == 0x7F800000
    r = np.where(special, u.astype(np.uint32), r)
    return r.view(np.float32)


def init_params(seed=0):
    """He-initialised fp32 parameters of the 32-64-64-1 network."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, np.float32)
    p[0:2048] = rng.normal(0, np.sqrt(2 / 32), 2048)
    p[2112:6208] = rng.normal(0, np.sqrt(2 / 64), 4096)
    p[6272:6336] = rng.normal(0, np.sqrt(1 / 64), 64)
    return p


class Batch:
    """Many graphs packed into device arenas and solved together."""

    def __init__(self, solver, graphs, max_arena_edges=0):
        self.solver = solver
        self.graphs = list(graphs)
        self._views = (_Graph * len(self.graphs))(*[g.view() for g in self.graphs])
        h = C.c_void_p()
        solver._err(_batch_create(solver._h, len(self.graphs), self._views, int(max_arena_edges), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        c, n, m = C.c_int32(), C.c_int64(), C.c_int64()
        self.solver._err(_batch_info(self._h, C.byref(c), C.byref(n), C.byref(m)))
        return c.value, n.value, m.value

    def solve(self, use_filter=FILTER_PLAIN, thresholds=None, use_graph=True):
        st = BatchStats()
        tp = None
        if thresholds is not None:
            self._thr = np.ascontiguousarray(thresholds, np.int32)
            tp = self._thr.ctypes.data_as(_p32)
        self.solver._err(_batch_solve(self._h, int(use_filter), tp, int(bool(use_graph)), C.byref(st)))
        return st

    def fetch(self, g):
        m = self.graphs[g].edges
        out = np.zeros(max(m, 1), np.uint8)
        self.solver._err(_batch_fetch(self._h, int(g), out.ctypes.data_as(_pu8)))
        return out[:m]

    def set_filter_model(self, params):
        """Network for FILTER_LEARNED solves of this batch; None uses the solver's."""
        if params is None:
            self.solver._err(_batch_set_model(self._h, None))
            return
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._model = p
        self.solver._err(_batch_set_model(self._h, p.ctypes.data_as(_pf32)))

    def arena_ms(self):
        """(device ms of each arena in the last solve, index of each arena's first graph)."""
        n = C.c_int32()
        self.solver._err(_batch_arena_ms(self._h, None, None, 0, C.byref(n)))
        ms = np.zeros(max(n.value, 1), np.float64)
        first = np.zeros(max(n.value, 1), np.int32)
        self.solver._err(_batch_arena_ms(self._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                         first.ctypes.data_as(_p32), n.value, C.byref(n)))
        return ms[:n.value], first[:n.value]

    def thresholds(self):
        t = np.zeros(len(self.graphs), np.int32)
        f = np.zeros(len(self.graphs), np.int32)
        self.solver._err(_batch_thresholds(self._h, t.ctypes.data_as(_p32), f.ctypes.data_as(_p32)))
        return t, f
