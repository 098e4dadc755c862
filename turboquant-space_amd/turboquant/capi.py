"""ctypes binding of the C ABI in include/tq_hip.h (libtq_hip.so).

This is the binding a Python host would add to call the library without the
pybind11 module (tests and benchmarks use it to drive device-resident
buffers). It raises ImportError when the library has not been built: there is
no fallback path.
"""
import ctypes
import os

from . import _runtime  # noqa: F401  (must run before libtq_hip.so loads)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtq_hip.so")

TQ_OK = 0
TQ_EINVAL = -1
TQ_EHIP = -2
TQ_ENOMEM = -3
TQ_EUNSUPPORTED = -4
TQ_MEM_HOST = 0
TQ_MEM_DEVICE = 1
TQ_MODE_FP32 = 0
TQ_MODE_MFMA = 1
TQ_SYM_MSE = 0
TQ_SYM_FULL = 1
TQ_SYM_LIGHT = 2

EXPORTED = [
    "tq_last_error", "tq_device_count", "tq_space_create", "tq_space_destroy",
    "tq_space_info_get", "tq_space_tables", "tq_encode_batch", "tq_prepare_queries",
    "tq_distance_1_to_n", "tq_distance_m_to_n", "tq_distance_m_to_n_noalloc",
    "tq_distance_prepared",
    "tq_distance_m_to_n_symmetric", "tq_distance_m_to_n_symmetric_ex", "tq_prepare_symmetric",
    "tq_distance_symmetric_prepared", "tq_rotated_coords_batch", "tq_encode_graph_create",
    "tq_encode_graph_run", "tq_encode_graph_destroy", "tq_search_plan_create", "tq_search_plan_run",
    "tq_search_plan_run_graph", "tq_search_plan_tune", "tq_search_tuning_lookup",
    "tq_search_plan_stats",
    "tq_search_plan_destroy", "tq_search", "tq_topk_rows", "tq_server_create",
    "tq_server_destroy", "tq_server_add_collection", "tq_server_submit_append",
    "tq_server_submit_search", "tq_server_wait", "tq_server_collection", "tq_server_load_codes", "tq_server_read_codes",
    "tq_server_stats_get",
]
TQ_SEARCH_MAX_K = 128
TQ_FORM_AUTO, TQ_FORM_FUSED, TQ_FORM_BLOCKED = 0, 1, 2


class SpaceInfo(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_uint64),
        ("padded_dim", ctypes.c_uint64),
        ("code_size", ctypes.c_uint64),
        ("packed_bytes", ctypes.c_uint64),
        ("bits_per_coord", ctypes.c_int32),
        ("num_levels", ctypes.c_int32),
        ("scale", ctypes.c_float),
        ("padded", ctypes.c_int32),
        ("prepared_stride", ctypes.c_uint64),
        ("sym_prepared_stride", ctypes.c_uint64),
    ]


class SearchJob(ctypes.Structure):
    _fields_ = [
        ("space", ctypes.c_void_p),
        ("queries", ctypes.c_void_p),
        ("m", ctypes.c_uint64),
        ("codes", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("out_dist", ctypes.c_void_p),
        ("out_ids", ctypes.c_void_p),
    ]


class SearchTuning(ctypes.Structure):
    _fields_ = [("query_tile", ctypes.c_int32), ("items_per_wg", ctypes.c_int32),
                ("form", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("blocked_query_tile", ctypes.c_int32), ("block_mb", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("hbm_budget", ctypes.c_uint64),
        ("slot_bytes", ctypes.c_uint64),
        ("max_inflight", ctypes.c_uint32),
        ("max_batch_queries", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "batches", "searches_done", "appends_done", "queries_done", "failed", "resident_bytes",
        "reserved_bytes", "max_batch_requests")]


class TQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (run turboquant-space_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.tq_last_error.restype = ctypes.c_char_p
        L.tq_last_error.argtypes = []
        L.tq_device_count.restype = I
        L.tq_device_count.argtypes = []
        L.tq_space_create.restype = I
        L.tq_space_create.argtypes = [U64, I, U64, U64, ctypes.POINTER(P)]
        L.tq_space_destroy.restype = I
        L.tq_space_destroy.argtypes = [P]
        L.tq_space_info_get.restype = I
        L.tq_space_info_get.argtypes = [P, ctypes.POINTER(SpaceInfo)]
        L.tq_space_tables.restype = I
        L.tq_space_tables.argtypes = [P, P, P, P, P]
        L.tq_encode_batch.restype = I
        L.tq_encode_batch.argtypes = [P, P, U64, P, I, P]
        L.tq_prepare_queries.restype = I
        L.tq_prepare_queries.argtypes = [P, P, U64, P, I, P]
        L.tq_distance_1_to_n.restype = I
        L.tq_distance_1_to_n.argtypes = [P, P, P, U64, P, I, P]
        L.tq_distance_m_to_n.restype = I
        L.tq_distance_m_to_n.argtypes = [P, P, U64, P, U64, P, I, I, P]
        L.tq_distance_m_to_n_noalloc.restype = I
        L.tq_distance_m_to_n_noalloc.argtypes = [P, P, U64, P, U64, P, I, P]
        L.tq_distance_prepared.restype = I
        L.tq_distance_prepared.argtypes = [P, P, U64, P, U64, P, I, P]
        L.tq_distance_m_to_n_symmetric.restype = I
        L.tq_distance_m_to_n_symmetric.argtypes = [P, P, U64, P, U64, P, I, P]
        L.tq_distance_m_to_n_symmetric_ex.restype = I
        L.tq_distance_m_to_n_symmetric_ex.argtypes = [P, P, U64, P, U64, P, I, I, I, P]
        L.tq_prepare_symmetric.restype = I
        L.tq_prepare_symmetric.argtypes = [P, P, U64, P, I, P]
        L.tq_distance_symmetric_prepared.restype = I
        L.tq_distance_symmetric_prepared.argtypes = [P, P, U64, P, U64, P, I, P]
        L.tq_rotated_coords_batch.restype = I
        L.tq_rotated_coords_batch.argtypes = [P, P, U64, P, I, P]
        JP = ctypes.POINTER(SearchJob)
        TP = ctypes.POINTER(SearchTuning)
        L.tq_search_plan_create.restype = I
        L.tq_search_plan_create.argtypes = [JP, U64, I, I, TP, ctypes.POINTER(P)]
        L.tq_search_plan_run.restype = I
        L.tq_search_plan_run.argtypes = [P, P]
        L.tq_search_plan_run_graph.restype = I
        L.tq_search_plan_run_graph.argtypes = [P, P]
        L.tq_search_plan_tune.restype = I
        L.tq_search_plan_tune.argtypes = [P, P, TP]
        L.tq_search_tuning_lookup.restype = I
        L.tq_search_tuning_lookup.argtypes = [ctypes.c_int32, ctypes.c_int32, U64, TP]
        L.tq_search_plan_stats.restype = I
        L.tq_search_plan_stats.argtypes = [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]
        L.tq_search_plan_destroy.restype = I
        L.tq_search_plan_destroy.argtypes = [P]
        L.tq_search.restype = I
        L.tq_search.argtypes = [JP, U64, I, I, P]
        L.tq_topk_rows.restype = I
        L.tq_topk_rows.argtypes = [P, U64, U64, I, P, P, P]
        L.tq_encode_graph_create.restype = I
        L.tq_encode_graph_create.argtypes = [P, P, U64, P, ctypes.POINTER(P)]
        L.tq_encode_graph_run.restype = I
        L.tq_encode_graph_run.argtypes = [P, P]
        L.tq_encode_graph_destroy.restype = I
        L.tq_encode_graph_destroy.argtypes = [P]
        U32 = ctypes.c_uint32
        L.tq_server_create.restype = I
        L.tq_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
        L.tq_server_destroy.restype = I
        L.tq_server_destroy.argtypes = [P]
        L.tq_server_add_collection.restype = I
        L.tq_server_add_collection.argtypes = [P, P, U64, ctypes.POINTER(U32)]
        L.tq_server_submit_append.restype = I
        L.tq_server_submit_append.argtypes = [P, U32, P, U64, ctypes.POINTER(U64)]
        L.tq_server_submit_search.restype = I
        L.tq_server_submit_search.argtypes = [P, U32, P, U64, P, P, ctypes.POINTER(U64)]
        L.tq_server_wait.restype = I
        L.tq_server_wait.argtypes = [P, U64, ctypes.POINTER(ctypes.c_double)]
        L.tq_server_collection.restype = I
        L.tq_server_collection.argtypes = [P, U32, ctypes.POINTER(U64), ctypes.POINTER(P)]
        L.tq_server_load_codes.restype = I
        L.tq_server_load_codes.argtypes = [P, U32, P, U64, I]
        L.tq_server_read_codes.restype = I
        L.tq_server_read_codes.argtypes = [P, U32, U64, U64, P]
        L.tq_server_stats_get.restype = I
        L.tq_server_stats_get.argtypes = [P, ctypes.POINTER(ServerStats)]
        _lib = L
    return _lib


def check(rc):
    if rc != TQ_OK:
        raise TQError(rc, lib().tq_last_error().decode())


def ptr(x):
    """Raw address of a numpy array, a torch tensor or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


class Space:
    """Owns a tq_space handle."""

    def __init__(self, dim, bits=4, rot_seed=42, qjl_seed=137):
        h = ctypes.c_void_p()
        check(lib().tq_space_create(dim, bits, rot_seed, qjl_seed, ctypes.byref(h)))
        self.h = h
        info = SpaceInfo()
        check(lib().tq_space_info_get(self.h, ctypes.byref(info)))
        self.info = info
        self.dim = int(info.dim)
        self.D = int(info.padded_dim)
        self.cs = int(info.code_size)
        self.bits = bits
        self.prep_stride = int(info.prepared_stride)
        self.sym_prep_stride = int(info.sym_prepared_stride)

    def close(self):
        if self.h:
            lib().tq_space_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode(self, X, n, codes, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_encode_batch(self.h, ptr(X), n, ptr(codes), mem, stream))

    def prepare(self, Q, m, prep, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_prepare_queries(self.h, ptr(Q), m, ptr(prep), mem, stream))

    def scan(self, prep, m, codes, n, out, mode=TQ_MODE_FP32, stream=None):
        check(lib().tq_distance_prepared(self.h, ptr(prep), m, ptr(codes), n, ptr(out), mode,
                                         stream))

    def m_to_n(self, Q, m, codes, n, out, mode=TQ_MODE_FP32, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_distance_m_to_n(self.h, ptr(Q), m, ptr(codes), n, ptr(out), mode, mem,
                                       stream))

    def m_to_n_noalloc(self, Q, m, codes, n, out, mode=TQ_MODE_FP32, stream=None):
        check(lib().tq_distance_m_to_n_noalloc(self.h, ptr(Q), m, ptr(codes), n, ptr(out), mode,
                                               stream))

    def one_to_n(self, q, codes, n, out, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_distance_1_to_n(self.h, ptr(q), ptr(codes), n, ptr(out), mem, stream))

    def symmetric(self, A, m, B, n, out, mem=TQ_MEM_DEVICE, stream=None, variant=TQ_SYM_MSE,
                  mode=TQ_MODE_FP32):
        check(lib().tq_distance_m_to_n_symmetric_ex(self.h, ptr(A), m, ptr(B), n, ptr(out),
                                                    variant, mode, mem, stream))

    def prepare_symmetric(self, A, m, prep, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_prepare_symmetric(self.h, ptr(A), m, ptr(prep), mem, stream))

    def symmetric_prepared(self, prep, m, B, n, out, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_distance_symmetric_prepared(self.h, ptr(prep), m, ptr(B), n, ptr(out), mem,
                                                   stream))

    def rotated(self, X, n, out, mem=TQ_MEM_DEVICE, stream=None):
        check(lib().tq_rotated_coords_batch(self.h, ptr(X), n, ptr(out), mem, stream))


def make_jobs(jobs):
    """ctypes array of SearchJob from (space, queries, m, codes, n, out_dist,
    out_ids) tuples; buffers are device tensors or raw addresses."""
    arr = (SearchJob * len(jobs))()
    for i, (sp, q, m, c, n, od, oi) in enumerate(jobs):
        arr[i] = SearchJob(sp.h.value if isinstance(sp, Space) else sp, ptr(q), m, ptr(c), n,
                           ptr(od), ptr(oi))
    return arr


class EncodeGraph:
    """Owns a tq_encode_graph: the encode of n rows at X into codes (device
    buffers), replayed by run(); each replay reads X when it runs."""

    def __init__(self, space, X, n, codes):
        self._keep = (space, X, codes)  # must outlive the graph
        h = ctypes.c_void_p()
        check(lib().tq_encode_graph_create(space.h, ptr(X), n, ptr(codes), ctypes.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(lib().tq_encode_graph_run(self.h, stream))

    def close(self):
        if self.h:
            lib().tq_encode_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_lookup(padded_dim, bits, m):
    """The autotuner's knobs for (padded dim, bits, m bucket), or None."""
    t = SearchTuning()
    if lib().tq_search_tuning_lookup(padded_dim, bits, m, ctypes.byref(t)) != TQ_OK:
        return None
    return {f: int(getattr(t, f)) for f, _ in SearchTuning._fields_}


class SearchPlan:
    """Owns a tq_search_plan over device-resident job buffers."""

    def __init__(self, jobs, k, mode=TQ_MODE_MFMA, query_tile=0, items_per_wg=0,
                 form=TQ_FORM_AUTO, wg_per_cu=0, blocked_query_tile=0, block_mb=0):
        self._jobs = make_jobs(jobs)
        self._keep = jobs  # the spaces must outlive the plan
        tun = SearchTuning(query_tile, items_per_wg, form, wg_per_cu, blocked_query_tile,
                           block_mb)
        h = ctypes.c_void_p()
        check(lib().tq_search_plan_create(self._jobs, len(jobs), k, mode, ctypes.byref(tun),
                                          ctypes.byref(h)))
        self.h = h
        self.k = k

    def run(self, stream=None):
        check(lib().tq_search_plan_run(self.h, stream))

    def run_graph(self, stream=None):
        check(lib().tq_search_plan_run_graph(self.h, stream))

    def tune(self, stream=None):
        """Autotunes every batch of the plan and rebuilds it; returns the
        fused query tile, items per workgroup and form of the costliest batch
        (all knobs of that batch are in `tuned` afterwards)."""
        t = SearchTuning()
        check(lib().tq_search_plan_tune(self.h, stream, ctypes.byref(t)))
        self.tuned = {f: int(getattr(t, f)) for f, _ in SearchTuning._fields_}
        return int(t.query_tile), int(t.items_per_wg), int(t.form)

    def stats(self):
        items, wgs = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().tq_search_plan_stats(self.h, ctypes.byref(items), ctypes.byref(wgs)))
        return int(items.value), int(wgs.value)

    def close(self):
        if self.h:
            lib().tq_search_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def topk_rows(dist, m, n, k, out_dist, out_ids, stream=None):
    check(lib().tq_topk_rows(ptr(dist), m, n, k, ptr(out_dist), ptr(out_ids), stream))


class Server:
    """Owns a tq_server. Inputs and outputs are host numpy arrays; the server
    reads and writes them asynchronously, so this wrapper holds a reference to
    each request's arrays until its ticket is waited for (the native server
    only has their addresses)."""

    def __init__(self, mode=TQ_MODE_MFMA, k=10, hbm_budget=8 << 30, slot_bytes=0,
                 max_inflight=0, max_batch_queries=0):
        cfg = ServerConfig(mode, k, hbm_budget, slot_bytes, max_inflight, max_batch_queries)
        h = ctypes.c_void_p()
        check(lib().tq_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.k = k
        self._spaces = []
        self._held = {}

    def add_collection(self, space, capacity):
        c = ctypes.c_uint32()
        check(lib().tq_server_add_collection(self.h, space.h, capacity, ctypes.byref(c)))
        self._spaces.append(space)
        return int(c.value)

    def append(self, coll, X):
        t = ctypes.c_uint64()
        check(lib().tq_server_submit_append(self.h, coll, ptr(X), len(X), ctypes.byref(t)))
        self._held[int(t.value)] = (X,)
        return int(t.value)

    def search(self, coll, Q, out_d, out_i):
        t = ctypes.c_uint64()
        check(lib().tq_server_submit_search(self.h, coll, ptr(Q), len(Q), ptr(out_d), ptr(out_i),
                                            ctypes.byref(t)))
        self._held[int(t.value)] = (Q, out_d, out_i)
        return int(t.value)

    def wait(self, ticket):
        lat = ctypes.c_double()
        try:
            check(lib().tq_server_wait(self.h, ticket, ctypes.byref(lat)))
        finally:
            self._held.pop(int(ticket), None)
        return float(lat.value)

    def collection(self, coll):
        n, codes = ctypes.c_uint64(), ctypes.c_void_p()
        check(lib().tq_server_collection(self.h, coll, ctypes.byref(n), ctypes.byref(codes)))
        return int(n.value), codes.value

    def load_codes(self, coll, codes, n, mem=TQ_MEM_DEVICE):
        check(lib().tq_server_load_codes(self.h, coll, ptr(codes), n, mem))

    def read_codes(self, coll, start, count, code_size):
        import numpy as np
        out = np.empty((count, code_size), np.uint8)
        check(lib().tq_server_read_codes(self.h, coll, start, count, ptr(out)))
        return out

    def stats(self):
        st = ServerStats()
        check(lib().tq_server_stats_get(self.h, ctypes.byref(st)))
        return {name: int(getattr(st, name)) for name, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().tq_server_destroy(self.h)
            self.h = None
        self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
