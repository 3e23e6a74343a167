"""Python front end of libdsm_amd.so (include/dsm.h) over ctypes.

Mirrors the reference's matching path (CDSM, main.cu:152-348): build the
engine once from the data graph and the standing queries, then for every update
batch apply it (updateHostGraph + clusterFilter + compact, cpuGraph.cpp:521,
filter.cu:1525, filter.cu:1692) and match it (seeding main.cu:268-329 +
parallelMatch gpu_match.cu:1121), which yields one positive and one negative
uint64 count per query.

There is no CPU fallback: every call goes to the HIP engine, and loading fails
loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DSM_OK = 0
DSM_EINVAL = -1
DSM_ENOMEM = -2
DSM_EDEVICE = -3
DSM_EFORMAT = -4
DSM_EIO = -5
DSM_ELIMIT = -6
DSM_ETRUNC = -7
DSM_EADMIT = -8
MAX_QUERY_VERTICES = 16
TUPLE_WORDS = MAX_QUERY_VERTICES + 2

# every entry point declared in include/dsm.h
EXPORTS = (
    "dsm_options_default", "dsm_create", "dsm_destroy", "dsm_last_error", "dsm_apply_batch",
    "dsm_match_batch", "dsm_process_batch", "dsm_get_tuples", "dsm_get_stats", "dsm_get_graph",
    "dsm_get_candidates", "dsm_get_query_time", "dsm_num_vertices", "dsm_num_queries", "dsm_file_read", "dsm_file_free",
    "dsm_file_graph", "dsm_file_query", "dsm_file_num_batches", "dsm_file_batch", "dsm_file_error",
    "dsm_file_edge_labels",
    "dsm_server_create", "dsm_server_submit", "dsm_server_wait", "dsm_server_error", "dsm_server_destroy",
    "dsm_admit", "dsm_estimate_request_bytes", "dsm_autotune", "dsm_process_batches",
    "dsm_get_refined_candidates", "dsm_cluster_plan", "dsm_intersect_sorted", "dsm_intersect_sorted_noalloc",
)


class DsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dsm error %d: %s" % (code, msg))
        self.code = code


class Change(C.Structure):
    _fields_ = [("add", C.c_uint32), ("src", C.c_uint32), ("dst", C.c_uint32), ("label", C.c_uint32)]


class GraphDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("labels", C.c_void_p), ("num_edges", C.c_uint64),
                ("edge_src", C.c_void_p), ("edge_dst", C.c_void_p)]


class QueryDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("labels", C.c_void_p), ("num_edges", C.c_uint32),
                ("edge_src", C.c_void_p), ("edge_dst", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("device", C.c_int), ("relabel", C.c_int), ("count_negatives", C.c_int),
                ("tuple_capacity", C.c_uint64), ("tuples_in_host_memory", C.c_int),
                ("waves_per_block", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("donate_min", C.c_uint32), ("queue_capacity", C.c_uint32), ("watchdog_ms", C.c_uint32),
                ("lane_work", C.c_uint32), ("leaf_sets", C.c_int), ("donate_every", C.c_uint32),
                ("donate_min_n2", C.c_uint32), ("poll_max", C.c_uint32), ("leaf_segments", C.c_int),
                ("graph_rounds", C.c_int), ("overlap_apply", C.c_int), ("leaf_budget", C.c_uint32),
                ("cluster_filter", C.c_int), ("bit_probe_last", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("donate_min", C.c_uint32),
                ("lane_work", C.c_uint32), ("leaf_sets", C.c_int), ("leaf_segments", C.c_int),
                ("donate_every", C.c_uint32), ("trials", C.c_uint32),
                ("ms_initial", C.c_float), ("ms_best", C.c_float), ("leaf_budget", C.c_uint32),
                ("bit_probe_last", C.c_int), ("order_changed", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("seeds_pos", C.c_uint64), ("seeds_neg", C.c_uint64), ("tasks_donated", C.c_uint64),
                ("tuples", C.c_uint64), ("touched_vertices", C.c_uint32),
                ("relabeled_vertices", C.c_uint32), ("dirty_groups", C.c_uint32),
                ("relocated_vertices", C.c_uint32), ("ms_apply", C.c_float), ("ms_filter", C.c_float),
                ("ms_seed", C.c_float), ("ms_dfs_pos", C.c_float), ("ms_dfs_neg", C.c_float),
                ("pool_entries_used", C.c_uint64), ("pool_entries_capacity", C.c_uint64),
                ("dfs_entries", C.c_uint64), ("dfs_search_loads", C.c_uint64), ("dfs_launches", C.c_uint32),
                ("dfs_merge_bytes", C.c_uint64), ("dfs_busy_ms", C.c_double), ("dfs_waves", C.c_uint64),
                ("dfs_block_imbalance", C.c_double), ("dfs_head_idle", C.c_double), ("dfs_tail_idle", C.c_double),
                ("ms_refine", C.c_float), ("filter_entries", C.c_uint64),
                ("filter_bits", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOptions(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("hbm_headroom", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("queue_depth", C.c_uint32), ("fused", C.c_uint32),
                ("num_changes", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def library_path():
    """The in-tree build; DSM_AMD_LIBRARY names another build of the same library."""
    return os.environ.get("DSM_AMD_LIBRARY") or os.path.join(_HERE, "libdsm_amd.so")


def lib():
    """Loads libdsm_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError("libdsm_amd.so not built (run __graft_entry__.build()); no fallback exists")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.dsm_options_default.argtypes = [C.POINTER(Options)]
    L.dsm_create.restype = C.c_int
    L.dsm_create.argtypes = [C.POINTER(GraphDesc), C.POINTER(QueryDesc), C.c_uint32, C.POINTER(Options),
                             C.POINTER(vp)]
    L.dsm_destroy.argtypes = [vp]
    L.dsm_last_error.restype = C.c_char_p
    L.dsm_last_error.argtypes = [vp]
    L.dsm_apply_batch.restype = C.c_int
    L.dsm_apply_batch.argtypes = [vp, C.POINTER(Change), C.c_uint32]
    L.dsm_match_batch.restype = C.c_int
    L.dsm_match_batch.argtypes = [vp, vp, vp]
    L.dsm_process_batch.restype = C.c_int
    L.dsm_process_batch.argtypes = [vp, C.POINTER(Change), C.c_uint32, vp, vp]
    L.dsm_get_tuples.restype = C.c_int
    L.dsm_get_tuples.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    L.dsm_get_stats.restype = C.c_int
    L.dsm_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.dsm_get_graph.restype = C.c_int
    L.dsm_get_graph.argtypes = [vp, vp, vp, C.c_uint64, vp, C.POINTER(C.c_uint64)]
    L.dsm_get_candidates.restype = C.c_int
    L.dsm_get_candidates.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
    L.dsm_get_refined_candidates.restype = C.c_int
    L.dsm_get_refined_candidates.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
    L.dsm_intersect_sorted.restype = C.c_int
    L.dsm_intersect_sorted.argtypes = [C.c_int, vp, C.c_uint32, vp, C.c_uint32, vp, C.POINTER(C.c_uint32)]
    L.dsm_intersect_sorted_noalloc.restype = C.c_int
    L.dsm_intersect_sorted_noalloc.argtypes = [C.c_int, vp, C.c_uint32, vp, C.c_uint32, vp, C.POINTER(C.c_uint32)]
    L.dsm_cluster_plan.restype = C.c_uint32
    L.dsm_cluster_plan.argtypes = [C.c_uint32, C.c_uint32, vp, vp, vp, C.c_uint32]
    L.dsm_get_query_time.restype = C.c_int
    L.dsm_get_query_time.argtypes = [vp, vp, C.c_uint32]
    L.dsm_num_vertices.restype = C.c_uint32
    L.dsm_num_vertices.argtypes = [vp]
    L.dsm_num_queries.restype = C.c_uint32
    L.dsm_num_queries.argtypes = [vp]
    L.dsm_file_read.restype = C.c_int
    L.dsm_file_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
    L.dsm_file_free.argtypes = [vp]
    L.dsm_file_graph.restype = C.c_int
    L.dsm_file_graph.argtypes = [vp, C.POINTER(GraphDesc)]
    L.dsm_file_query.restype = C.c_int
    L.dsm_file_query.argtypes = [vp, C.POINTER(QueryDesc)]
    L.dsm_file_num_batches.restype = C.c_uint32
    L.dsm_file_num_batches.argtypes = [vp]
    L.dsm_file_batch.restype = C.c_int
    L.dsm_file_batch.argtypes = [vp, C.c_uint32, C.POINTER(C.POINTER(Change)), C.POINTER(C.c_uint32)]
    L.dsm_file_edge_labels.restype = C.c_int
    L.dsm_file_edge_labels.argtypes = [vp, C.POINTER(C.POINTER(C.c_uint32))]
    L.dsm_file_error.restype = C.c_char_p
    L.dsm_file_error.argtypes = []
    L.dsm_server_create.restype = C.c_int
    L.dsm_server_create.argtypes = [vp, C.POINTER(ServerOptions), C.POINTER(vp)]
    L.dsm_server_submit.restype = C.c_int
    L.dsm_server_submit.argtypes = [vp, C.POINTER(Change), C.c_uint32, C.POINTER(C.c_uint64)]
    L.dsm_server_wait.restype = C.c_int
    L.dsm_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(RequestInfo), vp, vp]
    L.dsm_server_error.restype = C.c_char_p
    L.dsm_server_error.argtypes = [vp]
    L.dsm_server_destroy.argtypes = [vp]
    L.dsm_admit.restype = C.c_int
    L.dsm_admit.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.dsm_process_batches.restype = C.c_int
    L.dsm_process_batches.argtypes = [vp, C.POINTER(Change), vp, C.c_uint32, vp, vp, vp]
    L.dsm_autotune.restype = C.c_int
    L.dsm_autotune.argtypes = [vp, C.POINTER(Change), C.c_uint32, C.POINTER(TuneReport)]
    L.dsm_estimate_request_bytes.restype = C.c_uint64
    L.dsm_estimate_request_bytes.argtypes = [vp, C.c_uint32]
    _LIB = L
    return L


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def changes_array(add, src, dst, label):
    """Packs parallel arrays into a ctypes dsm_change[] (16 B records)."""
    k = len(src)
    rec = np.zeros((k, 4), np.uint32)
    rec[:, 0] = np.asarray(add, np.uint32)
    rec[:, 1] = _u32(src)
    rec[:, 2] = _u32(dst)
    rec[:, 3] = _u32(label)
    return rec


class DsmFile:
    """A DSM text file parsed by the engine's own reader (dsm_file_read)."""

    def __init__(self, path, is_query):
        L = lib()
        h = C.c_void_p()
        rc = L.dsm_file_read(path.encode(), int(bool(is_query)), C.byref(h))
        if rc != DSM_OK:
            raise DsmError(rc, L.dsm_file_error().decode())
        self._h = h
        self.is_query = bool(is_query)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().dsm_file_free(self._h)
            self._h = None

    def graph(self):
        g = GraphDesc()
        lib().dsm_file_graph(self._h, C.byref(g))
        n, m = g.num_vertices, g.num_edges
        lab = np.ctypeslib.as_array(C.cast(g.labels, C.POINTER(C.c_uint32)), (max(n, 1),))[:n].copy()
        eu = np.ctypeslib.as_array(C.cast(g.edge_src, C.POINTER(C.c_uint32)), (max(m, 1),))[:m].copy() if m else np.zeros(0, np.uint32)
        ev = np.ctypeslib.as_array(C.cast(g.edge_dst, C.POINTER(C.c_uint32)), (max(m, 1),))[:m].copy() if m else np.zeros(0, np.uint32)
        return n, lab, eu, ev

    def edge_labels(self):
        """The e lines' edge labels (0 where absent), or None when no line has one."""
        p = C.POINTER(C.c_uint32)()
        has = lib().dsm_file_edge_labels(self._h, C.byref(p))
        if has < 0:
            raise DsmError(has, "dsm_file_edge_labels")
        if not has:
            return None
        m = self.graph()[2].size
        return np.ctypeslib.as_array(p, (m,)).copy()

    @property
    def num_batches(self):
        return lib().dsm_file_num_batches(self._h)

    def batch(self, b):
        """(add u8, src u32, dst u32, label u32) arrays of batch b."""
        p = C.POINTER(Change)()
        k = C.c_uint32()
        rc = lib().dsm_file_batch(self._h, b, C.byref(p), C.byref(k))
        if rc != DSM_OK:
            raise DsmError(rc, "no batch %d" % b)
        if k.value == 0:
            z = np.zeros(0, np.uint32)
            return z.astype(np.uint8), z, z, z
        arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (k.value * 4,)).reshape(-1, 4).copy()
        return arr[:, 0].astype(np.uint8), arr[:, 1], arr[:, 2], arr[:, 3]


class Engine:
    """One continuous matching engine on one GPU.

    graph: (n, labels, eu, ev); queries: list of (n, labels, eu, ev)."""

    def __init__(self, graph, queries, device=0, relabel=True, count_negatives=True, tuple_capacity=0,
                 tuples_in_host_memory=False, waves_per_block=4, blocks_per_cu=0, donate_min=0,
                 queue_capacity=0, watchdog_ms=0, lane_work=0, leaf_sets=True, graph_rounds=True,
                 donate_every=0, donate_min_n2=0, leaf_segments=True, poll_max=0, overlap_apply=True,
                 leaf_budget=0, cluster_filter=False, bit_probe_last=False):
        L = lib()
        n, lab, eu, ev = graph
        self._keep = []
        lab, eu, ev = _u32(lab), _u32(eu), _u32(ev)
        self._keep += [lab, eu, ev]
        gd = GraphDesc(int(n), lab.ctypes.data, int(len(eu)), eu.ctypes.data, ev.ctypes.data)
        qds = (QueryDesc * len(queries))()
        for i, (qn, ql, qu, qv) in enumerate(queries):
            ql, qu, qv = _u32(ql), _u32(qu), _u32(qv)
            self._keep += [ql, qu, qv]
            qds[i] = QueryDesc(int(qn), ql.ctypes.data, int(len(qu)), qu.ctypes.data, qv.ctypes.data)
        opt = Options()
        L.dsm_options_default(C.byref(opt))
        opt.device = device
        opt.relabel = int(bool(relabel))
        opt.count_negatives = int(bool(count_negatives))
        opt.tuple_capacity = int(tuple_capacity)
        opt.tuples_in_host_memory = int(bool(tuples_in_host_memory))
        opt.waves_per_block = waves_per_block
        opt.blocks_per_cu = blocks_per_cu
        opt.donate_min = donate_min
        opt.queue_capacity = queue_capacity
        opt.watchdog_ms = watchdog_ms
        opt.lane_work = lane_work
        opt.leaf_sets = int(bool(leaf_sets))
        opt.graph_rounds = int(bool(graph_rounds))
        opt.donate_every = donate_every
        opt.donate_min_n2 = donate_min_n2
        opt.leaf_segments = int(bool(leaf_segments))
        opt.poll_max = poll_max
        opt.overlap_apply = int(bool(overlap_apply))
        opt.leaf_budget = leaf_budget
        opt.cluster_filter = int(bool(cluster_filter))
        opt.bit_probe_last = int(bool(bit_probe_last))
        h = C.c_void_p()
        rc = L.dsm_create(C.byref(gd), qds, len(queries), C.byref(opt), C.byref(h))
        self._keep = None
        if rc != DSM_OK:
            raise DsmError(rc, L.dsm_last_error(None).decode())
        self._h = h
        self.n = int(n)
        self.nq = len(queries)
        self.query_sizes = [int(q[0]) for q in queries]

    def close(self):
        if getattr(self, "_h", None):
            lib().dsm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != DSM_OK:
            raise DsmError(rc, lib().dsm_last_error(self._h).decode())

    def apply_batch(self, add, src, dst, label):
        rec = changes_array(add, src, dst, label)
        self._check(lib().dsm_apply_batch(self._h, rec.ctypes.data_as(C.POINTER(Change)), rec.shape[0]))

    def match_batch(self):
        pos = np.zeros(self.nq, np.uint64)
        neg = np.zeros(self.nq, np.uint64)
        self._check(lib().dsm_match_batch(self._h, pos.ctypes.data, neg.ctypes.data))
        return pos, neg

    def process_batch(self, add, src, dst, label):
        rec = changes_array(add, src, dst, label)
        pos = np.zeros(self.nq, np.uint64)
        neg = np.zeros(self.nq, np.uint64)
        self._check(lib().dsm_process_batch(self._h, rec.ctypes.data_as(C.POINTER(Change)), rec.shape[0],
                                            pos.ctypes.data, neg.ctypes.data))
        return pos, neg

    def process_batches(self, batches):
        """dsm_process_batches over a list of (add, src, dst, label) batches:
        (pos u64[nb, nq], neg u64[nb, nq], done_s f64[nb])."""
        recs = [changes_array(*b) for b in batches]
        nb = len(recs)
        off = np.zeros(nb + 1, np.uint32)
        off[1:] = np.cumsum([r.shape[0] for r in recs])
        allrec = np.concatenate(recs) if nb else np.zeros((0, 4), np.uint32)
        allrec = np.ascontiguousarray(allrec if len(allrec) else np.zeros((1, 4), np.uint32))
        pos = np.zeros((max(nb, 1), self.nq), np.uint64)
        neg = np.zeros((max(nb, 1), self.nq), np.uint64)
        done = np.zeros(max(nb, 1), np.float64)
        self._check(lib().dsm_process_batches(self._h, allrec.ctypes.data_as(C.POINTER(Change)), off.ctypes.data,
                                              nb, pos.ctypes.data, neg.ctypes.data, done.ctypes.data))
        return pos[:nb], neg[:nb], done[:nb]

    def autotune(self, add, src, dst, label):
        """dsm_autotune over these records; returns the report as a dict."""
        rec = changes_array(add, src, dst, label)
        rep = TuneReport()
        self._check(lib().dsm_autotune(self._h, rec.ctypes.data_as(C.POINTER(Change)), rec.shape[0],
                                       C.byref(rep)))
        return rep.as_dict()

    def tuples(self, cap=None):
        """Tuples of the last batch: dict (query, sign) -> (N, n) uint32 array."""
        L = lib()
        num = C.c_uint64()
        L.dsm_get_tuples(self._h, None, 0, C.byref(num))
        k = num.value if cap is None else min(cap, num.value)
        buf = np.zeros(max(k, 1) * TUPLE_WORDS, np.uint32)
        rc = L.dsm_get_tuples(self._h, buf.ctypes.data, k, C.byref(num))
        self._check(rc)
        rows = buf[: k * TUPLE_WORDS].reshape(k, TUPLE_WORDS)
        out = {}
        for q in range(self.nq):
            for sign in (0, 1):
                sel = rows[(rows[:, 0] == q) & (rows[:, 1] == sign)]
                out[(q, sign)] = sel[:, 2:2 + self.query_sizes[q]].copy()
        return out

    def stats(self):
        s = Stats()
        self._check(lib().dsm_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def graph(self):
        """Device graph as (offsets u64[n+1], neighbours u32[], labels u32[n])."""
        L = lib()
        off = np.zeros(self.n + 1, np.uint64)
        lab = np.zeros(self.n, np.uint32)
        ne = C.c_uint64()
        L.dsm_get_graph(self._h, off.ctypes.data, None, 0, lab.ctypes.data, C.byref(ne))
        nb = np.zeros(max(ne.value, 1), np.uint32)
        self._check(L.dsm_get_graph(self._h, off.ctypes.data, nb.ctypes.data, ne.value, lab.ctypes.data,
                                    C.byref(ne)))
        return off, nb[: ne.value], lab

    def candidates(self, q, u):
        """Candidate bitmap of query vertex u of query q as a bool array over vertices."""
        words = np.zeros((self.n + 63) // 64, np.uint64)
        self._check(lib().dsm_get_candidates(self._h, q, u, words.ctypes.data))
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return bits[: self.n].astype(bool)

    def refined_candidates(self, q, u):
        """The candidates the seeding and the DFS use for query vertex u of
        query q (the cluster-encoding refinement with cluster_filter, else the
        LDF/NLF set), as a bool array over vertices."""
        words = np.zeros((self.n + 63) // 64, np.uint64)
        self._check(lib().dsm_get_refined_candidates(self._h, q, u, words.ctypes.data))
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return bits[: self.n].astype(bool)

    def query_time(self):
        """DFS task time per query in ms since the last call (resets it)."""
        ms = np.zeros(self.nq, np.float64)
        self._check(lib().dsm_get_query_time(self._h, ms.ctypes.data, self.nq))
        return ms


def cluster_plan_words(n, edges):
    """The engine's cluster plan of a query (dsm_cluster_plan) as a word
    array, or None past the reference's limits."""
    eu = _u32([a for a, _ in edges])
    ev = _u32([b for _, b in edges])
    buf = np.zeros(1 << 16, np.uint32)
    w = lib().dsm_cluster_plan(int(n), len(eu), eu.ctypes.data, ev.ctypes.data, buf.ctypes.data, buf.shape[0])
    return buf[:w].copy() if w else None


def intersect_sorted(a, b, device=0, noalloc=False):
    """a ∩ b on the device (dsm_intersect_sorted, or with noalloc
    dsm_intersect_sorted_noalloc); a and b strictly ascending."""
    L = lib()
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, min(len(a), len(b))), np.uint32)
    n = C.c_uint32()
    fn = L.dsm_intersect_sorted_noalloc if noalloc else L.dsm_intersect_sorted
    rc = fn(device, a.ctypes.data, len(a), b.ctypes.data, len(b), out.ctypes.data, C.byref(n))
    if rc != DSM_OK:
        raise DsmError(rc, L.dsm_last_error(None).decode())
    return out[: n.value].copy()


def admit(request_bytes, in_flight_bytes, budget):
    """The server's admission rule (dsm_admit): 1 admit, 0 wait, DSM_EADMIT."""
    return int(lib().dsm_admit(int(request_bytes), int(in_flight_bytes), int(budget)))


class Server:
    """Served stream over one Engine (dsm_server_*): submit update batches from
    any thread, wait for each one's counts. The engine is not used directly
    while the server exists."""

    def __init__(self, engine, budget_bytes=0, hbm_headroom=0.0):
        o = ServerOptions(int(budget_bytes), float(hbm_headroom))
        h = C.c_void_p()
        rc = lib().dsm_server_create(engine._h, C.byref(o), C.byref(h))
        if rc != DSM_OK:
            raise DsmError(rc, lib().dsm_last_error(engine._h).decode())
        self._h = h
        self.engine = engine
        self.nq = engine.nq

    def estimate_bytes(self, num_changes):
        return int(lib().dsm_estimate_request_bytes(self.engine._h, int(num_changes)))

    def submit(self, add, src, dst, label):
        rec = changes_array(add, src, dst, label)
        rid = C.c_uint64()
        rc = lib().dsm_server_submit(self._h, rec.ctypes.data_as(C.POINTER(Change)), rec.shape[0], C.byref(rid))
        if rc != DSM_OK:
            raise DsmError(rc, "submit rejected")
        return rid.value

    def wait(self, rid):
        """(status, pos u64[nq], neg u64[nq], info dict) of request rid."""
        pos = np.zeros(self.nq, np.uint64)
        neg = np.zeros(self.nq, np.uint64)
        info = RequestInfo()
        st = lib().dsm_server_wait(self._h, int(rid), C.byref(info), pos.ctypes.data, neg.ctypes.data)
        return st, pos, neg, info.as_dict()

    def error(self):
        return lib().dsm_server_error(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().dsm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
