"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product path (dsm_amd) never imports it.

Semantics restated from the reference (file:line in oracle.c's header):
updateHostGraph (cpuGraph.cpp:521-607), LDF/NLF (filter.cu:394-566), edge
seeding (main.cu:268-329) and the DFS candidate rule (gpu_match.cu:96-325).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    L.og_new.restype = vp
    L.og_new.argtypes = [C.c_uint32, u32p, C.c_uint64, u32p, u32p]
    L.og_free.argtypes = [vp]
    L.og_num_entries.restype = C.c_uint64
    L.og_num_entries.argtypes = [vp]
    L.og_export.argtypes = [vp, u64p, u32p, u32p]
    L.og_apply.restype = C.c_int
    L.og_apply.argtypes = [vp, C.c_uint32, u8p, u32p, u32p, u32p, C.c_int]
    L.oq_new.restype = vp
    L.oq_new.argtypes = [C.c_uint32, u32p, C.c_uint32, u32p, u32p]
    L.oq_free.argtypes = [vp]
    L.og_ldf_nlf.argtypes = [vp, vp, u8p]
    L.og_count_records.argtypes = [vp, vp, u8p, C.c_uint32, u32p, u32p, u64p]
    L.og_tuples_records.restype = C.c_uint64
    L.og_tuples_records.argtypes = [vp, vp, u8p, C.c_uint32, u32p, u32p, vp, C.c_uint64]
    L.og_brute_record.restype = C.c_uint64
    L.og_brute_record.argtypes = [vp, vp, C.c_uint32, C.c_uint32]
    L.og_brute_total.restype = C.c_uint64
    L.og_brute_total.argtypes = [vp, vp, vp, C.c_uint64]
    L.oq_cluster_plan.restype = C.c_uint32
    L.oq_cluster_plan.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, u32p, C.c_uint32]
    L.og_cluster_filter.restype = C.c_int
    L.og_cluster_filter.argtypes = [vp, C.c_uint32, C.c_uint32, u32p, u32p, u8p, u8p]
    L.of_read.restype = vp
    L.of_read.argtypes = [C.c_char_p, C.c_int]
    L.of_free.argtypes = [vp]
    for f in ("of_n", "of_m", "of_nb"):
        getattr(L, f).restype = C.c_uint32
        getattr(L, f).argtypes = [vp]
    L.of_nchanges.restype = C.c_uint64
    L.of_nchanges.argtypes = [vp]
    L.of_copy.argtypes = [vp, u32p, u32p, u32p, u32p]
    L.of_copy_changes.argtypes = [vp, u32p, u8p, u32p, u32p, u32p]
    _LIB = L
    return L


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class Graph:
    """Data graph state: sorted, deduplicated adjacency sets + labels."""

    def __init__(self, n, labels, eu, ev):
        L = lib()
        labels, eu, ev = _u32(labels), _u32(eu), _u32(ev)
        assert labels.shape[0] == n and eu.shape == ev.shape
        self.n = int(n)
        self._h = L.og_new(n, labels, eu.shape[0], eu, ev)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().og_free(self._h)
            self._h = None

    def apply(self, add, src, dst, label, relabel=True):
        add = np.ascontiguousarray(np.asarray(add, dtype=np.uint8))
        src, dst, label = _u32(src), _u32(dst), _u32(label)
        rc = lib().og_apply(self._h, add.shape[0], add, src, dst, label, int(bool(relabel)))
        if rc != 0:
            raise ValueError("change endpoint out of range")

    def export(self):
        """(offsets uint64[n+1], neighbours uint32[], labels uint32[n])"""
        L = lib()
        ne = L.og_num_entries(self._h)
        off = np.zeros(self.n + 1, np.uint64)
        nb = np.zeros(max(ne, 1), np.uint32)
        lab = np.zeros(self.n, np.uint32)
        L.og_export(self._h, off, nb, lab)
        return off, nb[:ne], lab

    def candidates(self, q):
        out = np.zeros(q.n * self.n, np.uint8)
        lib().og_ldf_nlf(self._h, q._h, out)
        return out.reshape(q.n, self.n)

    def count_records(self, q, src, dst, cand=None):
        src, dst = _u32(src), _u32(dst)
        if cand is None:
            cand = self.candidates(q)
        out = np.zeros(max(src.shape[0], 1), np.uint64)
        lib().og_count_records(self._h, q._h, np.ascontiguousarray(cand.reshape(-1)),
                               src.shape[0], src, dst, out)
        return out[: src.shape[0]]

    def tuples_records(self, q, src, dst, cand=None):
        """All seeded embeddings as an (N, q.n) uint32 array, in seed order."""
        src, dst = _u32(src), _u32(dst)
        if cand is None:
            cand = self.candidates(q)
        c = np.ascontiguousarray(cand.reshape(-1))
        L = lib()
        n = L.og_tuples_records(self._h, q._h, c, src.shape[0], src, dst, None, 0)
        buf = np.zeros(max(n, 1) * q.n, np.uint32)
        L.og_tuples_records(self._h, q._h, c, src.shape[0], src, dst, buf.ctypes.data, n)
        return buf[: n * q.n].reshape(n, q.n)

    def cluster_candidates(self, q, ldf=None):
        """Cluster-encoding refinement of q's candidates (clusterFilter,
        filter.cu:1525-1654) as q.n rows of bytes; the plan follows q's edge
        lines in the order given."""
        if ldf is None:
            ldf = self.candidates(q)
        eu = _u32([a for a, _ in q.edges])
        ev = _u32([b for _, b in q.edges])
        out = np.zeros(q.n * self.n, np.uint8)
        rc = lib().og_cluster_filter(self._h, q.n, eu.shape[0], eu, ev,
                                     np.ascontiguousarray(ldf.reshape(-1), np.uint8), out)
        if rc != 0:
            raise ValueError("cluster plan exceeds a limit of the reference")
        return out.reshape(q.n, self.n)

    def brute_record(self, q, s, d):
        return int(lib().og_brute_record(self._h, q._h, int(s), int(d)))

    def brute_total(self, q):
        return int(lib().og_brute_total(self._h, q._h, None, 0))

    def brute_embeddings(self, q):
        """Every embedding of q in the graph, by brute force, as (N, q.n)."""
        L = lib()
        n = L.og_brute_total(self._h, q._h, None, 0)
        buf = np.zeros(max(n, 1) * q.n, np.uint32)
        L.og_brute_total(self._h, q._h, buf.ctypes.data, n)
        return buf[: n * q.n].reshape(n, q.n)


class Query:
    def __init__(self, n, labels, eu, ev):
        labels, eu, ev = _u32(labels), _u32(eu), _u32(ev)
        self.n = int(n)
        self.labels = labels
        self.edges = list(zip(eu.tolist(), ev.tolist()))
        self._h = lib().oq_new(n, labels, eu.shape[0], eu, ev)
        if not self._h:
            raise ValueError("invalid query")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oq_free(self._h)
            self._h = None


OP_NAMES = {0: "encode", 1: "merge", 2: "combine"}


def cluster_plan_words(n, edges):
    """The oracle's plan in the flat word layout (oq_cluster_plan)."""
    eu = _u32([a for a, _ in edges])
    ev = _u32([b for _, b in edges])
    buf = np.zeros(1 << 16, np.uint32)
    w = lib().oq_cluster_plan(n, eu.shape[0], eu, ev, buf, buf.shape[0])
    return buf[:w].copy() if w else None


def cluster_plan(n, edges):
    """The cluster plan of a query (decycle.cpp:6-135, filter.cu:23-392,
    encode()'s pass order filter.cu:936-1177, enc_pos_of_u filter.cu:1589-1609)
    as a dict, for edge lines in the order given; None past a reference limit."""
    eu = _u32([a for a, _ in edges])
    ev = _u32([b for _, b in edges])
    buf = np.zeros(1 << 16, np.uint32)
    w = lib().oq_cluster_plan(n, eu.shape[0], eu, ev, buf, buf.shape[0])
    if w == 0:
        return None
    it = iter(buf[:w].tolist())
    nx = lambda: next(it)  # noqa: E731
    out = {"n": nx()}
    out["cover"] = [nx() for _ in range(nx())]
    clusters = []
    for _ in range(nx()):
        valid = nx()
        clusters.append({"valid": bool(valid), "members": [nx() for _ in range(nx())]})
    out["clusters"] = clusters
    out["final"] = nx()
    out["enc_pos"] = [nx() for _ in range(out["n"])]
    out["ops"] = [(OP_NAMES[nx()], nx()) for _ in range(nx())]
    return out


class DsmFile:
    """Parsed DSM text file (io.cu:122-252 restated)."""

    def __init__(self, path, is_query):
        L = lib()
        h = L.of_read(path.encode(), int(bool(is_query)))
        if not h:
            raise ValueError("malformed DSM file: %s" % path)
        try:
            self.n, self.m, self.nb = L.of_n(h), L.of_m(h), L.of_nb(h)
            nc = L.of_nchanges(h)
            self.labels = np.zeros(max(self.n, 1), np.uint32)
            self.degs = np.zeros(max(self.n, 1), np.uint32)
            self.eu = np.zeros(max(self.m, 1), np.uint32)
            self.ev = np.zeros(max(self.m, 1), np.uint32)
            L.of_copy(h, self.labels, self.degs, self.eu, self.ev)
            self.labels, self.degs = self.labels[: self.n], self.degs[: self.n]
            self.eu, self.ev = self.eu[: self.m], self.ev[: self.m]
            self.bstart = np.zeros(self.nb + 1, np.uint32)
            self.add = np.zeros(max(nc, 1), np.uint8)
            self.src = np.zeros(max(nc, 1), np.uint32)
            self.dst = np.zeros(max(nc, 1), np.uint32)
            self.clab = np.zeros(max(nc, 1), np.uint32)
            L.of_copy_changes(h, self.bstart, self.add, self.src, self.dst, self.clab)
        finally:
            L.of_free(h)

    def batch(self, b):
        s, e = int(self.bstart[b]), int(self.bstart[b + 1])
        return self.add[s:e], self.src[s:e], self.dst[s:e], self.clab[s:e]


def run_stream_threaded(graph, queries, batches, relabel=True, threads=None, parts=8, log=None):
    """run_stream's counts (no tuples) with the counting spread over host
    threads: per phase, the candidate sets of every query, then (query, 1/parts
    of the records) tasks. A count is a sum over records (main.cu:268-329), so
    the split does not change it. The oracle's C calls release the GIL. log
    (optional) is called with a line as tasks finish."""
    import time
    from concurrent.futures import ThreadPoolExecutor, as_completed
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    out = []
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        def phase(name, s_, d_):
            cands = list(ex.map(graph.candidates, queries))
            chunks = [c for c in np.array_split(np.arange(len(s_)), parts) if len(c)]
            tasks = [(qi, c) for qi in range(len(queries)) for c in chunks]
            futs = {ex.submit(lambda t: int(graph.count_records(queries[t[0]], s_[t[1]], d_[t[1]],
                                                                cands[t[0]]).sum()), t): t for t in tasks}
            tot = [0] * len(queries)
            for i, f in enumerate(as_completed(futs)):
                tot[futs[f][0]] += f.result()
                if log and (i + 1) % max(1, len(tasks) // 8) == 0:
                    log("  oracle %s: %d/%d tasks, %.1f s" % (name, i + 1, len(tasks), time.perf_counter() - t0))
            return tot

        for add, src, dst, lab in batches:
            add = np.asarray(add, np.uint8)
            src, dst = _u32(src), _u32(dst)
            dels = add == 0
            res = {"neg": phase("negatives", src[dels], dst[dels])}
            graph.apply(add, src, dst, lab, relabel)
            res["pos"] = phase("positives", src[~dels], dst[~dels])
            out.append(res)
    return out


def run_stream(graph, queries, batches, relabel=True, tuples=False):
    """Per batch: negatives (delete records seeded in the pre-batch graph), then
    the update rule, then positives (insert records seeded in the post-batch
    graph). Returns a list of dicts {pos, neg[, pos_tuples, neg_tuples]} with
    one entry per query."""
    out = []
    for add, src, dst, lab in batches:
        add = np.asarray(add, np.uint8)
        src, dst = _u32(src), _u32(dst)
        dels = add == 0
        res = {"pos": [], "neg": []}
        if tuples:
            res["pos_tuples"], res["neg_tuples"] = [], []
        for q in queries:
            cand = graph.candidates(q)
            res["neg"].append(int(graph.count_records(q, src[dels], dst[dels], cand).sum()))
            if tuples:
                res["neg_tuples"].append(graph.tuples_records(q, src[dels], dst[dels], cand))
        graph.apply(add, src, dst, lab, relabel)
        ins = ~dels
        for q in queries:
            cand = graph.candidates(q)
            res["pos"].append(int(graph.count_records(q, src[ins], dst[ins], cand).sum()))
            if tuples:
                res["pos_tuples"].append(graph.tuples_records(q, src[ins], dst[ins], cand))
        out.append(res)
    return out
