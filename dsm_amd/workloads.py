"""Seeded synthetic workloads for the continuous subgraph matching engine.

A workload is an initial data graph (vertex labels + undirected edges), a
stream of update batches in the DSM change format (``a|d src dst label``) and a
set of standing queries. Everything is derived from the parameters and seeds
below by libdsm_gen.so (csrc/gen.cpp), so the same call yields the same
workload on any machine.

Shapes follow the workload definitions of the reference's benchmark plan:
  small   1,000 V / 8,000 uniform edges, 4 uniform labels, triangle+pendant
          query, 100 single-edge insertions of held-out edges then 100
          single-edge deletions.
  rmat1m  R-MAT (a,b,c,d) = (.57,.19,.19,.05), 1,000,000 V, 10,000,000
          undirected edges, 16 Zipf(1.0) labels; 10% of edges held out as
          inserts plus 10% of the initial edges as deletes, interleaved, in
          batches of 4,096; one 6-vertex query (5-cycle plus chord).
  rmat1m_q64  the same graph and stream with 64 standing queries requested
          as 4-12-vertex trees, cycles, 4-cliques and chorded cycles; the
          committed set holds 4-7 vertices (21 of 4, 10 of 5, 32 of 6, 1 of 7):
          every 8-12-vertex draw exceeded the enumeration budget (see
          draw_queries and queries/rmat1m_q64.json, rejected_sizes).
  rmat16m R-MAT (.57,.19,.19,.05), 16,000,000 V, 256,000,000 undirected
          edges (the parallel generator), 32 Zipf(0.8) labels; the same stream
          construction in batches of 32,768; dense queries: 4-clique, 5-cycle
          with two chords, and a tree requested at 7 vertices that fits the
          budget only at 5 (queries/rmat16m.json, rejected_sizes).

Label field of an insert record: the ORIGINAL label of its source vertex. Under
the reference update rule (cpuGraph.cpp:564) an insert relabels both endpoints
with it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEN = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def _lib():
    global _GEN
    if _GEN is not None:
        return _GEN
    path = os.path.join(_HERE, "libdsm_gen.so")
    if not os.path.exists(path):
        raise RuntimeError("libdsm_gen.so is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    L.gen_edges_uniform.restype = C.c_int64
    L.gen_edges_uniform.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, u32p, u32p]
    L.gen_edges_rmat.restype = C.c_int64
    L.gen_edges_rmat.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                 C.c_uint64, u32p, u32p]
    L.gen_edges_rmat_chunked.restype = C.c_int64
    L.gen_edges_rmat_chunked.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                         C.c_uint64, C.c_uint32, u32p, u32p]
    L.gen_labels.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, u32p]
    L.gen_set_progress.argtypes = [C.c_int]
    # graphs of more than 2^25 edges report generation progress on stderr
    L.gen_set_progress(1)
    L.gen_permutation.argtypes = [C.c_uint64, C.c_uint64, u64p]
    L.gen_csr.argtypes = [C.c_uint32, C.c_uint64, u32p, u32p, u64p, u32p]
    L.gen_find_pattern.restype = C.c_int
    L.gen_find_pattern.argtypes = [C.c_uint32, u64p, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p,
                                   u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, u32p]
    L.gen_pilot_nodes.restype = C.c_uint64
    L.gen_pilot_nodes.argtypes = [C.c_uint32, u64p, u32p, u32p, C.c_uint32, C.c_uint32, u32p, u32p, u32p,
                                  C.c_uint64, u32p, u32p, C.c_uint64]
    L.gen_write_dsm.restype = C.c_int
    L.gen_write_dsm.argtypes = [C.c_char_p, C.c_uint32, u32p, C.c_uint64, u32p, u32p, C.c_uint32,
                                u32p, u8p, u32p, u32p, u32p]
    _GEN = L
    return L


class Query:
    def __init__(self, n, labels, edges, shape=""):
        self.cost = None
        self.draw = None
        self.n = int(n)
        self.labels = np.asarray(labels, np.uint32)
        self.edges = [(int(a), int(b)) for a, b in edges]
        self.shape = shape

    @property
    def eu(self):
        return np.array([a for a, _ in self.edges], np.uint32)

    @property
    def ev(self):
        return np.array([b for _, b in self.edges], np.uint32)


class Workload:
    """Initial graph + batches + queries. Batches are flat arrays with
    ``bstart`` offsets: batch b is records bstart[b]:bstart[b+1]."""

    def __init__(self, name, n, labels, eu, ev, bstart, add, src, dst, clab, queries):
        self.name = name
        self.n = int(n)
        self.labels = labels
        self.eu, self.ev = eu, ev
        self.bstart = bstart
        self.add, self.src, self.dst, self.clab = add, src, dst, clab
        self.queries = queries

    @property
    def num_batches(self):
        return len(self.bstart) - 1

    def batch(self, b):
        s, e = int(self.bstart[b]), int(self.bstart[b + 1])
        return self.add[s:e], self.src[s:e], self.dst[s:e], self.clab[s:e]

    def batches(self, lo=0, hi=None):
        hi = self.num_batches if hi is None else min(hi, self.num_batches)
        return [self.batch(b) for b in range(lo, hi)]

    def csr(self):
        L = _lib()
        off = np.zeros(self.n + 1, np.uint64)
        nbr = np.zeros(2 * len(self.eu), np.uint32)
        L.gen_csr(self.n, len(self.eu), self.eu, self.ev, off, nbr)
        return off, nbr

    def write(self, data_path, query_paths):
        L = _lib()
        rc = L.gen_write_dsm(data_path.encode(), self.n, self.labels, len(self.eu), self.eu, self.ev,
                             self.num_batches, self.bstart.astype(np.uint32), self.add, self.src,
                             self.dst, self.clab)
        if rc:
            raise IOError(data_path)
        for q, p in zip(self.queries, query_paths):
            write_query(p, q)


def write_query(path, q):
    L = _lib()
    z32 = np.zeros(1, np.uint32)
    rc = L.gen_write_dsm(path.encode(), q.n, q.labels, len(q.edges), q.eu, q.ev, 0xFFFFFFFF,
                         z32, np.zeros(1, np.uint8), z32, z32, z32)
    if rc:
        raise IOError(path)


def permutation(n, seed):
    out = np.zeros(n, np.uint64)
    _lib().gen_permutation(n, seed, out)
    return out


# ----------------------------------------------------------------- shapes

def shape_edges(kind, k, seed=0):
    """Edge list of a query shape on k vertices."""
    if kind == "tri_pendant":
        return [(0, 1), (1, 2), (0, 2), (1, 3)]
    if kind == "cycle5_chord":
        # 6 vertices: a 5-cycle with the chord 0-2, plus vertex 5 pendant on 3
        return [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2), (3, 5)]
    if kind == "cycle5_2chords":
        return [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2), (0, 3)]
    if kind == "clique4":
        return [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    if kind == "cycle":
        return [(i, (i + 1) % k) for i in range(k)]
    if kind == "chord_cycle":
        return [(i, (i + 1) % k) for i in range(k)] + [(0, k // 2)]
    if kind == "tree":
        # random recursive tree: vertex i attaches to a seeded earlier vertex
        p = permutation(1 << 16, seed)
        return [(int(p[i] % i), i) for i in range(1, k)]
    raise ValueError(kind)


def extract_query(n, off, nbr, labels, kind, k, seed, max_start_deg=1 << 30, max_walk_deg=1 << 30,
                  budget=200000, attempts=2000, min_label=0):
    """Query of the given shape whose vertices are a randomized-walk embedding
    in the data graph (through vertices with label >= min_label); labels copied
    from the data vertices it landed on."""
    edges = shape_edges(kind, k, seed)
    pu = np.array([a for a, _ in edges], np.uint32)
    pv = np.array([b for _, b in edges], np.uint32)
    out = np.zeros(k, np.uint32)
    rc = _lib().gen_find_pattern(n, off, nbr, np.ascontiguousarray(labels, np.uint32), min_label, k,
                                 len(edges), pu, pv, seed, max_start_deg, max_walk_deg, budget, attempts,
                                 out)
    if rc != 0:
        return None
    return Query(k, labels[out], edges, shape="%s%d" % (kind, k))


def pilot_nodes(n, off, nbr, labels, q, rs, rd, budget):
    """Backtracking-tree nodes of q over the records (rs, rd) in the graph
    (off, nbr), stopping at budget + 1 (gen.cpp gen_pilot_nodes)."""
    return int(_lib().gen_pilot_nodes(n, off, nbr, np.ascontiguousarray(labels, np.uint32), q.n, len(q.edges),
                                      q.eu, q.ev, q.labels, len(rs), np.ascontiguousarray(rs, np.uint32),
                                      np.ascontiguousarray(rd, np.uint32), budget))


# -------------------------------------------------------------- workloads

def _stream(m, n_ins, n_del, batch, seed, interleave):
    """Indices of held-out inserts (into the full edge list) and deletes (into
    the initial edges), cut into batches. Returns (keep_mask, order list of
    (is_add, edge index))."""
    perm = permutation(m, seed)
    ins = perm[:n_ins]
    keep = np.ones(m, bool)
    keep[ins] = False
    init_idx = np.nonzero(keep)[0]
    dperm = permutation(len(init_idx), seed + 1)
    dels = init_idx[dperm[:n_del]]
    recs = np.concatenate([np.stack([np.ones(n_ins, np.uint64), ins]),
                           np.stack([np.zeros(n_del, np.uint64), dels])], axis=1)
    if interleave:
        mix = permutation(recs.shape[1], seed + 2)
        recs = recs[:, mix]
    return keep, recs


def _assemble(name, n, labels, eu, ev, keep, recs, batch, queries, seed):
    add = recs[0].astype(np.uint8)
    idx = recs[1].astype(np.int64)
    # orientation of each record: seeded coin so src/dst are not always u<v
    flip = (permutation(len(idx), seed + 3) & 1).astype(bool)
    src = np.where(flip, ev[idx], eu[idx]).astype(np.uint32)
    dst = np.where(flip, eu[idx], ev[idx]).astype(np.uint32)
    clab = labels[src].astype(np.uint32)
    nb = (len(idx) + batch - 1) // batch
    bstart = np.minimum(np.arange(nb + 1, dtype=np.int64) * batch, len(idx)).astype(np.uint32)
    return Workload(name, n, labels, eu[keep].copy(), ev[keep].copy(), bstart, add, src, dst, clab,
                    queries)


def small(seed=1):
    """1,000 V / 8,000 uniform edges, 4 uniform labels, triangle + pendant."""
    L = _lib()
    n, m = 1000, 8000
    eu = np.zeros(m, np.uint32)
    ev = np.zeros(m, np.uint32)
    assert L.gen_edges_uniform(n, m, seed, eu, ev) == m
    labels = np.zeros(n, np.uint32)
    L.gen_labels(n, 4, 0.0, seed + 10, labels)
    keep, recs = _stream(m, 100, 100, 1, seed + 20, interleave=False)
    off = np.zeros(n + 1, np.uint64)
    nbr = np.zeros(2 * int(keep.sum()), np.uint32)
    L.gen_csr(n, int(keep.sum()), eu[keep].copy(), ev[keep].copy(), off, nbr)
    q = extract_query(n, off, nbr, labels, "tri_pendant", 4, seed + 30, min_label=0)
    assert q is not None
    return _assemble("small", n, labels, eu, ev, keep, recs, 1, [q], seed + 40)


def dense(seed=3):
    """400 V / 8,000 uniform edges (mean degree 40), 3 uniform labels,
    triangle + pendant; 120 inserts and 120 deletes interleaved, 8 records per
    batch, so that most batches carry matches of both signs."""
    L = _lib()
    n, m = 400, 8000
    eu = np.zeros(m, np.uint32)
    ev = np.zeros(m, np.uint32)
    assert L.gen_edges_uniform(n, m, seed, eu, ev) == m
    labels = np.zeros(n, np.uint32)
    L.gen_labels(n, 3, 0.0, seed + 10, labels)
    keep, recs = _stream(m, 120, 120, 8, seed + 20, interleave=True)
    off = np.zeros(n + 1, np.uint64)
    nbr = np.zeros(2 * int(keep.sum()), np.uint32)
    L.gen_csr(n, int(keep.sum()), eu[keep].copy(), ev[keep].copy(), off, nbr)
    q = extract_query(n, off, nbr, labels, "tri_pendant", 4, seed + 30, min_label=0)
    assert q is not None
    return _assemble("dense", n, labels, eu, ev, keep, recs, 8, [q], seed + 40)


_RMAT_CACHE = {}


def _rmat_graph(n, m, nl, s, seed, chunked=False):
    key = (n, m, nl, s, seed, chunked)
    if key not in _RMAT_CACHE:
        L = _lib()
        eu = np.zeros(m, np.uint32)
        ev = np.zeros(m, np.uint32)
        if chunked:
            threads = max(1, min(16, os.cpu_count() or 1))
            rc = L.gen_edges_rmat_chunked(n, m, 0.57, 0.19, 0.19, seed, threads, eu, ev)
        else:
            rc = L.gen_edges_rmat(n, m, 0.57, 0.19, 0.19, seed, eu, ev)
        if rc != m:
            raise RuntimeError("R-MAT generation failed: %d" % rc)
        labels = np.zeros(n, np.uint32)
        L.gen_labels(n, nl, s, seed + 10, labels)
        _RMAT_CACHE.clear()
        _RMAT_CACHE[key] = (eu, ev, labels)
    return _RMAT_CACHE[key]


MIN_SIZE = {"tree": 3, "cycle": 3, "chord_cycle": 4, "clique4": 4, "cycle5_chord": 6, "tri_pendant": 4,
            "cycle5_2chords": 5}


def _pick_query(n, off, nbr, labels, nl, kind, k, seed0, rs, rd, budget, max_tries, walk=(256, 1024, 5000, 50),
                rejected=None):
    """Draws until the pilot fits the budget; every 8 redraws the walk moves
    to rarer labels; after max_tries draws the shape shrinks by one vertex.
    walk = (max start degree, max walk degree, walk budget, attempts).
    rejected (a list) receives one summary per size given up: draws made,
    draws whose walk found no embedding, and the sorted pilots of the others
    (budget + 1 = over the budget)."""
    while k >= MIN_SIZE.get(kind, 3):
        pilots, missed = [], 0
        for t in range(max_tries):
            q = extract_query(n, off, nbr, labels, kind, k, seed0 + k * 4096 + t, max_start_deg=walk[0],
                              max_walk_deg=walk[1], budget=walk[2], attempts=walk[3],
                              min_label=min(nl - 1, t // 8))
            if q is None:
                missed += 1
                continue
            q.cost = pilot_nodes(n, off, nbr, labels, q, rs, rd, budget)
            if q.cost <= budget:
                q.draw = t
                return q
            pilots.append(q.cost)
        if rejected is not None:
            rejected.append({"shape": kind, "vertices": k, "draws": max_tries, "no_embedding": missed,
                             "over_budget": len(pilots), "min_pilot": min(pilots) if pilots else None})
        k -= 1
    raise RuntimeError("no query of shape %s within the budget" % kind)


def draw_queries(w, specs, seed, nl, budget, window=12, max_tries=160, threads=8, log=None, walk=None,
                 fallback=True, rejected=None):
    """Draws one query per (shape, size) spec by randomized walk on the initial
    graph of w. A draw whose backtracking tree over the records of the first
    `window` batches (pilot on the initial graph, pilot_nodes) has more than
    `budget` nodes is redrawn; every 8 redraws the walk moves to rarer labels,
    and after max_tries draws the shape shrinks by one vertex.
    This keeps the stream enumerable: R-MAT hubs give some edges astronomically
    many embeddings of trees and long paths. The pilot visits the records with
    the highest endpoint degrees first, so explosive draws are rejected early."""
    from concurrent.futures import ThreadPoolExecutor
    off, nbr = w.csr()
    hi = int(w.bstart[min(window, w.num_batches)])
    deg = np.diff(off).astype(np.int64)
    rs, rd = w.src[:hi], w.dst[:hi]
    order = np.argsort(-(deg[rs] + deg[rd]), kind="stable")
    rs, rd = np.ascontiguousarray(rs[order]), np.ascontiguousarray(rd[order])

    rej = [[] for _ in specs]

    def one(i):
        kind, k = specs[i]
        q = None
        # a shape with no draw within the budget falls back to a cycle, then a
        # tree, of the same size (recorded in q.shape)
        for alt in ((kind, "cycle", "tree") if fallback else (kind,)):
            try:
                q = _pick_query(w.n, off, nbr, w.labels, nl, alt, k, seed * 1000 + i * 1024 + 30, rs, rd,
                                budget, max_tries, rejected=rej[i], **({"walk": walk} if walk else {}))
                break
            except RuntimeError:
                continue
        if q is None:
            raise RuntimeError("no query for spec %d (%s%d) within the budget" % (i, kind, k))
        if log:
            log("query %d: %s labels %s pilot %d nodes (draw %d)" % (i, q.shape, q.labels.tolist(), q.cost,
                                                                     q.draw))
        return q

    with ThreadPoolExecutor(max_workers=threads) as ex:
        out = list(ex.map(one, range(len(specs))))
    if rejected is not None:
        for i, r in enumerate(rej):
            for x in r:
                x["spec"] = i
                rejected.append(x)
    return out


QUERY_DIR = os.path.join(_HERE, "queries")


def load_queries(name):
    import json
    with open(os.path.join(QUERY_DIR, name + ".json")) as f:
        d = json.load(f)
    qs = []
    for x in d["queries"]:
        q = Query(len(x["labels"]), x["labels"], x["edges"], shape=x["shape"])
        q.cost = x.get("pilot_nodes")
        qs.append(q)
    return qs


def rmat(n=1000000, m=10000000, nl=16, zipf=1.0, batch=4096, seed=2, frac=0.10, name="rmat1m",
         queries=None, chunked=False):
    """R-MAT graph with a held-out insert stream and random deletes; the
    standing queries are the committed set `queries/<name>.json` (drawn by
    draw_queries, see queries/draw.py) unless a list is given. chunked: the
    parallel generator (gen_edges_rmat_chunked; a different, equally seeded
    edge sequence)."""
    eu, ev, labels = _rmat_graph(n, m, nl, zipf, seed, chunked)
    n_ins = int(m * frac)
    keep, recs = _stream(m, n_ins, n_ins, batch, seed + 20, interleave=True)
    w = _assemble(name, n, labels, eu, ev, keep, recs, batch, [], seed + 40)
    w.queries = list(queries) if queries is not None else load_queries(name)
    return w


def q64_specs(seed=7):
    """64 requested query shapes of 4-12 vertices: trees, cycles, 4-cliques,
    chorded cycles (draw_queries shrinks those with no draw within the budget)."""
    p = permutation(4096, seed)
    kinds = ["tree", "cycle", "clique4", "chord_cycle"]
    specs = []
    for i in range(64):
        kind = kinds[i % 4]
        if kind == "clique4":
            k = 4
        else:
            k = 4 + int(p[i] % 9)  # 4..12
        specs.append((kind, k))
    return specs


def rmat_q64(**kw):
    """The same graph and stream with the 64 committed standing queries."""
    kw.setdefault("name", "rmat1m_q64")
    return rmat(**kw)


RMAT16M_SPECS = [("clique4", 4), ("cycle5_2chords", 5), ("tree", 7)]


def rmat16m(**kw):
    """R-MAT 16,000,000 V / 256,000,000 undirected edges, 32 Zipf(0.8) labels,
    batches of 32,768, with the committed dense queries (4-clique, 5-cycle with
    two chords, 7-vertex tree)."""
    kw.setdefault("name", "rmat16m")
    return rmat(n=16_000_000, m=256_000_000, nl=32, zipf=0.8, batch=32768, seed=5, chunked=True, **kw)
