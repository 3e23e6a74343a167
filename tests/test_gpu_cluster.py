"""Parity of the device cluster-encoding refinement (dsm_options.cluster_filter)
with the oracle's restatement of clusterFilter (filter.cu:1525-1654):
bit-exact refined candidate sets of every query vertex after every batch, and
counts that equal the oracle's (the refinement only prunes candidates no
embedding uses)."""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from dsm_amd import engine as E
from dsm_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STAR3 = [(0, 1), (0, 2), (0, 3), (1, 4), (2, 5), (3, 6)]
SHAPES = {
    "edge": [(0, 1)],
    "path3": [(0, 1), (1, 2)],
    "triangle": [(0, 1), (1, 2), (0, 2)],
    "tri_pendant": [(0, 1), (1, 2), (0, 2), (1, 3)],
    "square": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "clique4": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],
    "house": [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4)],
    "path6": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)],
    "star3legs": STAR3,
    "combine_keep": [(0, 1), (4, 2), (3, 4), (3, 1), (1, 2), (0, 5), (0, 3)],
    "combine_late": [(1, 2), (0, 5), (0, 1), (1, 3), (2, 4), (0, 6), (5, 6)],
    "tree7": [(0, 1), (1, 2), (1, 3), (3, 4), (0, 5), (5, 6)],
}


def say(msg):
    print(msg, flush=True)


def qtuple(edges, labels):
    e = np.array(edges, np.uint32)
    return (len(labels), np.asarray(labels, np.uint32), e[:, 0].copy(), e[:, 1].copy())


def check_refined(eng, g, queries, oq, tag):
    for qi, q in enumerate(oq):
        want = g.cluster_candidates(q)
        for u in range(q.n):
            got = eng.refined_candidates(qi, u)
            assert np.array_equal(got, want[u].astype(bool)), "%s query %d vertex %d refined candidates" % (tag, qi, u)


def test_example_fixture_refined(golden_dir):
    say("case: reference example fixture with the cluster filter")
    d = E.DsmFile(os.path.join(golden_dir, "example_g_01.in"), False)
    q = E.DsmFile(os.path.join(golden_dir, "example_q_01.in"), True)
    exp = json.load(open(os.path.join(golden_dir, "example_expected.json")))
    eng = E.Engine(d.graph(), [q.graph()], cluster_filter=True)
    try:
        pos, neg = eng.process_batch(*d.batch(0))
        assert pos.tolist() == [exp["batch0_pos"]]
        got = {u: np.nonzero(eng.refined_candidates(0, u))[0].tolist() for u in range(5)}
        # hand-derived (tests/test_cluster_oracle.py): vertex 1 leaves u0 and u2
        assert got == {0: [0, 4, 5, 6, 7, 8], 1: [1, 6, 8], 2: [0, 4, 5, 6, 7, 8], 3: [2, 9], 4: [3, 10]}
        st = eng.stats()
        say("  ms_refine %.3f" % st["ms_refine"])
    finally:
        eng.close()


def random_graph(n, m, nl, rng):
    edges = set()
    while len(edges) < m:
        a, b = map(int, rng.integers(0, n, 2))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    return rng.integers(0, nl, n).astype(np.uint32), e[:, 0].copy(), e[:, 1].copy()


def random_batch(n, k, rng, nl):
    add = (rng.random(k) < 0.6).astype(np.uint8)
    src = rng.integers(0, n, k).astype(np.uint32)
    dst = rng.integers(0, n, k).astype(np.uint32)
    return add, src, dst, rng.integers(0, nl, k).astype(np.uint32)


@pytest.mark.parametrize("relabel", [True, False])
def test_random_stream_refined_and_counts(relabel):
    say("case: random stream, 12 shapes, relabel=%s" % relabel)
    rng = np.random.default_rng(31 if relabel else 32)
    n, nl = 400, 3
    lab, eu, ev = random_graph(n, 2400, nl, rng)
    queries = []
    for name in sorted(SHAPES):
        ed = SHAPES[name]
        k = max(max(e) for e in ed) + 1
        queries.append(qtuple(ed, rng.integers(0, nl, k)))
    g = O.Graph(n, lab, eu, ev)
    oq = [O.Query(*q) for q in queries]
    eng = E.Engine((n, lab, eu, ev), queries, relabel=relabel, cluster_filter=True)
    try:
        check_refined(eng, g, queries, oq, "initial")
        for bi in range(4):
            b = random_batch(n, 60, rng, nl)
            want = O.run_stream(g, oq, [b], relabel=relabel)[0]
            pos, neg = eng.process_batch(*b)
            assert pos.tolist() == want["pos"], "batch %d positives" % bi
            assert neg.tolist() == want["neg"], "batch %d negatives" % bi
            check_refined(eng, g, queries, oq, "batch %d" % bi)
    finally:
        eng.close()


def test_pipelined_with_refinement_equals_oracle():
    say("case: pipelined batches with the cluster filter (refinement on the second copy)")
    rng = np.random.default_rng(77)
    n, nl = 600, 3
    lab, eu, ev = random_graph(n, 4000, nl, rng)
    queries = []
    for name in ("star3legs", "combine_keep", "combine_late", "house", "tree7", "square"):
        ed = SHAPES[name]
        k = max(max(e) for e in ed) + 1
        queries.append(qtuple(ed, rng.integers(0, nl, k)))
    batches = [random_batch(n, int(rng.integers(1, 200)), rng, nl) for _ in range(6)]
    g = O.Graph(n, lab, eu, ev)
    oq = [O.Query(*q) for q in queries]
    want = O.run_stream(g, oq, batches)
    eng = E.Engine((n, lab, eu, ev), queries, cluster_filter=True)
    try:
        pos, neg, _ = eng.process_batches(batches)
        for bi in range(len(batches)):
            assert pos[bi].tolist() == want[bi]["pos"], "batch %d positives" % bi
            assert neg[bi].tolist() == want[bi]["neg"], "batch %d negatives" % bi
        check_refined(eng, g, queries, oq, "after the stream")
    finally:
        eng.close()


def test_hub_candidates_use_heavy_chunks():
    say("case: hub candidates with more than 4,096 neighbours (chunked neighbour walks)")
    rng = np.random.default_rng(5)
    n = 20000
    hubs = [0, 1, 2]
    edges = set()
    for h in hubs:
        for v in rng.choice(np.arange(3, n), 9000, replace=False):
            edges.add((h, int(v)))
    while len(edges) < 60000:
        a, b = map(int, rng.integers(3, n, 2))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    lab = rng.integers(0, 2, n).astype(np.uint32)
    lab[hubs] = 0
    queries = [qtuple(SHAPES["star3legs"], [0, 1, 1, 0, 1, 0, 1]), qtuple(SHAPES["path3"], [1, 0, 1]),
               qtuple(SHAPES["combine_keep"], [0, 1, 0, 1, 0, 1])]
    g = O.Graph(n, lab, e[:, 0], e[:, 1])
    oq = [O.Query(*q) for q in queries]
    # refined sets only: around hubs of degree 9,000 these shapes have far too
    # many embeddings for the oracle to enumerate, so nothing is matched here
    eng = E.Engine((n, lab, e[:, 0], e[:, 1]), queries, cluster_filter=True, count_negatives=False)
    try:
        check_refined(eng, g, queries, oq, "initial")
        for bi in range(2):
            b = random_batch(n, 300, rng, 2)
            g.apply(*b)
            eng.apply_batch(*b)
            check_refined(eng, g, queries, oq, "batch %d" % bi)
    finally:
        eng.close()


@pytest.mark.timeout(900)
def test_rmat1m_q64_refined_sets_and_counts():
    say("case: rmat1m_q64 with the cluster filter: refined sets of every query vertex, counts, 2 batches")
    w = W.rmat_q64()
    qs = [(q.n, q.labels, q.eu, q.ev) for q in w.queries]
    batches = w.batches(0, 2)
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    oq = [O.Query(*q) for q in qs]
    threads = max(1, min(16, os.cpu_count() or 1))
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qs, cluster_filter=True)
    try:
        for bi, b in enumerate(batches):
            want = O.run_stream_threaded(g, oq, [b], threads=threads, log=say)[0]
            pos, neg = eng.process_batch(*b)
            assert pos.tolist() == want["pos"], "batch %d positives" % bi
            assert neg.tolist() == want["neg"], "batch %d negatives" % bi
            say("  batch %d: counts equal; ms_refine %.2f, dfs %.2f ms" % (bi, eng.stats()["ms_refine"],
                                                                          eng.stats()["ms_dfs_pos"]))
        with ThreadPoolExecutor(max_workers=threads) as ex:
            refined = list(ex.map(g.cluster_candidates, oq))
        for qi, q in enumerate(oq):
            for u in range(q.n):
                assert np.array_equal(eng.refined_candidates(qi, u), refined[qi][u].astype(bool)), (qi, u)
        say("  refined sets equal for all %d query vertices" % sum(q.n for q in oq))
    finally:
        eng.close()
