"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Bar: bit-exact. Per batch and query the uint64 positive and negative counts,
the sorted tuple multisets, the device adjacency after the update rule and the
LDF/NLF candidate bitmaps must equal the oracle's on the same seeded inputs.
"""
import json
import os

import numpy as np
import pytest

from dsm_amd import engine as E
from dsm_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = {
    "edge": [(0, 1)],
    "path3": [(0, 1), (1, 2)],
    "triangle": [(0, 1), (1, 2), (0, 2)],
    "tri_pendant": [(0, 1), (1, 2), (0, 2), (1, 3)],
    "square": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "star4": [(0, 1), (0, 2), (0, 3), (0, 4)],
    "clique4": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],
    "house": [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4)],
    "path6": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)],
    "cycle6_chord": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (0, 3)],
}


def say(msg):
    print(msg, flush=True)


def qtuple(shape, labels):
    e = np.array(SHAPES[shape], np.uint32)
    return (len(labels), np.asarray(labels, np.uint32), e[:, 0].copy(), e[:, 1].copy())


def random_graph(n, m, nl, rng):
    edges = set()
    while len(edges) < m:
        a, b = rng.integers(0, n, 2)
        if a != b:
            edges.add((int(min(a, b)), int(max(a, b))))
    e = np.array(sorted(edges), np.uint32)
    return rng.integers(0, nl, n).astype(np.uint32), e[:, 0].copy(), e[:, 1].copy()


def conflict_free_batch(n, k, present, rng, p_add=0.5):
    """k records, no undirected pair touched twice: inserts of absent pairs,
    deletes of present ones; updates `present`."""
    add, src, dst = [], [], []
    used = set()
    pl = sorted(present)
    while len(src) < k:
        if rng.random() < p_add or not pl:
            a, b = map(int, rng.integers(0, n, 2))
            key = (min(a, b), max(a, b))
            if a == b or key in present or key in used:
                continue
            used.add(key)
            present.add(key)
            add.append(1)
        else:
            key = pl[int(rng.integers(0, len(pl)))]
            if key in used or key not in present:
                continue
            used.add(key)
            present.discard(key)
            a, b = key if rng.random() < 0.5 else key[::-1]
            add.append(0)
        src.append(a)
        dst.append(b)
    return (np.array(add, np.uint8), np.array(src, np.uint32), np.array(dst, np.uint32),
            rng.integers(0, 4, k).astype(np.uint32))


def oracle_queries(queries):
    return [O.Query(n, l, u, v) for n, l, u, v in queries]


def compare_stream(n, labels, eu, ev, queries, batches, relabel=True, tuples=False, check_graph=False,
                   **engine_kw):
    g = O.Graph(n, labels, eu, ev)
    oq = oracle_queries(queries)
    eng = E.Engine((n, labels, eu, ev), queries, relabel=relabel,
                   tuple_capacity=(1 << 20) if tuples else 0, **engine_kw)
    used = set(int(x) for q in queries for x in np.asarray(q[1]))
    try:
        for bi, (add, src, dst, lab) in enumerate(batches):
            want = O.run_stream(g, oq, [(add, src, dst, lab)], relabel=relabel, tuples=tuples)[0]
            say("  batch %d: %d records, oracle done" % (bi, len(src)))
            pos, neg = eng.process_batch(add, src, dst, lab)
            say("  batch %d: engine done %s" % (bi, {k: v for k, v in eng.stats().items() if k.startswith("ms_")}))
            assert pos.tolist() == want["pos"], "batch %d positive counts" % bi
            assert neg.tolist() == want["neg"], "batch %d negative counts" % bi
            if tuples:
                got = eng.tuples()
                for qi in range(len(queries)):
                    for sign, key in ((1, "pos_tuples"), (0, "neg_tuples")):
                        a = sorted(map(tuple, got[(qi, sign)].tolist()))
                        b = sorted(map(tuple, want[key][qi].tolist()))
                        assert a == b, "batch %d query %d sign %d tuples" % (bi, qi, sign)
            if check_graph:
                off, nb, dlab = eng.graph()
                ooff, onb, olab = g.export()
                assert off.tolist() == ooff.tolist(), "batch %d offsets" % bi
                assert np.array_equal(nb, onb), "batch %d neighbours" % bi
                clamp = np.array([l if (l < 255 and l in used) else 255 for l in olab.tolist()], np.uint32)
                assert np.array_equal(dlab, clamp), "batch %d labels" % bi
                for qi, q in enumerate(oq):
                    cand = g.candidates(q)
                    for u in range(q.n):
                        assert np.array_equal(eng.candidates(qi, u), cand[u].astype(bool)), \
                            "batch %d query %d vertex %d candidates" % (bi, qi, u)
        return eng.stats()
    finally:
        eng.close()


# ----------------------------------------------------------------- fixtures

def test_example_fixture(golden_dir):
    say("case: reference example fixture")
    exp = json.load(open(os.path.join(golden_dir, "example_expected.json")))
    d = E.DsmFile(os.path.join(golden_dir, "example_g_01.in"), False)
    q = E.DsmFile(os.path.join(golden_dir, "example_q_01.in"), True)
    eng = E.Engine(d.graph(), [q.graph()], tuple_capacity=1024)
    try:
        pos, neg = eng.process_batch(*d.batch(0))
        assert pos.tolist() == [exp["batch0_pos"]] and neg.tolist() == [exp["batch0_neg"]]
        t = eng.tuples()
        assert sorted(t[(0, 1)].tolist()) == exp["batch0_tuples"]
        for u, vs in exp["candidates_after_batch0"].items():
            assert np.nonzero(eng.candidates(0, int(u)))[0].tolist() == vs
    finally:
        eng.close()


def test_small_workload_golden(golden_dir):
    """1,000 V / 8,000 E, triangle+pendant, 100 inserts then 100 deletes."""
    say("case: small workload golden fixture (200 single-record batches)")
    exp = json.load(open(os.path.join(golden_dir, "small_expected.json")))
    d = E.DsmFile(os.path.join(golden_dir, "small_g.in"), False)
    q = E.DsmFile(os.path.join(golden_dir, "small_q.in"), True)
    eng = E.Engine(d.graph(), [q.graph()], tuple_capacity=4096)
    try:
        for b, want in enumerate(exp["batches"]):
            pos, neg = eng.process_batch(*d.batch(b))
            assert (int(pos[0]), int(neg[0])) == (want["pos"], want["neg"]), "batch %d" % b
            t = eng.tuples()
            assert sorted(t[(0, 1)].tolist()) == want["pos_tuples"], "batch %d pos tuples" % b
            assert sorted(t[(0, 0)].tolist()) == want["neg_tuples"], "batch %d neg tuples" % b
        off, nb, lab = eng.graph()
        fin = exp["final_graph"]
        assert len(nb) == fin["entries"]
        assert int((nb.astype(np.uint64) * 2654435761 % (1 << 32)).sum() % (1 << 61)) == fin["checksum"]
    finally:
        eng.close()


def test_dense_workload_golden(golden_dir):
    """400 V, mean degree 40, triangle+pendant, 8 interleaved inserts and
    deletes per batch (23 of 30 batches carry positive matches, 16 negative
    ones): counts and a digest of each batch's tuples against the committed
    fixture."""
    import sys
    sys.path.insert(0, golden_dir)
    import make_fixtures as MF
    say("case: dense workload golden fixture (30 batches of 8 records)")
    exp = json.load(open(os.path.join(golden_dir, "dense_expected.json")))
    w = W.dense()
    q = w.queries[0]
    assert q.labels.tolist() == exp["query"]["labels"]
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), [(q.n, q.labels, q.eu, q.ev)], tuple_capacity=1 << 16)
    try:
        for b, want in enumerate(exp["batches"]):
            pos, neg = eng.process_batch(*w.batch(b))
            assert (int(pos[0]), int(neg[0])) == (want["pos"], want["neg"]), "batch %d" % b
            t = eng.tuples()
            assert MF.tuple_digest(t[(0, 1)]) == want["pos_digest"], "batch %d pos tuples" % b
            assert MF.tuple_digest(t[(0, 0)]) == want["neg_digest"], "batch %d neg tuples" % b
    finally:
        eng.close()


# ------------------------------------------------------------ random streams

@pytest.mark.parametrize("relabel", [True, False])
def test_random_multi_query_stream(relabel):
    say("case: random multi-query stream relabel=%s" % relabel)
    rng = np.random.default_rng(101 if relabel else 202)
    n, m, nl = 300, 2400, 3
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple(s, rng.integers(0, nl, len(set(x for e in SHAPES[s] for x in e))))
               for s in sorted(SHAPES)]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 40, present, rng) for _ in range(8)]
    compare_stream(n, labels, eu, ev, queries, batches, relabel=relabel, tuples=True, check_graph=True)


def test_dense_graph_work_donation():
    """Dense labelled graph where single seeds have large subtrees: exercises
    the donation queue (small donate_min)."""
    say("case: dense graph with work donation")
    rng = np.random.default_rng(303)
    n, m, nl = 400, 10000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("clique4", [0, 0, 1, 1]), qtuple("path6", [0, 1, 0, 1, 0, 1]),
               qtuple("house", [0, 0, 0, 1, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 64, present, rng) for _ in range(2)]
    st = compare_stream(n, labels, eu, ev, queries, batches, relabel=False, donate_min=64)
    say("  donated tasks: %d" % st["tasks_donated"])


def test_hub_vertex_many_changes_and_pool_growth():
    """One vertex gets > 1024 inserts in one batch (several apply rounds),
    then every vertex gains edges beyond its slack (relocation of every list
    into the pool's reserve)."""
    say("case: hub vertex with 3000 inserts in one batch")
    rng = np.random.default_rng(404)
    n = 5000
    labels = rng.integers(0, 2, n).astype(np.uint32)
    eu = np.arange(0, n - 1, dtype=np.uint32)
    ev = eu + 1
    k = 3000
    dst = rng.choice(np.arange(2, n), k, replace=False).astype(np.uint32)
    hub = [(np.ones(k, np.uint8), np.zeros(k, np.uint32), dst, np.zeros(k, np.uint32))]
    # then many vertices each gaining edges beyond their slack
    more_s = rng.integers(0, n, 80000).astype(np.uint32)
    more_d = rng.integers(0, n, 80000).astype(np.uint32)
    keep = more_s != more_d
    grow = [(np.ones(int(keep.sum()), np.uint8), more_s[keep], more_d[keep],
             np.zeros(int(keep.sum()), np.uint32))]
    queries = [qtuple("triangle", [0, 0, 1]), qtuple("star4", [0, 1, 1, 0, 1])]
    compare_stream(n, labels, eu, ev, queries, hub + grow, relabel=True, check_graph=True)


def test_edge_cases_batches():
    say("case: empty / delete-only / duplicate / self-loop / absent-edge batches")
    rng = np.random.default_rng(505)
    n, m = 120, 700
    labels, eu, ev = random_graph(n, m, 2, rng)
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("edge", [1, 1])]
    z = np.zeros(0, np.uint32)
    present = sorted(zip(eu.tolist(), ev.tolist()))
    a, b = present[0]
    c, d = present[1]
    batches = [
        (z.astype(np.uint8), z, z, z),                                             # empty
        (np.zeros(2, np.uint8), np.array([a, d], np.uint32), np.array([b, c], np.uint32), np.zeros(2, np.uint32)),
        (np.ones(2, np.uint8), np.array([a, a], np.uint32), np.array([b, b], np.uint32), np.zeros(2, np.uint32)),  # duplicate insert
        (np.ones(1, np.uint8), np.array([5], np.uint32), np.array([5], np.uint32), np.ones(1, np.uint32)),       # self loop
        (np.zeros(1, np.uint8), np.array([7], np.uint32), np.array([8], np.uint32), np.zeros(1, np.uint32)),     # maybe absent
    ]
    compare_stream(n, labels, eu, ev, queries, batches, relabel=True, tuples=True, check_graph=True)


def test_invalid_batch_leaves_state_unchanged():
    say("case: out-of-range change rejected")
    rng = np.random.default_rng(606)
    n, m = 100, 500
    labels, eu, ev = random_graph(n, m, 2, rng)
    queries = [qtuple("triangle", [0, 0, 1])]
    g = O.Graph(n, labels, eu, ev)
    eng = E.Engine((n, labels, eu, ev), queries)
    try:
        with pytest.raises(E.DsmError) as ei:
            eng.process_batch(np.array([1, 1], np.uint8), np.array([1, 2], np.uint32),
                              np.array([2, n], np.uint32), np.zeros(2, np.uint32))
        assert ei.value.code == E.DSM_EINVAL
        present = set(zip(eu.tolist(), ev.tolist()))
        batch = conflict_free_batch(n, 30, present, rng)
        want = O.run_stream(g, oracle_queries(queries), [batch])[0]
        pos, neg = eng.process_batch(*batch)
        assert pos.tolist() == want["pos"] and neg.tolist() == want["neg"]
    finally:
        eng.close()


def test_rmat_scaled_stream():
    """R-MAT (.57,.19,.19) 20,000 V / 200,000 E, 16 Zipf labels, random-walk
    queries of mixed shapes, batches of 1,024 mixed records."""
    say("case: scaled R-MAT stream")
    w = W.rmat(n=20000, m=200000, batch=1024, seed=9, name="rmat20k", queries=[])
    w.queries = W.draw_queries(w, [("cycle5_chord", 6), ("clique4", 4), ("tree", 5), ("cycle", 5),
                                   ("chord_cycle", 6), ("tree", 4)], seed=9, nl=16, budget=3_000_000,
                               window=3)
    queries = [(q.n, q.labels, q.eu, q.ev) for q in w.queries]
    compare_stream(w.n, w.labels, w.eu, w.ev, queries, w.batches(0, 3), relabel=True, check_graph=False)


def test_each_query_alone_equals_in_mix():
    say("case: alone == in the mix")
    rng = np.random.default_rng(707)
    n, m, nl = 500, 4000, 3
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple(s, rng.integers(0, nl, len(set(x for e in SHAPES[s] for x in e))))
               for s in ("triangle", "square", "house", "path6")]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 100, present, rng) for _ in range(3)]
    mix = E.Engine((n, labels, eu, ev), queries)
    alone = [E.Engine((n, labels, eu, ev), [q]) for q in queries]
    try:
        for b in batches:
            pm, nm = mix.process_batch(*b)
            for i, e in enumerate(alone):
                pa, na = e.process_batch(*b)
                assert (int(pa[0]), int(na[0])) == (int(pm[i]), int(nm[i]))
    finally:
        mix.close()
        for e in alone:
            e.close()


@pytest.mark.parametrize("wpb", [1, 2, 8])
def test_launch_shapes_agree(wpb):
    say("case: DFS launch shape waves_per_block=%d" % wpb)
    rng = np.random.default_rng(808)
    n, m, nl = 300, 3000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("house", [0, 1, 0, 1, 0]), qtuple("clique4", [0, 0, 1, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 60, present, rng) for _ in range(2)]
    compare_stream(n, labels, eu, ev, queries, batches, relabel=False, waves_per_block=wpb, donate_min=64)


def test_tuples_in_pinned_host_memory():
    say("case: tuple spill to pinned host memory")
    rng = np.random.default_rng(909)
    n, m, nl = 200, 1500, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 50, present, rng) for _ in range(2)]
    compare_stream(n, labels, eu, ev, queries, batches, tuples=True, tuples_in_host_memory=True)


def test_uint64_counts_beyond_32_bits_closed_form():
    """All 19,900 edges of K_200 inserted in one batch into an empty graph,
    4-clique query, one label: every ordered embedding (200*199*198*197) uses
    six inserted edges and counts once per edge, so the positive count is
    6 * 200*199*198*197 = 9,314,632,800 > 2^32 (closed form, no oracle)."""
    say("case: K_200 batch, 4-clique, count beyond 32 bits")
    n = 200
    labels = np.zeros(n, np.uint32)
    z = np.zeros(0, np.uint32)
    iu, ju = np.triu_indices(n, 1)
    k = len(iu)
    batch = (np.ones(k, np.uint8), iu.astype(np.uint32), ju.astype(np.uint32), np.zeros(k, np.uint32))
    eng = E.Engine((n, labels, z, z), [qtuple("clique4", [0, 0, 0, 0])])
    try:
        pos, neg = eng.process_batch(*batch)
        want = 6 * 200 * 199 * 198 * 197
        assert want > 2 ** 32
        assert int(pos[0]) == want and int(neg[0]) == 0
        # deleting them all again: the same count as negatives
        pos, neg = eng.process_batch(np.zeros(k, np.uint8), batch[1], batch[2], batch[3])
        assert int(neg[0]) == want and int(pos[0]) == 0
    finally:
        eng.close()


def test_autotune_keeps_counts_and_state():
    """dsm_autotune times launch configurations on a sample (every one must give
    the same counts) and keeps the fastest; the stream that follows still
    matches the oracle, graph and bitmaps included."""
    say("case: autotune, then a stream against the oracle")
    rng = np.random.default_rng(1111)
    n, m, nl = 400, 5000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("house", [0, 1, 0, 1, 0]), qtuple("clique4", [0, 0, 1, 1]), qtuple("path6", [0, 1, 0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    sample = conflict_free_batch(n, 200, set(present), rng)
    batches = [conflict_free_batch(n, 80, present, rng) for _ in range(3)]
    g = O.Graph(n, labels, eu, ev)
    oq = oracle_queries(queries)
    eng = E.Engine((n, labels, eu, ev), queries)
    try:
        rep = eng.autotune(*sample)
        say("  tuned: %s" % rep)
        assert rep["waves_per_block"] in (1, 2, 4, 8) and rep["blocks_per_cu"] >= 1
        assert rep["trials"] >= 5 and rep["ms_best"] > 0 and rep["ms_initial"] > 0
        for bi, b in enumerate(batches):
            want = O.run_stream(g, oq, [b])[0]
            pos, neg = eng.process_batch(*b)
            assert pos.tolist() == want["pos"] and neg.tolist() == want["neg"], "batch %d" % bi
        off, nb, _ = eng.graph()
        ooff, onb, _ = g.export()
        assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb)
    finally:
        eng.close()


@pytest.mark.parametrize("leaf_sets,lane_work,budget", [(False, 0, 0), (True, 1, 0), (True, 1000, 0),
                                                        (True, 0, 16), (True, 0, 0xFFFFFFFF)])
def test_leaf_level_strategies_agree(leaf_sets, lane_work, budget):
    """Leaf sets off and on, lane work limits, and leaf budgets per window
    (16: candidates descend constantly; 0xFFFFFFFF: never) against the oracle."""
    say("case: leaf level leaf_sets=%s lane_work=%d budget=%d" % (leaf_sets, lane_work, budget))
    rng = np.random.default_rng(1212)
    n, m, nl = 300, 4000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("house", [0, 1, 0, 1, 0]), qtuple("clique4", [0, 0, 1, 1]), qtuple("square", [0, 1, 0, 1]),
               qtuple("star4", [0, 1, 1, 0, 1]), qtuple("cycle6_chord", [0, 1, 0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 60, present, rng) for _ in range(2)]
    compare_stream(n, labels, eu, ev, queries, batches, relabel=True, leaf_sets=leaf_sets, lane_work=lane_work,
                   leaf_budget=budget)


@pytest.mark.parametrize("tuples", [False, True])
def test_graph_and_eager_rounds_agree(tuples):
    """The DFS round as a HIP graph (built once, node arguments updated per
    round, replayed with eager apply work in between) against separate
    launches and the oracle, batch by batch, tuples included."""
    say("case: graph rounds vs eager rounds, tuples=%s" % tuples)
    rng = np.random.default_rng(1313)
    n, m, nl = 300, 3000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("house", [0, 1, 0, 1, 0]), qtuple("path6", [0, 1, 0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, int(k), present, rng) for k in (1, 40, 3, 120, 0, 60)]
    for graph_rounds in (True, False):
        compare_stream(n, labels, eu, ev, queries, batches, relabel=True, tuples=tuples, graph_rounds=graph_rounds)


def test_label_sorted_copy_of_big_hub():
    """A hub of degree > 4,096 (its label-sorted copy is built by a whole
    block) and relabels of its neighbours every batch (the copy is rebuilt
    for the dirty set): counts, graph and bitmaps against the oracle."""
    say("case: hub of degree 9,000 with relabelled neighbours")
    rng = np.random.default_rng(1414)
    n, nl = 10000, 4
    labels = rng.integers(0, nl, n).astype(np.uint32)
    hub_nbrs = rng.choice(np.arange(1, n), 9000, replace=False)
    edges = set((0, int(v)) for v in hub_nbrs)
    while len(edges) < 9000 + 20000:
        a, b = map(int, rng.integers(1, n, 2))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    eu, ev = e[:, 0].copy(), e[:, 1].copy()
    queries = [qtuple("triangle", [int(labels[0]), 1, 2]), qtuple("square", [int(labels[0]), 0, 3, 1]),
               qtuple("house", [int(labels[0]), 1, 0, 2, 3])]
    present = set(edges)
    batches = []
    for _ in range(3):
        add, src, dst, lab = conflict_free_batch(n, 300, present, rng)
        lab = rng.integers(0, nl, len(src)).astype(np.uint32)   # relabels both endpoints of inserts
        batches.append((add, src, dst, lab))
    # inserts at the hub so that its list and its neighbours' labels change
    k = 50
    nb = rng.choice(np.setdiff1d(np.arange(1, n), hub_nbrs), k, replace=False).astype(np.uint32)
    batches.append((np.ones(k, np.uint8), np.zeros(k, np.uint32), nb, rng.integers(0, nl, k).astype(np.uint32)))
    compare_stream(n, labels, eu, ev, queries, batches, relabel=True, check_graph=True)


def test_label_sorted_copy_of_huge_hub():
    """A hub of degree 140,000 (its label-sorted copy is built in 65,536-entry
    segments spread over the grid), relabelled neighbours and inserts at the
    hub: counts, graph and bitmaps against the oracle."""
    say("case: hub of degree 140,000")
    rng = np.random.default_rng(1515)
    n, nl = 150000, 4
    labels = rng.integers(0, nl, n).astype(np.uint32)
    hub_nbrs = rng.choice(np.arange(1, n), 140000, replace=False)
    a = rng.integers(1, n, 40000)
    b = rng.integers(1, n, 40000)
    keep = a != b
    e = np.concatenate([np.stack([np.zeros(len(hub_nbrs), np.int64), hub_nbrs], 1),
                        np.stack([np.minimum(a[keep], b[keep]), np.maximum(a[keep], b[keep])], 1)])
    e = np.unique(e, axis=0).astype(np.uint32)
    eu, ev = e[:, 0].copy(), e[:, 1].copy()
    x = int(labels[0])
    # one label everywhere: inserts relabel both endpoints alike, so only
    # same-label query edges can be seeded by them
    queries = [qtuple("triangle", [x, x, x]), qtuple("square", [x, x, x, x])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = []
    for _ in range(2):
        add, src, dst, lab = conflict_free_batch(n, 400, present, rng)
        batches.append((add, src, dst, rng.integers(0, nl, len(src)).astype(np.uint32)))
    k = 100
    nb = rng.choice(np.setdiff1d(np.arange(1, n), hub_nbrs), k, replace=False).astype(np.uint32)
    batches.append((np.ones(k, np.uint8), np.zeros(k, np.uint32), nb, rng.integers(0, nl, k).astype(np.uint32)))
    compare_stream(n, labels, eu, ev, queries, batches, relabel=True, check_graph=True)


@pytest.mark.parametrize("relabel,overlap", [(True, True), (False, True), (True, False)])
def test_process_batches_pipelined_equals_one_by_one(relabel, overlap):
    """dsm_process_batches (the positives of batch b share a launch with the
    negatives of batch b+1; with overlap, batch b+1 is applied to the second
    copy of the graph meanwhile) against dsm_process_batch one batch at a time
    and the oracle, with empty, insert-only, delete-only and mixed batches."""
    say("case: pipelined batches, relabel=%s overlap=%s" % (relabel, overlap))
    rng = np.random.default_rng(1616 if relabel else 1717)
    n, m, nl = 400, 4000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("house", [0, 1, 0, 1, 0]), qtuple("path6", [0, 1, 0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batches = [conflict_free_batch(n, 60, present, rng),
               conflict_free_batch(n, 0, present, rng),
               conflict_free_batch(n, 30, present, rng, p_add=1.0),
               conflict_free_batch(n, 30, present, rng, p_add=0.0),
               conflict_free_batch(n, 90, present, rng),
               conflict_free_batch(n, 1, present, rng)]
    g = O.Graph(n, labels, eu, ev)
    want = O.run_stream(g, oracle_queries(queries), batches, relabel=relabel)
    one = E.Engine((n, labels, eu, ev), queries, relabel=relabel)
    pipe = E.Engine((n, labels, eu, ev), queries, relabel=relabel, overlap_apply=overlap)
    try:
        pos, neg, done = pipe.process_batches(batches)
        assert np.all(np.diff(done) >= 0)
        for b, bt in enumerate(batches):
            p1, n1 = one.process_batch(*bt)
            assert pos[b].tolist() == p1.tolist() == want[b]["pos"], "batch %d positives" % b
            assert neg[b].tolist() == n1.tolist() == want[b]["neg"], "batch %d negatives" % b
        off, nb, _ = pipe.graph()
        ooff, onb, _ = g.export()
        assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb)
        # and the engine continues with single batches afterwards
        more = conflict_free_batch(n, 40, present, rng)
        w = O.run_stream(g, oracle_queries(queries), [more], relabel=relabel)[0]
        p2, n2 = pipe.process_batch(*more)
        assert p2.tolist() == w["pos"] and n2.tolist() == w["neg"]
    finally:
        one.close()
        pipe.close()


@pytest.mark.parametrize("overlap", [True, False])
def test_process_batches_grows_pool(overlap):
    """Pipelined batches, where one update relocates nearly every
    list past the pool's slack. The pool starts at 1.5x the initial lists
    (about 6 entries per vertex here) plus 2^20 entries. 600,000 inserts raise
    the mean degree from 2 to 14, and the relocated lists need about 25
    entries per vertex, so the pool grows in the middle of the update (with
    overlap: the copy being updated on the second stream, while the other
    copy is matched). Counts, the graph after
    the stream, and a later pipelined call (the other copy, grown in turn)
    are checked against the oracle."""
    say("case: pipelined batches, pool growth, overlap=%s" % overlap)
    rng = np.random.default_rng(1818)
    n, m, nl = 100000, 100000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("square", [0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    k = 600000
    a = rng.integers(0, n, 2 * k)
    b = rng.integers(0, n, 2 * k)
    keys = np.unique(np.minimum(a, b)[a != b].astype(np.int64) * n + np.maximum(a, b)[a != b])
    old = np.array(sorted(x * n + y for x, y in present), np.int64)
    keys = keys[~np.isin(keys, old)]
    keys = rng.permutation(keys)[:k]
    assert len(keys) == k
    for key in keys.tolist():
        present.add((key // n, key % n))
    big = (np.ones(k, np.uint8), (keys // n).astype(np.uint32), (keys % n).astype(np.uint32),
           rng.integers(0, nl, k).astype(np.uint32))
    first = [conflict_free_batch(n, 2000, present, rng), big, conflict_free_batch(n, 3000, present, rng)]
    then = [conflict_free_batch(n, 3000, present, rng), conflict_free_batch(n, 1000, present, rng)]
    g = O.Graph(n, labels, eu, ev)
    want = O.run_stream(g, oracle_queries(queries), first, relabel=True)
    say("  oracle done")
    eng = E.Engine((n, labels, eu, ev), queries, relabel=True, overlap_apply=overlap)
    try:
        pos, neg, _ = eng.process_batches(first[:1])
        cap0 = eng.stats()["pool_entries_capacity"]
        p2, n2, _ = eng.process_batches(first[1:])
        pos, neg = np.concatenate([pos, p2]), np.concatenate([neg, n2])
        assert eng.stats()["pool_entries_capacity"] > cap0, "the pool did not grow"
        for bi in range(len(first)):
            assert pos[bi].tolist() == want[bi]["pos"], "batch %d positives" % bi
            assert neg[bi].tolist() == want[bi]["neg"], "batch %d negatives" % bi
        off, nb, _ = eng.graph()
        ooff, onb, _ = g.export()
        assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb)
        want2 = O.run_stream(g, oracle_queries(queries), then, relabel=True)
        pos, neg, _ = eng.process_batches(then)
        for bi in range(len(then)):
            assert pos[bi].tolist() == want2[bi]["pos"] and neg[bi].tolist() == want2[bi]["neg"], "later %d" % bi
        off, nb, _ = eng.graph()
        ooff, onb, _ = g.export()
        assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb)
    finally:
        eng.close()


def test_query_time_reports_and_resets():
    """dsm_get_query_time: per-query DFS task time since the last call; a
    query that seeds nothing costs nothing, and the call resets the sums."""
    say("case: per-query DFS time")
    rng = np.random.default_rng(1919)
    n, m, nl = 400, 4000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    # the third query's label 7 exists nowhere in the graph: no seeds
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("square", [0, 1, 0, 1]), qtuple("triangle", [7, 7, 7])]
    present = set(zip(eu.tolist(), ev.tolist()))
    eng = E.Engine((n, labels, eu, ev), queries, relabel=False)
    try:
        eng.query_time()
        eng.process_batch(*conflict_free_batch(n, 200, present, rng))
        t = eng.query_time()
        assert t.shape == (3,) and t[0] > 0 and t[1] > 0 and t[2] == 0
        assert np.all(eng.query_time() == 0)
    finally:
        eng.close()


@pytest.mark.parametrize("stats,neg", [(False, True), (True, True), (False, False)])
def test_cdsm_cli_counts(golden_dir, stats, neg):
    """The CDSM command line on the 200-batch small fixture prints, per batch,
    the expected `res:` counts (and `neg:` counts with --neg; without it the
    engine matches no deletes), pipelined (no --stats) or batch by batch
    (--stats)."""
    import subprocess
    say("case: CDSM command line, stats=%s neg=%s" % (stats, neg))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dsm_amd", "bin", "CDSM")
    exp = json.load(open(os.path.join(golden_dir, "small_expected.json")))
    cmd = [exe, "-q", os.path.join(golden_dir, "small_q.in"), "-d", os.path.join(golden_dir, "small_g.in"),
           "--gpu", "0"] + (["--neg"] if neg else []) + (["--stats"] if stats else [])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = [int(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("res:")]
    negs = [int(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("neg:")]
    assert res == [b["pos"] for b in exp["batches"]]
    assert negs == ([b["neg"] for b in exp["batches"]] if neg else [])


def test_tuple_capacity_truncation():
    """More tuples than tuple_capacity: the counts stay exact, dsm_get_tuples
    reports the number found and DSM_ETRUNC, and every tuple kept is one of
    the oracle's embeddings."""
    import ctypes
    say("case: tuple capacity truncation")
    rng = np.random.default_rng(2020)
    n, m, nl = 300, 3000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 0])]
    present = set(zip(eu.tolist(), ev.tolist()))
    batch = conflict_free_batch(n, 300, present, rng, p_add=1.0)
    g = O.Graph(n, labels, eu, ev)
    want = O.run_stream(g, oracle_queries(queries), [batch], relabel=True, tuples=True)[0]
    total = int(sum(want["pos"]) + sum(want["neg"]))
    cap = 16
    assert total > cap
    eng = E.Engine((n, labels, eu, ev), queries, relabel=True, tuple_capacity=cap)
    try:
        pos, neg = eng.process_batch(*batch)
        assert pos.tolist() == want["pos"] and neg.tolist() == want["neg"]
        num = ctypes.c_uint64()
        buf = np.zeros(cap * E.TUPLE_WORDS, np.uint32)
        rc = E.lib().dsm_get_tuples(eng._h, buf.ctypes.data, cap, ctypes.byref(num))
        assert rc == E.DSM_ETRUNC and num.value == total
        real = set(map(tuple, want["pos_tuples"][0].tolist()))
        for row in buf.reshape(cap, E.TUPLE_WORDS):
            assert row[0] == 0 and row[1] == 1
            assert tuple(row[2:5].tolist()) in real
    finally:
        eng.close()


def test_insert_batch_larger_than_merge_scratch():
    """One batch of 1.5M inserts on a 3,000-vertex graph: an apply round
    merges up to 1,024 inserts into each of the 3,000 lists, more than the
    pool's entries plus 2^20, so the merge scratch must grow before the merge
    (before the fix such merges were dropped silently). Graph and counts
    against the oracle; the query's labels keep the count small."""
    say("case: 1.5M inserts into 3,000 vertices (merge scratch growth)")
    rng = np.random.default_rng(4242)
    n = 3000
    labels, eu, ev = random_graph(n, 6000, 2, rng)
    labels[:5] = 2  # five vertices carry the query's label
    queries = [qtuple("triangle", [2, 2, 2])]
    k = 1_500_000
    src = rng.integers(0, n, k).astype(np.uint32)
    dst = rng.integers(0, n, k).astype(np.uint32)
    big = (np.ones(k, np.uint8), src, dst, np.zeros(k, np.uint32))
    compare_stream(n, labels, eu, ev, queries, [big], relabel=False, check_graph=True)


@pytest.mark.parametrize("overlap", [True, False])
def test_hub_copy_current_after_relabel_to_used_label(overlap):
    """Label-sorted copies and candidacy are kept only for vertices with a
    label some query uses. Two hubs (degree 5,000, built by a block, and 300,
    built by a wave) carry a label no query uses while their lists and their
    neighbours' labels change; then inserts relabel them to used labels, and
    their copies must be current from that batch on. Counts, graph and bitmaps
    against the oracle."""
    say("case: hubs with an unused label, relabelled to used labels (overlap %s)" % overlap)
    rng = np.random.default_rng(5151)
    n = 8000
    labels = rng.integers(0, 3, n).astype(np.uint32)
    labels[:2] = 3  # no query uses label 3
    edges = set()
    for hub, deg in ((0, 5000), (1, 300)):
        for v in rng.choice(np.arange(2, n), deg, replace=False):
            edges.add((hub, int(v)))
    while len(edges) < 5300 + 24000:
        a, b = map(int, rng.integers(2, n, 2))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    eu, ev = e[:, 0].copy(), e[:, 1].copy()
    queries = [qtuple("triangle", [0, 1, 2]), qtuple("square", [0, 1, 0, 2]), qtuple("tri_pendant", [1, 0, 2, 2])]
    present = set(edges)
    nbr = {0: [v for (h, v) in edges if h == 0], 1: [v for (h, v) in edges if h == 1]}

    def batch(records):
        add, src, dst, lab = zip(*records)
        return (np.array(add, np.uint8), np.array(src, np.uint32), np.array(dst, np.uint32),
                np.array(lab, np.uint32))

    def fresh(hub, k, lab):
        out = []
        while len(out) < k:
            v = int(rng.integers(2, n))
            if (hub, v) not in present and all(r[2] != v for r in out):
                present.add((hub, v))
                out.append((1, hub, v, lab))
        return out

    def neighbour_relabels(hub, k):
        # inserts between the hub's neighbours and other vertices: the
        # neighbours take used labels, so the hub's groups change
        out, seen = [], set()
        for v in rng.choice(nbr[hub], k, replace=False):
            w = int(rng.integers(2, n))
            key = (min(int(v), w), max(int(v), w))
            if int(v) == w or key in present or key in seen:
                continue
            seen.add(key)
            present.add(key)
            out.append((1, int(v), w, int(rng.integers(0, 3))))
        return out

    batches = []
    # 1-2: the hubs keep label 3 (their own inserts carry it) while their
    # lists grow and their neighbours are relabelled
    for _ in range(2):
        recs = fresh(0, 40, 3) + fresh(1, 20, 3) + neighbour_relabels(0, 300) + neighbour_relabels(1, 60)
        dels = [(0, h, v, 0) for h in (0, 1) for v in rng.choice(nbr[h], 10, replace=False)
                if (h, int(v)) in present]
        for _, h, v, _ in dels:
            present.discard((h, int(v)))
        batches.append(batch(recs + [(0, h, int(v), l) for _, h, v, l in dels]))
    # 3: inserts relabel the hubs to used labels
    batches.append(batch(fresh(0, 5, 1) + fresh(1, 5, 2) + neighbour_relabels(0, 100)))
    # 4: more changes around the now-used hubs
    batches.append(batch(fresh(0, 30, 1) + fresh(1, 30, 2) + neighbour_relabels(0, 200) + neighbour_relabels(1, 50)))
    compare_stream(n, labels, eu, ev, queries, batches, relabel=True, check_graph=True, overlap_apply=overlap)


def test_second_copy_catch_up_paths():
    """With overlap_apply, an update first brings the second copy of the graph
    up to date: by copying only what the previous update changed when the copy
    lags by exactly that update, else the whole graph. Insert-only batches
    (applied to the current copy directly, no negatives round to hide behind),
    mixed batches (overlapped), a hub whose list relocates, and an autotune
    call (whose trial update on the second copy is dropped) alternate here;
    counts, graph and bitmaps against the oracle after every batch."""
    say("case: second-copy catch-up across direct, overlapped and dropped updates")
    rng = np.random.default_rng(6161)
    n = 3000
    labels, eu, ev = random_graph(n, 15000, 3, rng)
    queries = [qtuple("triangle", [0, 1, 2]), qtuple("tri_pendant", [1, 0, 2, 0])]
    present = set(zip(eu.tolist(), ev.tolist()))
    g = O.Graph(n, labels, eu, ev)
    oq = oracle_queries(queries)
    eng = E.Engine((n, labels, eu, ev), queries, relabel=True)
    try:
        for bi in range(12):
            if bi % 3 == 1:
                add, src, dst, lab = conflict_free_batch(n, 200, present, rng, p_add=1.0)
            else:
                add, src, dst, lab = conflict_free_batch(n, 300, present, rng)
            if bi == 6:
                # a hub: many inserts at vertex 0 relocate its list
                batch_pairs = set(zip(np.minimum(src, dst).tolist(), np.maximum(src, dst).tolist()))
                extra = [(1, 0, int(v), 1) for v in rng.choice(np.arange(1, n), 400, replace=False)
                         if (0, int(v)) not in present and (0, int(v)) not in batch_pairs]
                for _, a, b, _ in extra:
                    present.add((a, b))
                ea = np.array([x[0] for x in extra], np.uint8)
                add = np.concatenate([add, ea])
                src = np.concatenate([src, np.zeros(len(extra), np.uint32)])
                dst = np.concatenate([dst, np.array([x[2] for x in extra], np.uint32)])
                lab = np.concatenate([lab, np.ones(len(extra), np.uint32)])
            want = O.run_stream(g, oq, [(add, src, dst, lab)], relabel=True)[0]
            pos, neg = eng.process_batch(add, src, dst, lab)
            assert pos.tolist() == want["pos"] and neg.tolist() == want["neg"], "batch %d counts" % bi
            off, nb, dlab = eng.graph()
            ooff, onb, olab = g.export()
            assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb), "batch %d graph" % bi
            for qi, q in enumerate(oq):
                cand = g.candidates(q)
                for u in range(q.n):
                    assert np.array_equal(eng.candidates(qi, u), cand[u].astype(bool)), \
                        "batch %d query %d vertex %d candidates" % (bi, qi, u)
            if bi == 8:
                # trial rounds run beside a dropped update of the second copy
                s_add, s_src, s_dst, s_lab = conflict_free_batch(n, 200, set(present), rng)
                eng.autotune(s_add, s_src, s_dst, s_lab)
    finally:
        eng.close()


def test_label_sorted_copies_dirty_in_consecutive_overlapped_updates():
    """The second copy's label-sorted lists when hubs are relabelled batch
    after batch: an overlapped update copies the previous update's dirty
    copies into its graph only after marking its own dirty vertices, and skips
    those, whose copies it rebuilds. Hub 0 is relabelled in every batch (its
    long-list neighbours are dirty in consecutive updates); hubs 1-2 and 3-4
    in alternate batches (their neighbours are dirty in one update and clean
    in the next, so their copies are copied). Dense graph, so most lists have
    label-sorted copies; pipelined batches against the oracle per batch."""
    say("case: label-sorted copies dirty in consecutive overlapped updates")
    rng = np.random.default_rng(7373)
    n, hubs = 1500, 5
    labels, eu, ev = random_graph(n, 45000, 3, rng)
    edges = set(zip(eu.tolist(), ev.tolist()))
    for h in range(hubs):
        for v in rng.choice(np.arange(hubs, n), 500, replace=False):
            edges.add((h, int(v)))
    e = np.array(sorted(edges), np.uint32)
    eu, ev = e[:, 0].copy(), e[:, 1].copy()
    present = set(edges)
    queries = [qtuple("triangle", [0, 1, 2]), qtuple("tri_pendant", [1, 0, 2, 0])]
    batches = []
    for bi in range(8):
        add, src, dst, lab = conflict_free_batch(n, 150, present, rng)
        pairs = set(zip(np.minimum(src, dst).tolist(), np.maximum(src, dst).tolist()))
        relabelled = [0] + ([1, 2] if bi % 2 == 0 else [3, 4])
        extra = []
        for h in relabelled:
            while True:
                v = int(rng.integers(hubs, n))
                if (h, v) not in present and (h, v) not in pairs:
                    break
            present.add((h, v))
            pairs.add((h, v))
            extra.append((h, v, (bi + h) % 3))
        add = np.concatenate([add, np.ones(len(extra), np.uint8)])
        src = np.concatenate([src, np.array([x[0] for x in extra], np.uint32)])
        dst = np.concatenate([dst, np.array([x[1] for x in extra], np.uint32)])
        lab = np.concatenate([lab, np.array([x[2] for x in extra], np.uint32)])
        batches.append((add, src, dst, lab))
    g = O.Graph(n, labels, eu, ev)
    want = O.run_stream(g, oracle_queries(queries), batches, relabel=True)
    say("  oracle done")
    eng = E.Engine((n, labels, eu, ev), queries, relabel=True, overlap_apply=True)
    try:
        pos, neg, _ = eng.process_batches(batches)
        for b in range(len(batches)):
            assert pos[b].tolist() == want[b]["pos"], "batch %d positives" % b
            assert neg[b].tolist() == want[b]["neg"], "batch %d negatives" % b
        off, nb, _ = eng.graph()
        ooff, onb, _ = g.export()
        assert off.tolist() == ooff.tolist() and np.array_equal(nb, onb)
    finally:
        eng.close()
