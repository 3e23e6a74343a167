"""Pins the CPU oracle before it is trusted as the checker.

Known answers come from the reference's own fixture (test_data/example_*.in)
worked out by hand from the definitions, and from hand-derived update-rule
cases (cpuGraph.cpp:521-607). Counts and tuples are pinned against an
independent brute-force enumerator (no filter, no matching order, no
adjacency-list intersection). The LDF/NLF filter is pinned by the example's
candidate sets and by soundness: every embedding's image lies in it.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(12345)

SHAPES = {
    "edge": (2, [(0, 1)]),
    "path3": (3, [(0, 1), (1, 2)]),
    "triangle": (3, [(0, 1), (1, 2), (0, 2)]),
    "tri_pendant": (4, [(0, 1), (1, 2), (0, 2), (1, 3)]),
    "square": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "star4": (5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
    "clique4": (4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]),
    "house": (5, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4)]),
}


def random_graph(n, m, nl, rng):
    edges = set()
    while len(edges) < m:
        a, b = rng.integers(0, n, 2)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    return rng.integers(0, nl, n).astype(np.uint32), e[:, 0].copy(), e[:, 1].copy()


def random_query(shape, nl, rng):
    k, edges = SHAPES[shape]
    lab = rng.integers(0, nl, k).astype(np.uint32)
    e = np.array(edges, np.uint32)
    return O.Query(k, lab, e[:, 0], e[:, 1])


# ------------------------------------------------------------ the example

@pytest.fixture(scope="module")
def example(golden_dir):
    d = O.DsmFile(os.path.join(golden_dir, "example_g_01.in"), False)
    q = O.DsmFile(os.path.join(golden_dir, "example_q_01.in"), True)
    exp = json.load(open(os.path.join(golden_dir, "example_expected.json")))
    return d, q, exp


def test_example_parse_closed_form(example):
    d, q, _ = example
    assert (d.n, d.m, d.nb) == (11, 13, 1)
    assert (q.n, q.m, q.nb) == (5, 5, 0)
    assert q.labels.tolist() == [0, 0, 0, 1, 2]
    add, src, dst, lab = d.batch(0)
    assert add.tolist() == [1] and src.tolist() == [5] and dst.tolist() == [6] and lab.tolist() == [0]
    # v-line degrees equal the e-line incidence counts
    deg = np.bincount(np.concatenate([d.eu, d.ev]), minlength=d.n)
    assert deg.tolist() == d.degs.tolist()


def test_example_known_answers(example):
    d, q, exp = example
    g = O.Graph(d.n, d.labels, d.eu, d.ev)
    Q = O.Query(q.n, q.labels, q.eu, q.ev)
    assert g.brute_total(Q) == exp["embeddings_before"]
    res = O.run_stream(g, [Q], [d.batch(0)], relabel=True, tuples=True)[0]
    assert res["pos"] == [exp["batch0_pos"]] and res["neg"] == [exp["batch0_neg"]]
    assert sorted(res["pos_tuples"][0].tolist()) == exp["batch0_tuples"]
    assert g.brute_total(Q) == exp["embeddings_after"]
    cand = g.candidates(Q)
    for u, vs in exp["candidates_after_batch0"].items():
        assert np.nonzero(cand[int(u)])[0].tolist() == vs


# ------------------------------------------------------------ update rule

def _adj(g):
    off, nb, lab = g.export()
    return {v: nb[off[v]:off[v + 1]].tolist() for v in range(g.n)}, lab.tolist()


def test_update_relabel_both_endpoints_last_in_order_wins():
    # records (0,2,7) and (2,3,9): symmetrised and sorted (0,2),(2,0),(2,3),(3,2);
    # vertex 2's last add record in that order is (2,3) -> label 9
    g = O.Graph(4, [0, 1, 2, 3], [0], [1])
    g.apply([1, 1], [0, 2], [2, 3], [7, 9])
    adj, lab = _adj(g)
    assert lab == [7, 1, 9, 9]
    assert adj == {0: [1, 2], 1: [0], 2: [0, 3], 3: [2]}


def test_update_order_is_by_destination_not_file_order():
    # adds (0,3,label 1) then (0,1,label 2): sorted (0,1) before (0,3) -> label[0] = 1
    g = O.Graph(4, [5, 5, 5, 5], [], [])
    g.apply([1, 1], [0, 0], [3, 1], [1, 2])
    _, lab = _adj(g)
    assert lab == [1, 2, 5, 1]


def test_update_same_key_last_record_wins():
    g = O.Graph(3, [0, 0, 0], [0], [1])
    g.apply([0, 1], [0, 0], [1, 1], [5, 4])  # delete then add: present, relabelled 4
    adj, lab = _adj(g)
    assert adj[0] == [1] and lab[:2] == [4, 4]
    g.apply([1, 0], [1, 1], [2, 2], [6, 0])  # add then delete: absent, but relabelled 6
    adj, lab = _adj(g)
    assert adj[1] == [0] and adj[2] == [] and lab == [4, 6, 6]


def test_update_idempotent_and_noop_cases():
    g = O.Graph(3, [0, 1, 2], [0], [1])
    g.apply([0], [1], [2], [9])        # delete of an absent edge: nothing, no relabel
    g.apply([1], [0], [1], [0])        # re-insert of an existing edge: relabel only
    adj, lab = _adj(g)
    assert adj == {0: [1], 1: [0], 2: []} and lab == [0, 0, 2]
    g.apply([1], [2], [0], [3], relabel=False)
    adj, lab = _adj(g)
    assert adj == {0: [1, 2], 1: [0], 2: [0]} and lab == [0, 0, 2]


def test_update_conflicting_records_leave_one_direction():
    # 'a 0 1' and 'd 1 0' in one batch: key (0,1) sees add(pos 0) then the
    # reversed delete (pos 3) -> absent; key (1,0) sees delete(pos 1) then the
    # reversed add (pos 2) -> present
    g = O.Graph(2, [0, 0], [], [])
    g.apply([1, 0], [0, 1], [1, 0], [0, 0])
    adj, _ = _adj(g)
    assert adj == {0: [], 1: [0]}


def test_update_rejects_out_of_range():
    g = O.Graph(3, [0, 1, 2], [0], [1])
    with pytest.raises(ValueError):
        g.apply([1], [0], [3], [0])


def test_update_invariants_random():
    rng = np.random.default_rng(7)
    n = 60
    lab, eu, ev = random_graph(n, 150, 3, rng)
    g = O.Graph(n, lab, eu, ev)
    ref = {v: set() for v in range(n)}
    for a, b in zip(eu.tolist(), ev.tolist()):
        ref[a].add(b)
        ref[b].add(a)
    for _ in range(20):
        k = 30
        add = rng.integers(0, 2, k)
        src = rng.integers(0, n, k)
        dst = rng.integers(0, n, k)
        # records on one undirected pair carry one operation: the rule then keeps
        # the graph symmetric (a pair inserted and deleted in the same batch
        # can leave one direction only, as the reference's sort decides)
        op = {}
        for i in range(k):
            key = (min(src[i], dst[i]), max(src[i], dst[i]))
            add[i] = op.setdefault(key, add[i])
        g.apply(add, src, dst, rng.integers(0, 3, k))
        adj, _ = _adj(g)
        for v, lst in adj.items():
            assert lst == sorted(set(lst))               # sorted, deduplicated
            for w in lst:
                assert v in adj[w]                       # symmetric
        # set semantics: replay per directed key in (src,dst,record order)
        recs = sorted([(s, d, i, a) for i, (a, s, d) in enumerate(zip(add, src, dst))] +
                      [(d, s, k + i, a) for i, (a, s, d) in enumerate(zip(add, src, dst))])
        for s, d, _, a in recs:
            if a:
                ref[s].add(d)
            else:
                ref[s].discard(d)
        assert all(sorted(ref[v]) == adj[v] for v in range(n))


# ------------------------------------------------------------ filter

def test_ldf_nlf_sound_and_label_exact():
    rng = np.random.default_rng(11)
    for shape in SHAPES:
        n = 40
        lab, eu, ev = random_graph(n, 160, 2, rng)
        g = O.Graph(n, lab, eu, ev)
        q = random_query(shape, 2, rng)
        cand = g.candidates(q)
        emb = g.brute_embeddings(q)
        for u in range(q.n):
            assert np.all(cand[u][emb[:, u]]) if len(emb) else True
            # never a candidate with another label
            assert not np.any(cand[u] & (lab != q.labels[u]))


def test_ldf_nlf_hand_case():
    # path 0-1-2 with labels a,b,a; query edge u0(a)-u1(b)
    g = O.Graph(4, [0, 1, 0, 1], [0, 1], [1, 2])
    q = O.Query(2, [0, 1], [0], [1])
    cand = g.candidates(q)
    assert np.nonzero(cand[0])[0].tolist() == [0, 2]   # label a, one b-neighbour
    assert np.nonzero(cand[1])[0].tolist() == [1]      # vertex 3 has degree 0


# ------------------------------------------------------------ counts

@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_counts_match_brute_force(shape):
    rng = np.random.default_rng(abs(hash(shape)) % (1 << 31))
    n = 45
    lab, eu, ev = random_graph(n, 180, 2, rng)
    g = O.Graph(n, lab, eu, ev)
    q = random_query(shape, 2, rng)
    recs = [(a, b) for a, b in zip(eu.tolist(), ev.tolist())][:40]
    recs += [(b, a) for a, b in recs[:10]]                       # reversed orientation
    recs += [(int(rng.integers(0, n)), int(rng.integers(0, n))) for _ in range(10)]  # mostly non-edges
    src = np.array([r[0] for r in recs], np.uint32)
    dst = np.array([r[1] for r in recs], np.uint32)
    got = g.count_records(q, src, dst)
    want = [g.brute_record(q, s, d) for s, d in recs]
    assert got.tolist() == want
    # tuples: one per (record, embedding using that edge), all distinct per record
    tot = 0
    for (s, d), c in zip(recs, got.tolist()):
        t = g.tuples_records(q, [s], [d])
        assert len(t) == c
        assert len({tuple(x) for x in t.tolist()}) == c
        for f in t.tolist():
            assert len(set(f)) == q.n
            assert any({f[a], f[b]} == {s, d} for a, b in q.edges)
        tot += c
    assert tot == int(got.sum())


def test_counts_match_totals_identity():
    """For single-edge batches without relabelling, inserting a new edge adds
    exactly the embeddings that use it, and deleting one removes them."""
    rng = np.random.default_rng(5)
    n = 40
    lab, eu, ev = random_graph(n, 150, 2, rng)
    g = O.Graph(n, lab, eu, ev)
    q = random_query("tri_pendant", 2, rng)
    before = g.brute_total(q)
    pos = neg = 0
    present = set(zip(eu.tolist(), ev.tolist()))
    for _ in range(30):
        if rng.random() < 0.5:
            while True:
                a, b = map(int, rng.integers(0, n, 2))
                if a != b and (min(a, b), max(a, b)) not in present:
                    break
            present.add((min(a, b), max(a, b)))
            r = O.run_stream(g, [q], [([1], [a], [b], [0])], relabel=False)[0]
        else:
            a, b = sorted(present)[int(rng.integers(0, len(present)))]
            present.discard((a, b))
            r = O.run_stream(g, [q], [([0], [b], [a], [0])], relabel=False)[0]
        pos += r["pos"][0]
        neg += r["neg"][0]
    assert pos - neg == g.brute_total(q) - before


def test_multiset_over_records_counts_duplicates():
    # the same insert twice in a batch counts twice (no dedup, main.cu:268-329)
    g = O.Graph(3, [0, 0, 0], [0, 1], [1, 2])
    q = O.Query(3, [0, 0, 0], [0, 1, 0], [1, 2, 2])  # triangle
    r = O.run_stream(g, [q], [([1, 1], [0, 0], [2, 2], [0, 0])], relabel=True)[0]
    assert r["pos"] == [2 * 6]  # 6 ordered embeddings of a triangle use edge (0,2)


# ------------------------------------------------------ committed fixtures

@pytest.mark.parametrize("name", ["small", "dense"])
def test_oracle_reproduces_committed_fixtures(golden_dir, name):
    """The oracle (each record's count checked against the brute force again
    on the way) still writes the committed small and dense fixtures exactly."""
    import sys
    sys.path.insert(0, golden_dir)
    import make_fixtures as MF
    got = json.loads(json.dumps(getattr(MF, name + "_fixture")()))
    want = json.load(open(os.path.join(golden_dir, name + "_expected.json")))
    assert got == want


def test_tuple_digest_is_order_free_and_sensitive():
    """The digest of the dense fixture: row order does not matter; a changed
    entry, a swapped pair inside a row or a dropped row changes it."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_fixtures as MF
    t = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]], np.uint32)
    d = MF.tuple_digest(t)
    assert MF.tuple_digest(t[::-1]) == d
    assert MF.tuple_digest(t[[0, 1]]) != d
    u = t.copy()
    u[1, 2] = 70
    assert MF.tuple_digest(u) != d
    u = t.copy()
    u[0, [0, 1]] = u[0, [1, 0]]
    assert MF.tuple_digest(u) != d
    assert MF.tuple_digest(np.zeros((0, 4), np.uint32)) == 0
