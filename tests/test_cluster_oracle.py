"""Pins of the oracle's cluster-encoding refinement (clusterFilter,
filter.cu:1525-1654) on the CPU.

Known answers derived by hand from the reference's code:
  - the plans (spanning tree, vertex cover, clusters, pass order, positions)
    of the reference's example query, a path, a three-legged star (a combine
    that creates a cluster), a shape whose combine keeps an existing cluster,
    and a shape whose combine runs after the next layer's merges (its layer's
    checkpoint is never reached, filter.cu:1015-1066, :1132-1168);
  - the refined candidates of the example after its batch, and of the
    three-legged star on a 14-vertex graph (a merge that removes a centre
    candidate, then the combine's AND).
Closed forms: for a star (and a single edge) the refinement is
C'(centre) = C(centre), C'(leaf) = C(leaf) restricted to neighbours of C(centre).
Invariants on random graphs and shapes: LDF/NLF >= refined >= the projections
of the brute-force embeddings.
"""
import numpy as np
import pytest

from oracle import oracle as O

EXAMPLE_EDGES = [(0, 1), (0, 2), (1, 2), (1, 3), (3, 4)]  # example_q_01.in, file order
STAR3 = [(0, 1), (0, 2), (0, 3), (1, 4), (2, 5), (3, 6)]


def members(plan):
    return [(c["valid"], c["members"]) for c in plan["clusters"]]


def test_plan_of_the_example_query():
    p = O.cluster_plan(5, EXAMPLE_EDGES)
    # BFS tree from vertex 1 drops 0-2; cover {1, 3} in stack order
    assert p["cover"] == [1, 3]
    assert members(p) == [(True, [1, 0, 2, 3]), (True, [3, 1, 4]), (True, [3, 1])]
    assert p["final"] == 2
    assert p["enc_pos"] == [1, 8, 2, 7, 6]
    assert p["ops"] == [("encode", 0), ("encode", 1), ("merge", 2)]


def test_plan_of_a_path():
    p = O.cluster_plan(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    assert p["cover"] == [1, 3]
    assert members(p) == [(True, [1, 0, 2]), (True, [3, 2, 4]), (True, [2, 1, 3])]
    assert p["enc_pos"] == [1, 7, 6, 8, 5]
    assert p["ops"] == [("encode", 0), ("encode", 1), ("merge", 2)]


def test_plan_combine_creates_a_cluster():
    p = O.cluster_plan(7, STAR3)
    assert p["cover"] == [3, 2, 1]
    assert members(p) == [(True, [3, 0, 6]), (True, [2, 0, 5]), (True, [1, 0, 4]),
                          (False, [0, 3, 2]), (False, [0, 3, 1]), (False, [0, 2, 1]),
                          (True, [0, 1, 2, 3])]
    assert p["final"] == 6
    assert p["enc_pos"] == [18, 19, 20, 21, 8, 5, 2]
    assert p["ops"] == [("encode", 0), ("encode", 1), ("encode", 2), ("merge", 3), ("merge", 4),
                        ("merge", 5), ("combine", 0)]


def test_plan_combine_keeps_an_existing_cluster():
    # layer 1: [3,0] [1,0] [0,3,1]; layer 2: [0,3,1] [0,3] [0,1]: the last two
    # are folded into the first, which already holds all of them
    p = O.cluster_plan(6, [(0, 1), (4, 2), (3, 4), (3, 1), (1, 2), (0, 5), (0, 3)])
    assert p["cover"] == [0, 3, 1]
    assert members(p) == [(True, [0, 1, 5, 3]), (True, [3, 4, 0]), (True, [1, 0, 2]),
                          (True, [3, 0]), (True, [1, 0]), (True, [0, 3, 1]),
                          (True, [0, 3, 1]), (False, [0, 3]), (False, [0, 1])]
    assert p["final"] == 6
    assert p["enc_pos"] == [17, 19, 9, 18, 5, 2]
    assert p["ops"] == [("encode", 0), ("encode", 1), ("encode", 2)] + \
        [("merge", c) for c in range(3, 9)] + [("combine", 0)]


def test_plan_combine_after_the_next_layer():
    # the layer-1 combine keeps cluster 4 and creates nothing, so that layer's
    # checkpoint is its end, which the pass loop never tests; the combine runs
    # in the final pass, after the layer-2 merge (cluster 5 and 6)
    p = O.cluster_plan(7, [(1, 2), (0, 5), (0, 1), (1, 3), (2, 4), (0, 6), (5, 6)])
    assert p["cover"] == [0, 1, 2]
    assert members(p) == [(True, [0, 5, 1, 6]), (True, [1, 2, 0, 3]), (True, [2, 1, 4]),
                          (False, [1, 0]), (True, [1, 0, 2]), (True, [2, 1]), (True, [2, 1])]
    assert p["ops"] == [("encode", 0), ("encode", 1), ("encode", 2), ("merge", 3), ("merge", 4),
                        ("merge", 5), ("merge", 6), ("combine", 0)]


def sets(rows):
    return {u: np.nonzero(r)[0].tolist() for u, r in enumerate(rows)}


def test_refined_candidates_of_the_example(golden_dir):
    d = O.DsmFile(golden_dir + "/example_g_01.in", False)
    q = O.DsmFile(golden_dir + "/example_q_01.in", True)
    assert list(zip(q.eu.tolist(), q.ev.tolist())) == EXAMPLE_EDGES
    g = O.Graph(d.n, d.labels, d.eu, d.ev)
    oq = O.Query(q.n, q.labels, q.eu, q.ev)
    g.apply(*d.batch(0))
    ldf = g.candidates(oq)
    assert sets(ldf)[0] == [0, 1, 4, 5, 6, 7, 8]
    # vertex 1 has no neighbour among C(u1) = {1, 6, 8}, so it leaves u0 and u2
    assert sets(g.cluster_candidates(oq, ldf)) == {0: [0, 4, 5, 6, 7, 8], 1: [1, 6, 8],
                                                   2: [0, 4, 5, 6, 7, 8], 3: [2, 9], 4: [3, 10]}


def test_refined_candidates_of_the_three_legged_star():
    # centre 0 (label 0), legs 1,2,3 (labels 1,2,3), leaves 4,5,6 (label 4).
    # Data: a full star on 0..6, and a second centre 7 whose leg 10 has no
    # leaf (so 10 fails LDF for leg 3). The first merge [0,3,2] drops 7 from
    # the centre's list; the later passes and the combine never see it.
    lab = [0, 1, 2, 3, 4, 4, 4, 0, 1, 2, 3, 4, 4, 9]
    e = [(0, 1), (0, 2), (0, 3), (1, 4), (2, 5), (3, 6), (7, 8), (7, 9), (7, 10), (8, 11), (9, 12)]
    g = O.Graph(14, lab, [a for a, _ in e], [b for _, b in e])
    q = O.Query(7, [0, 1, 2, 3, 4, 4, 4], [a for a, _ in STAR3], [b for _, b in STAR3])
    ldf = g.candidates(q)
    assert sets(ldf) == {0: [0, 7], 1: [1, 8], 2: [2, 9], 3: [3], 4: [4, 11], 5: [5, 12], 6: [6]}
    assert sets(g.cluster_candidates(q, ldf)) == {0: [0], 1: [1], 2: [2], 3: [3], 4: [4, 11], 5: [5, 12],
                                                  6: [6]}


def random_graph(n, m, nl, rng):
    edges = set()
    while len(edges) < m:
        a, b = map(int, rng.integers(0, n, 2))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.uint32)
    return rng.integers(0, nl, n).astype(np.uint32), e[:, 0].copy(), e[:, 1].copy()


def test_star_and_edge_closed_form():
    rng = np.random.default_rng(5)
    n = 300
    lab, eu, ev = random_graph(n, 1200, 3, rng)
    g = O.Graph(n, lab, eu, ev)
    off, nb, _ = g.export()
    for k in (1, 2, 3, 4):  # a single edge, then stars with 2..4 leaves
        for trial in range(4):
            ql = rng.integers(0, 3, k + 1).astype(np.uint32)
            edges = [(0, i) for i in range(1, k + 1)]
            perm = rng.permutation(k)
            edges = [edges[i] if rng.random() < 0.5 else edges[i][::-1] for i in perm]
            q = O.Query(k + 1, ql, [a for a, _ in edges], [b for _, b in edges])
            ldf = g.candidates(q)
            ref = g.cluster_candidates(q, ldf)
            # the centre: vertex 0 of a star; for one edge, the end the plan
            # roots at (the first vertex of highest degree, i.e. 0)
            c = ldf[0].astype(bool)
            near = np.zeros(n, bool)
            for v in np.nonzero(c)[0]:
                near[nb[off[v]:off[v + 1]]] = True
            assert np.array_equal(ref[0].astype(bool), c), (k, trial)
            for u in range(1, k + 1):
                assert np.array_equal(ref[u].astype(bool), ldf[u].astype(bool) & near), (k, trial, u)


SHAPES = {
    "path3": [(0, 1), (1, 2)],
    "triangle": [(0, 1), (1, 2), (0, 2)],
    "tri_pendant": [(0, 1), (1, 2), (0, 2), (1, 3)],
    "square": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "clique4": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],
    "house": [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4)],
    "path6": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)],
    "star3legs": STAR3,
    "cycle6_chord": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (0, 3)],
    "tree7": [(0, 1), (1, 2), (1, 3), (3, 4), (0, 5), (5, 6)],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_refinement_is_sound(shape):
    rng = np.random.default_rng(hash(shape) % 1000)
    edges = SHAPES[shape]
    k = max(max(e) for e in edges) + 1
    for trial in range(3):
        n, nl = 60, 2
        lab, eu, ev = random_graph(n, 260, nl, rng)
        g = O.Graph(n, lab, eu, ev)
        order = rng.permutation(len(edges))
        ed = [edges[i] for i in order]
        q = O.Query(k, rng.integers(0, nl, k).astype(np.uint32), [a for a, _ in ed], [b for _, b in ed])
        ldf = g.candidates(q).astype(bool)
        ref = g.cluster_candidates(q).astype(bool)
        emb = g.brute_embeddings(q)
        proj = np.zeros_like(ldf)
        for u in range(k):
            proj[u, emb[:, u]] = True
        assert not (ref & ~ldf).any(), "refined outside LDF/NLF"
        assert not (proj & ~ref).any(), "an embedding's vertex was filtered out"
        if trial == 0:
            assert len(emb) > 0 or shape in ("clique4", "house", "cycle6_chord")


def test_engine_plan_equals_oracle_plan():
    # the engine's planner (dsm_cluster_plan, host code of libdsm_amd.so, no GPU
    # needed) against the oracle's, on random connected queries of 2..16
    # vertices with edge lines in random order and orientation
    from dsm_amd import engine as E
    rng = np.random.default_rng(2024)
    for t in range(3000):
        n = int(rng.integers(2, 17))
        edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
        for _ in range(int(rng.integers(0, n))):
            a, b = map(int, rng.integers(0, n, 2))
            if a != b and (a, b) not in edges and (b, a) not in edges:
                edges.append((a, b))
        perm = rng.permutation(len(edges))
        edges = [edges[i] if rng.random() < 0.5 else edges[i][::-1] for i in perm]
        want = O.cluster_plan_words(n, edges)
        got = E.cluster_plan_words(n, edges)
        assert want is not None and got is not None, (n, edges)
        assert np.array_equal(got, want), (n, edges)
    for name, edges in SHAPES.items():
        k = max(max(e) for e in edges) + 1
        assert np.array_equal(E.cluster_plan_words(k, edges), O.cluster_plan_words(k, edges)), name
