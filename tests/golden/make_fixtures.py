"""Writes the golden fixtures of the GPU parity tests.

  example_g_01.in / example_q_01.in   the reference's own input pair
                                      (test_data/, copied as data)
  example_expected.json               hand-derived known answers for it: the
                                      LDF/NLF candidate sets and the batch-0
                                      count and tuples, worked out from the
                                      definitions (no reference output exists)
  small_g.in / small_q.in             the 1,000-vertex workload (uniform
                                      8,000 edges, 4 labels, triangle+pendant
                                      query, 100 inserts then 100 deletes,
                                      one record per batch), DSM text format
  small_expected.json                 per-batch positive/negative counts and
                                      sorted tuples of that workload from the
                                      CPU oracle, each record's count checked
                                      here against the brute-force enumerator
  dense_expected.json                 the same for the 400-vertex, mean-degree
                                      40 workload (8 interleaved records per
                                      batch), with a digest of each batch's
                                      tuples in place of the tuples; the graph
                                      is regenerated from its seed by
                                      dsm_amd.workloads.dense()

Run from the repository root: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from dsm_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

# Known answers for the reference's example (data graph 11 V / 13 E, query:
# triangle 0-1-2 plus path 1-3-4, labels 0,0,0,1,2; one batch `a 5 6 0`),
# derived by hand from the LDF/NLF definition (filter.cu:475-547) and the
# seeding rule (main.cu:268-329).
EXAMPLE = {
    "candidates_after_batch0": {
        "0": [0, 1, 4, 5, 6, 7, 8],
        "1": [1, 6, 8],
        "2": [0, 1, 4, 5, 6, 7, 8],
        "3": [2, 9],
        "4": [3, 10],
    },
    "embeddings_before": 4,
    "embeddings_after": 8,
    "batch0_pos": 4,
    "batch0_neg": 0,
    # embeddings f = (f(u0),..,f(u4)) that use the inserted edge (5,6)
    "batch0_tuples": [[5, 6, 8, 9, 10], [5, 8, 6, 9, 10], [6, 8, 5, 9, 10], [8, 6, 5, 9, 10]],
}


def small_fixture():
    w = W.small()
    data = os.path.join(HERE, "small_g.in")
    query = os.path.join(HERE, "small_q.in")
    w.write(data, [query])
    q = w.queries[0]
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    oq = O.Query(q.n, q.labels, q.eu, q.ev)
    batches = []
    for b in range(w.num_batches):
        add, src, dst, lab = w.batch(b)
        dels = add == 0
        cand = g.candidates(oq)
        neg_rec = g.count_records(oq, src[dels], dst[dels], cand)
        for s, d, c in zip(src[dels], dst[dels], neg_rec):
            assert c == g.brute_record(oq, s, d), "oracle != brute force (neg)"
        neg_t = g.tuples_records(oq, src[dels], dst[dels], cand)
        g.apply(add, src, dst, lab, relabel=True)
        cand = g.candidates(oq)
        pos_rec = g.count_records(oq, src[~dels], dst[~dels], cand)
        for s, d, c in zip(src[~dels], dst[~dels], pos_rec):
            assert c == g.brute_record(oq, s, d), "oracle != brute force (pos)"
        pos_t = g.tuples_records(oq, src[~dels], dst[~dels], cand)
        batches.append({
            "pos": int(pos_rec.sum()), "neg": int(neg_rec.sum()),
            "pos_tuples": sorted(map(list, pos_t.tolist())),
            "neg_tuples": sorted(map(list, neg_t.tolist())),
        })
    off, nbr, lab = g.export()
    final = {"entries": int(len(nbr)),
             "checksum": int((nbr.astype(np.uint64) * 2654435761 % (1 << 32)).sum() % (1 << 61)),
             "labels_checksum": int((lab.astype(np.uint64) * (np.arange(len(lab), dtype=np.uint64) + 1)).sum())}
    return {"workload": "small", "relabel": True, "query": {"labels": q.labels.tolist(), "edges": q.edges},
            "batches": batches, "final_graph": final}


def tuple_digest(t):
    """Order-independent digest of a set of tuples (rows): the sum over rows of
    a 64-bit mix of the row, mod 2^61."""
    t = np.asarray(t, np.uint64)
    if t.size == 0:
        return 0
    t = t.reshape(len(t), -1)
    h = np.zeros(len(t), np.uint64)
    with np.errstate(over="ignore"):
        for c in range(t.shape[1]):
            h = (h ^ t[:, c]) * np.uint64(0x9E3779B97F4A7C15)
            h ^= h >> np.uint64(29)
    return int(sum(int(x) for x in h) % (1 << 61))


def dense_fixture():
    w = W.dense()
    q = w.queries[0]
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    oq = O.Query(q.n, q.labels, q.eu, q.ev)
    batches = []
    for b in range(w.num_batches):
        add, src, dst, lab = w.batch(b)
        dels = add == 0
        cand = g.candidates(oq)
        neg_rec = g.count_records(oq, src[dels], dst[dels], cand)
        for s_, d_, c in zip(src[dels], dst[dels], neg_rec):
            assert c == g.brute_record(oq, s_, d_), "oracle != brute force (neg)"
        neg_t = g.tuples_records(oq, src[dels], dst[dels], cand)
        g.apply(add, src, dst, lab, relabel=True)
        cand = g.candidates(oq)
        pos_rec = g.count_records(oq, src[~dels], dst[~dels], cand)
        for s_, d_, c in zip(src[~dels], dst[~dels], pos_rec):
            assert c == g.brute_record(oq, s_, d_), "oracle != brute force (pos)"
        pos_t = g.tuples_records(oq, src[~dels], dst[~dels], cand)
        assert len(pos_t) == pos_rec.sum() and len(neg_t) == neg_rec.sum()
        batches.append({"pos": int(pos_rec.sum()), "neg": int(neg_rec.sum()),
                        "pos_digest": tuple_digest(pos_t), "neg_digest": tuple_digest(neg_t)})
    return {"workload": "dense", "relabel": True, "query": {"labels": q.labels.tolist(), "edges": q.edges},
            "batches": batches}


def main():
    with open(os.path.join(HERE, "example_expected.json"), "w") as f:
        json.dump(EXAMPLE, f, indent=1)
    with open(os.path.join(HERE, "small_expected.json"), "w") as f:
        json.dump(small_fixture(), f)
    with open(os.path.join(HERE, "dense_expected.json"), "w") as f:
        json.dump(dense_fixture(), f)
    print("fixtures written")


if __name__ == "__main__":
    main()
