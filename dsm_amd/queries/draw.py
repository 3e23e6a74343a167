"""Draws the committed standing-query sets of the R-MAT workloads.

  rmat1m.json      one 6-vertex query (5-cycle with the chord 0-2 plus a
                   pendant vertex), pilot budget 5e7 nodes
  rmat1m_q64.json  64 queries of 4-12 vertices (trees, cycles, 4-cliques,
                   chorded cycles; shapes from workloads.q64_specs, shrunk by
                   a vertex at a time when no draw fits), pilot budget 2e7
                   nodes each

Both over the records of the first 12 batches of the R-MAT 1M / 10M stream
(see workloads.draw_queries for the rule). Run from the repository root:
    python -m dsm_amd.queries.draw [rmat1m|rmat1m_q64]
"""
import json
import os
import sys
import time

from dsm_amd import workloads as W

SETS = {
    "rmat1m": {"specs": [("cycle5_chord", 6)], "budget": 50_000_000},
    "rmat1m_q64": {"specs": W.q64_specs(), "budget": 20_000_000},
}


def main(names):
    for name in names:
        cfg = SETS[name]
        t = time.time()
        w = W.rmat(name=name, queries=[])
        qs = W.draw_queries(w, cfg["specs"], seed=2, nl=16, budget=cfg["budget"],
                            log=lambda m: print(m, flush=True))
        out = {"workload": name, "rule": "randomized walk; redrawn while the pilot over the first 12 batches "
                                          "exceeds the budget; rarer labels every 8 redraws; one vertex smaller "
                                          "after 160 draws",
               "budget_nodes": cfg["budget"], "window_batches": 12,
               "queries": [{"shape": q.shape, "labels": q.labels.tolist(), "edges": q.edges,
                            "pilot_nodes": q.cost, "draw": q.draw} for q in qs]}
        with open(os.path.join(W.QUERY_DIR, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print("%s: %d queries in %.1f s" % (name, len(qs), time.time() - t), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or sorted(SETS))
