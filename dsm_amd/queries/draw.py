"""Draws the committed standing-query sets of the R-MAT workloads.

  rmat1m.json      one 6-vertex query (5-cycle with the chord 0-2 plus a
                   pendant vertex), pilot budget 5e7 nodes
  rmat1m_q64.json  64 queries of 4-12 vertices (trees, cycles, 4-cliques,
                   chorded cycles; shapes from workloads.q64_specs, shrunk by
                   a vertex at a time when no draw fits), pilot budget 2e7
                   nodes each
  (both over the records of the first 12 batches of the R-MAT 1M / 10M stream)
  rmat16m.json     4-clique, 5-cycle with two chords, 7-vertex tree on the
                   R-MAT 16M / 256M graph, pilot budget 2e8 nodes each over
                   the records of the first 2 batches (65,536 records)

See workloads.draw_queries for the rule. Run from the repository root:
    python -m dsm_amd.queries.draw [rmat1m|rmat1m_q64]
"""
import json
import os
import sys
import time

from dsm_amd import workloads as W

SETS = {
    "rmat1m": {"make": W.rmat, "specs": [("cycle5_chord", 6)], "budget": 50_000_000, "window": 12, "nl": 16},
    "rmat1m_q64": {"make": W.rmat, "specs": W.q64_specs(), "budget": 20_000_000, "window": 12, "nl": 16},
    # dense shapes sit around hubs: the walk may start and step anywhere, and a
    # shape never falls back to a cycle or tree
    "rmat16m": {"make": W.rmat16m, "specs": W.RMAT16M_SPECS, "budget": 200_000_000, "window": 2, "nl": 32,
                "walk": (1 << 30, 1 << 30, 200000, 400), "fallback": False},
}


def main(names):
    for name in names:
        cfg = SETS[name]
        t = time.time()
        w = cfg["make"](name=name, queries=[])
        print("%s: workload built in %.1f s" % (name, time.time() - t), flush=True)
        qs = W.draw_queries(w, cfg["specs"], seed=2, nl=cfg["nl"], budget=cfg["budget"], window=cfg["window"],
                            log=lambda m: print(m, flush=True), walk=cfg.get("walk"),
                            fallback=cfg.get("fallback", True))
        out = {"workload": name, "rule": "randomized walk; redrawn while the pilot over the first %d batches "
                                          "exceeds the budget; rarer labels every 8 redraws; one vertex smaller "
                                          "after 160 draws" % cfg["window"],
               "budget_nodes": cfg["budget"], "window_batches": cfg["window"],
               "queries": [{"shape": q.shape, "labels": q.labels.tolist(), "edges": q.edges,
                            "pilot_nodes": q.cost, "draw": q.draw} for q in qs]}
        with open(os.path.join(W.QUERY_DIR, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print("%s: %d queries in %.1f s" % (name, len(qs), time.time() - t), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or sorted(SETS))
