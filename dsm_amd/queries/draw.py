"""Draws the committed standing-query sets of the R-MAT workloads.

  rmat1m.json      one 6-vertex query (5-cycle with the chord 0-2 plus a
                   pendant vertex), pilot budget 5e7 nodes
  rmat1m_q64.json  64 queries requested at 4-12 vertices (trees, cycles,
                   4-cliques, chorded cycles; shapes from workloads.q64_specs,
                   shrunk by a vertex at a time when no draw fits), pilot
                   budget 2e7 nodes each; the kept set has 4-7 vertices, and
                   rejected_sizes records every size given up
  (both over the records of the first 12 batches of the R-MAT 1M / 10M stream)
  rmat16m.json     4-clique, 5-cycle with two chords and a tree requested at
                   7 vertices on the R-MAT 16M / 256M graph, pilot budget 2e8
                   nodes each over the records of the first 2 batches (65,536
                   records); the tree fits only at 5 vertices
  rmat1m_q64_large_pilots.json
                   pilots of 8-, 10- and 12-vertex cycles, chorded cycles and
                   trees on rarer labels with a budget of 2e9 nodes

See workloads.draw_queries for the rule. Run from the repository root:
    python -m dsm_amd.queries.draw [rmat1m|rmat1m_q64|rmat16m]
    python -m dsm_amd.queries.draw large_pilots
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
        rejected = []
        qs = W.draw_queries(w, cfg["specs"], seed=2, nl=cfg["nl"], budget=cfg["budget"], window=cfg["window"],
                            log=lambda m: print(m, flush=True), walk=cfg.get("walk"),
                            fallback=cfg.get("fallback", True), rejected=rejected)
        out = {"workload": name, "rule": "randomized walk; redrawn while the pilot over the first %d batches "
                                          "exceeds the budget; rarer labels every 8 redraws; one vertex smaller "
                                          "after 160 draws" % cfg["window"],
               "budget_nodes": cfg["budget"], "window_batches": cfg["window"],
               "queries": [{"shape": q.shape, "labels": q.labels.tolist(), "edges": q.edges,
                            "pilot_nodes": q.cost, "draw": q.draw} for q in qs],
               # every (spec, shape, size) given up before the committed draw:
               # draws made, walks that found no embedding, draws over the
               # pilot budget and the smallest such pilot (budget + 1 means the
               # pilot stopped at the budget)
               "rejected_sizes": rejected}
        with open(os.path.join(W.QUERY_DIR, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print("%s: %d queries in %.1f s" % (name, len(qs), time.time() - t), flush=True)


def large_pilots(budget=2_000_000_000, min_label=11):
    """Evidence for the sizes the 64-query set could not keep: for cycles,
    chorded cycles and trees of 8, 10 and 12 vertices, the first randomized
    walk on the 1M / 10M graph through vertices of the rarer labels (>=
    min_label, about 11% of the vertices) that finds an embedding, and its pilot
    over the records of the first 12 batches with a budget of `budget` nodes
    (budget + 1 = more). Written to rmat1m_q64_large_pilots.json."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    w = W.rmat(name="rmat1m_q64", queries=[])
    off, nbr = w.csr()
    hi = int(w.bstart[12])
    deg = np.diff(off).astype(np.int64)
    rs, rd = w.src[:hi], w.dst[:hi]
    order = np.argsort(-(deg[rs] + deg[rd]), kind="stable")
    rs, rd = np.ascontiguousarray(rs[order]), np.ascontiguousarray(rd[order])

    def one(job):
        kind, k = job
        for t in range(24):
            q = W.extract_query(w.n, off, nbr, w.labels, kind, k, 7000 + k * 131 + t * 17 + min_label * 1000,
                                max_start_deg=256, max_walk_deg=1024, budget=20000, attempts=200,
                                min_label=min_label)
            if q is None:
                continue
            t0 = time.time()
            c = W.pilot_nodes(w.n, off, nbr, w.labels, q, rs, rd, budget)
            return {"shape": "%s%d" % (kind, k), "labels": q.labels.tolist(), "walk": t, "pilot_nodes": c,
                    "over_budget": c > budget, "seconds": round(time.time() - t0, 1)}
        return {"shape": "%s%d" % (kind, k), "labels": None, "walk": None, "pilot_nodes": None}

    jobs = [(kind, k) for kind in ("cycle", "chord_cycle", "tree") for k in (8, 10, 12)]
    with ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(one, jobs))
    for r in res:
        print(r, flush=True)
    out = {"workload": "rmat1m_q64", "budget_nodes": budget, "window_batches": 12, "min_label": min_label,
           "committed_budget_nodes": SETS["rmat1m_q64"]["budget"], "draws": res}
    with open(os.path.join(W.QUERY_DIR, "rmat1m_q64_large_pilots.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if sys.argv[1:] == ["large_pilots"]:
        large_pilots()
    else:
        main(sys.argv[1:] or sorted(SETS))
