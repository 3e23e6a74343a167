"""Benchmark of the continuous subgraph matching path on one MI355X.

Default workload: the 1M-vertex R-MAT 64-query stream (R-MAT (.57,.19,.19,.05),
1,000,000 vertices, 10,000,000 undirected edges, 16 Zipf(1.0) labels; 10% of
the edges held out as inserts plus 10% of the initial edges as deletes,
interleaved, in batches of 4,096 records; 64 standing queries requested at
4-12 vertices, kept at 4-7: every larger draw exceeded the enumeration budget,
see the line's config). A step is one update batch: negatives in the pre-batch graph,
the in-place update, candidate refresh and positives in the post-batch graph,
for all 64 queries.

Legs (one JSON line on stdout, details on stderr):
  batched  one engine, all 64 queries matched per launch (the headline value)
  plain    the drop-in path as the reference runs it: one query per engine
           (one CDSM process per query), each engine doing the whole batch
           pipeline; run over the same timed batches as the batched leg
  served   the same stream cut into requests of log-uniform size (1..16,384
           records) arriving as a Poisson process at swept edge rates,
           served by one dsm_server (continuous batching, admission by
           estimated memory with 30% HBM headroom); per rate: achieved
           updates/s, p50/p99 request latency (arrival to counts ready)
  cpu      the CPU oracle (single core) on one batch of the same stream
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else None


def make_workload(name):
    from dsm_amd import workloads as W
    if name == "rmat1m_q64":
        return W.rmat_q64()
    if name == "rmat1m":
        return W.rmat()
    if name == "rmat16m":
        return W.rmat16m()
    if name == "small":
        return W.small()
    raise SystemExit("unknown workload %s" % name)


GRAPH_TEXT = {"rmat1m_q64": "R-MAT 1M V / 10M E, 16 Zipf(1.0) labels", "rmat1m": "R-MAT 1M V / 10M E, 16 Zipf(1.0) labels",
              "rmat16m": "R-MAT 16M V / 256M E, 32 Zipf(0.8) labels", "small": "uniform 1k V / 8k E, 4 labels"}


QUERY_RULE = {
    "rmat1m_q64": "requested as 64 shapes of 4-12 vertices (trees, cycles, 4-cliques, chorded cycles); each drawn by "
                  "randomized walk with labels copied from the graph and redrawn while its pilot (backtracking-tree "
                  "nodes over the first 12 batches) exceeds 2e7; rarer labels every 8 redraws, one vertex fewer after "
                  "160 draws; every 8-12-vertex draw exceeded the budget (dsm_amd/queries/rmat1m_q64.json "
                  "rejected_sizes; rmat1m_q64_large_pilots.json: pilots above 1.7e9 nodes at 8-12 vertices)",
    "rmat1m": "one 5-cycle with a chord plus a pendant (6 vertices), randomized walk, pilot budget 5e7",
    "rmat16m": "4-clique, 5-cycle with two chords and a tree requested at 7 vertices; redrawn while the pilot over "
               "the first 2 batches exceeds 2e8; 7- and 6-vertex trees never fit, the committed tree has 5 "
               "(dsm_amd/queries/rmat16m.json rejected_sizes)",
}


def query_histogram(w):
    h = {}
    for q in w.queries:
        h[str(q.n)] = h.get(str(q.n), 0) + 1
    return dict(sorted(h.items(), key=lambda kv: int(kv[0])))


def qtuples(w):
    return [(q.n, q.labels, q.eu, q.ev) for q in w.queries]


TUNE_SAMPLE = {"kind": "edges", "records": 16384}


def tune_sample(w):
    """The tuner's sample, matched read-only (every record that is an edge of
    the current graph seeds DFS roots, as a delete would). "edges": random
    edges of the initial graph, drawn with a fixed seed and independent of the
    stream's batches (the stream's inserts and deletes are random edges of the
    same graph); "batches": the records of the first 4 batches."""
    if TUNE_SAMPLE["kind"] == "batches":
        parts = [w.batch(b) for b in range(min(4, w.num_batches))]
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(4))
    rng = np.random.default_rng(777)
    k = min(TUNE_SAMPLE["records"], len(w.eu))
    idx = rng.choice(len(w.eu), k, replace=False)
    z = np.zeros(k, np.uint32)
    return np.zeros(k, np.uint8), w.eu[idx].astype(np.uint32), w.ev[idx].astype(np.uint32), z


def _agg_add(agg, st):
    agg["ms_dfs"] += st["ms_dfs_pos"] + st["ms_dfs_neg"]
    agg["launches"] += st["dfs_launches"]
    agg["entries"] += st["dfs_entries"]
    agg["loads"] += st["dfs_search_loads"]
    agg["merge_bytes"] += st["dfs_merge_bytes"]
    agg["busy_ms"] += st["dfs_busy_ms"]
    agg["waves"] += st["dfs_waves"]
    agg["imbalance"] += st["dfs_block_imbalance"]
    agg["head_idle"] += st["dfs_head_idle"]
    agg["tail_idle"] += st["dfs_tail_idle"]
    agg["ms_apply"] += st["ms_apply"]
    agg["ms_filter"] += st["ms_filter"]
    agg["ms_seed"] += st["ms_seed"]
    agg["ms_refine"] += st["ms_refine"]
    agg["seeds"] += st["seeds_pos"] + st["seeds_neg"]
    agg["donated"] += st["tasks_donated"]
    agg["filter_entries"] += st["filter_entries"]
    agg["dirty"] += st["dirty_groups"]
    agg["filter_bits"] += st["filter_bits"]


def run_batched(w, warmup, steps, engine_kw, autotune=True, pipelined=True):
    """One engine for all queries. Warm-up and timed batches go through
    dsm_process_batches (pipelined: the positives of a batch share a launch
    with the next batch's negatives, and the next batch is applied to the
    second copy of the graph meanwhile) or, with pipelined=False,
    dsm_process_batch one at a time. lat holds, pipelined, the interval between
    a batch's counts and the previous batch's; one at a time, each call's
    latency (records in, counts out)."""
    from dsm_amd import engine as E
    t0 = time.perf_counter()
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qtuples(w), **engine_kw)
    log("batched engine set-up %.2f s" % (time.perf_counter() - t0))
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        tuned = eng.autotune(*tune_sample(w))
        log("autotune %.2f s: %s" % (time.perf_counter() - t0, tuned))
    lat, recs = [], 0
    agg = {"ms_dfs": 0.0, "launches": 0, "entries": 0, "loads": 0, "merge_bytes": 0, "busy_ms": 0.0, "waves": 0,
           "imbalance": 0.0, "head_idle": 0.0, "tail_idle": 0.0, "ms_apply": 0.0, "ms_filter": 0.0, "ms_seed": 0.0, "seeds": 0, "donated": 0,
           "ms_refine": 0.0, "filter_entries": 0, "dirty": 0, "filter_bits": 0}
    pos_tot = neg_tot = 0
    per_batch = []
    try:
        if pipelined and warmup:
            t = time.perf_counter()
            pos, neg, _ = eng.process_batches([w.batch(b) for b in range(warmup)])
            per_batch += [(pos[b].copy(), neg[b].copy()) for b in range(warmup)]
            log("warmup: %d pipelined batches %.2f ms" % (warmup, (time.perf_counter() - t) * 1e3))
        for b in range(0 if pipelined else warmup):
            t = time.perf_counter()
            pos, neg = eng.process_batch(*w.batch(b))
            per_batch.append((pos.copy(), neg.copy()))
            log("warmup batch %d: %.2f ms" % (b, (time.perf_counter() - t) * 1e3))
        timed = [w.batch(b) for b in range(warmup, warmup + steps)]
        recs = sum(len(x[1]) for x in timed)
        qms = None
        if pipelined:
            eng.query_time()
            t0m = time.monotonic()
            pos, neg, done = eng.process_batches(timed)
            qms = eng.query_time()
            prev = t0m
            for b in range(steps):
                lat.append(done[b] - prev)
                prev = done[b]
                per_batch.append((pos[b].copy(), neg[b].copy()))
                pos_tot += int(pos[b].sum())
                neg_tot += int(neg[b].sum())
            st = eng.stats()
            _agg_add(agg, st)
            log("pipelined %d batches: %.2f ms, dfs %.2f ms (%d launches), apply %.2f ms, donated %d"
                % (steps, (done[-1] - t0m) * 1e3, st["ms_dfs_pos"], st["dfs_launches"], st["ms_apply"],
                   st["tasks_donated"]))
        else:
            for i, batch in enumerate(timed):
                t = time.perf_counter()
                pos, neg = eng.process_batch(*batch)
                dt = time.perf_counter() - t
                st = eng.stats()
                per_batch.append((pos.copy(), neg.copy()))
                lat.append(dt)
                pos_tot += int(pos.sum())
                neg_tot += int(neg.sum())
                _agg_add(agg, st)
                log("batch %d: %.2f ms  pos %d neg %d  dfs %.2f ms (%d launches) apply %.2f seeds %d donated %d"
                    % (warmup + i, dt * 1e3, int(pos.sum()), int(neg.sum()), st["ms_dfs_pos"] + st["ms_dfs_neg"],
                       st["dfs_launches"], st["ms_apply"], st["seeds_pos"] + st["seeds_neg"], st["tasks_donated"]))
    finally:
        eng.close()
    return {"lat": lat, "records": recs, "pos": pos_tot, "neg": neg_tot, "agg": agg, "per_batch": per_batch,
            "tuned": tuned, "query_ms": qms}


def query_share(w, qms, top=8):
    """The queries with the most DFS task time in the timed region, with their
    share of it."""
    if qms is None or qms.sum() <= 0:
        return None
    order = np.argsort(-qms)[:top]
    return [{"query": int(q), "vertices": int(w.queries[q].n), "edges": int(len(w.queries[q].eu)),
             "shape": getattr(w.queries[q], "shape", ""), "share": round(float(qms[q] / qms.sum()), 4)}
            for q in order]


def run_plain(w, warmup, steps, engine_kw):
    """One engine per query (the reference's CDSM runs one query per process);
    a batch is done when every query's engine has processed it."""
    from dsm_amd import engine as E
    nb = warmup + steps
    lat = np.zeros(steps)
    counts = np.zeros((nb, len(w.queries), 2), np.uint64)
    for qi, q in enumerate(qtuples(w)):
        eng = E.Engine((w.n, w.labels, w.eu, w.ev), [q], **engine_kw)
        try:
            for b in range(nb):
                batch = w.batch(b)
                t = time.perf_counter()
                pos, neg = eng.process_batch(*batch)
                dt = time.perf_counter() - t
                counts[b, qi] = (pos[0], neg[0])
                if b >= warmup:
                    lat[b - warmup] += dt
        finally:
            eng.close()
        if qi % 16 == 15:
            log("plain: %d/%d queries" % (qi + 1, len(w.queries)))
    recs = sum(len(w.batch(b)[1]) for b in range(warmup, nb))
    return {"lat": lat.tolist(), "records": recs, "counts": counts}


def run_served(w, rates, first_batch, engine_kw, seed=7, max_requests=160, offered_s=6.0, hi=16384,
               autotune=True):
    """Poisson arrivals of log-uniform-size requests at each rate in turn, on
    one server; the graph keeps evolving from one rate to the next."""
    import queue
    import threading
    from dsm_amd import engine as E
    rng = np.random.default_rng(seed)
    mean_size = (hi - 1) / np.log(hi)
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qtuples(w), **engine_kw)
    if autotune:
        log("served engine autotune: %s" % eng.autotune(*tune_sample(w)))
    srv = E.Server(eng)
    cur = {"b": first_batch, "buf": None, "pos": 0}

    def take(k):
        parts, need = [], k
        while need > 0:
            if cur["buf"] is None or cur["pos"] >= len(cur["buf"][1]):
                cur["buf"] = w.batch(cur["b"] % w.num_batches)
                cur["b"] += 1
                cur["pos"] = 0
            a, s_, d, l = cur["buf"]
            j = min(len(s_), cur["pos"] + need)
            parts.append((a[cur["pos"]:j], s_[cur["pos"]:j], d[cur["pos"]:j], l[cur["pos"]:j]))
            need -= j - cur["pos"]
            cur["pos"] = j
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(4))

    out = []
    try:
        # one untimed request: the server's first update allocates the
        # second copy of the graph
        st = srv.wait(srv.submit(*take(64)))[0]
        if st != E.DSM_OK:
            raise RuntimeError("served warm-up request failed: %s" % srv.error())
        for rate in rates:
            nreq = int(min(max_requests, max(20, rate * offered_s / mean_size)))
            sizes = np.exp(rng.uniform(0, np.log(hi + 1), nreq)).astype(np.int64).clip(1, hi)
            reqs = [take(int(k)) for k in sizes]
            gaps = rng.exponential(mean_size / rate, nreq)
            arrive = np.cumsum(gaps) - gaps[0]
            ids = queue.Queue()

            def submitter():
                t0 = time.perf_counter()
                for i, r in enumerate(reqs):
                    dt = arrive[i] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    ids.put(srv.submit(*r))

            th = threading.Thread(target=submitter)
            th.start()
            infos = []
            for _ in range(nreq):
                st, pos, neg, info = srv.wait(ids.get())
                if st != E.DSM_OK and st != E.DSM_EADMIT:
                    raise RuntimeError("served request failed: %s" % srv.error())
                infos.append(info)
            th.join()
            ok = [i for i in infos if i["status"] == E.DSM_OK]
            lat = [i["t_done"] - i["t_submit"] for i in ok]
            span = max(i["t_done"] for i in ok) - min(i["t_submit"] for i in infos)
            recs = int(sum(i["num_changes"] for i in ok))
            # the sizes drawn, over the arrival span: the load actually offered
            realized = float(sizes.sum()) / float(arrive[-1]) if arrive[-1] > 0 else None
            r = {"offered": rate, "offered_realized": round(realized, 1) if realized else None,
                 "requests": nreq, "rejected": nreq - len(ok), "records": recs,
                 "value": round(recs / span, 1), "p50_ms": round(pct(lat, 50) * 1e3, 3),
                 "p99_ms": round(pct(lat, 99) * 1e3, 3),
                 "fused_frac": round(float(np.mean([i["fused"] for i in ok])), 3),
                 "mean_queue_depth": round(float(np.mean([i["queue_depth"] for i in ok])), 2)}
            log("served rate %.0e: %s" % (rate, r))
            out.append(r)
    finally:
        srv.close()
        eng.close()
    return out


def run_cpu_threads(w, batch_index, threads):
    """CPU oracle on one whole batch with `threads` threads, one query at a
    time per thread (the oracle's count is reentrant and releases the GIL):
    negatives of every query in the pre-batch graph, the update (one thread),
    positives of every query."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    for b in range(batch_index):
        add, src, dst, lab = w.batch(b)
        g.apply(add, src, dst, lab, True)
    oq = [O.Query(q.n, q.labels, q.eu, q.ev) for q in w.queries]
    add, src, dst, lab = w.batch(batch_index)
    dels = add == 0

    def phase(ex, sel):
        # candidate sets per query, then (query, 1/8 of the records) tasks so
        # that one heavy query does not serialise the phase; per-query sums
        cands = list(ex.map(g.candidates, oq))
        s_, d_ = src[sel], dst[sel]
        parts = np.array_split(np.arange(len(s_)), 8)
        tasks = [(qi, p) for qi in range(len(oq)) for p in parts if len(p)]
        sums = list(ex.map(lambda t: int(g.count_records(oq[t[0]], s_[t[1]], d_[t[1]], cands[t[0]]).sum()),
                           tasks))
        tot = np.zeros(len(oq), np.uint64)
        for (qi, _), v in zip(tasks, sums):
            tot[qi] += np.uint64(v)
        return tot

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        neg = phase(ex, dels)
        g.apply(add, src, dst, lab, True)
        pos = phase(ex, ~dels)
    t = time.perf_counter() - t0
    return {"value": len(src) / t, "seconds": t, "pos": pos, "neg": neg,
            "sample": "batch %d (%d records), all %d queries, %d threads over (query, record share) tasks"
                      % (batch_index, len(src), len(oq), threads)}


def run_cpu(w, batch_index, budget_s):
    """CPU oracle on one batch (all queries, until the budget is spent)."""
    from oracle import oracle as O
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    for b in range(batch_index):
        add, src, dst, lab = w.batch(b)
        g.apply(add, src, dst, lab, True)
    oq = [O.Query(q.n, q.labels, q.eu, q.ev) for q in w.queries]
    add, src, dst, lab = w.batch(batch_index)
    dels = add == 0
    t0 = time.perf_counter()
    done = 0
    # negatives of every query, the update, positives of every query
    for q in oq:
        cand = g.candidates(q)
        g.count_records(q, src[dels], dst[dels], cand)
        done += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    frac_neg = done / len(oq)
    t_neg = time.perf_counter() - t0
    t1 = time.perf_counter()
    g.apply(add, src, dst, lab, True)
    t_apply = time.perf_counter() - t1
    t2 = time.perf_counter()
    done = 0
    for q in oq:
        cand = g.candidates(q)
        g.count_records(q, src[~dels], dst[~dels], cand)
        done += 1
        if time.perf_counter() - t2 > budget_s / 2:
            break
    frac_pos = done / len(oq)
    t_pos = time.perf_counter() - t2
    # extrapolate the per-query loops to all queries
    t_batch = t_neg / frac_neg + t_apply + t_pos / frac_pos
    return {"value": len(src) / t_batch, "seconds": t_batch,
            "sample": "batch %d (%d records), %d/%d queries timed for negatives and %d/%d for positives, "
                      "scaled to all queries" % (batch_index, len(src), round(frac_neg * len(oq)), len(oq),
                                                 round(frac_pos * len(oq)), len(oq))}


def calibrate_copy(nbytes):
    """One device-to-device copy of nbytes (the HIP runtime's copy kernel), the
    known-size reference for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration_copy_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat1m_q64")
    ap.add_argument("--plain-steps", type=int, default=0, help="0: the same batches as the batched leg (--steps)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-one-by-one", action="store_true",
                    help="skip the leg that calls dsm_process_batch once per batch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--served-rates", default="1e4,3e4,5e4,7e4,1e5,1.5e5,2e5,3e5,1e6,1e7")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--waves-per-block", type=int, default=4)
    ap.add_argument("--donate-min", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--lane-work", type=int, default=0)
    ap.add_argument("--leaf-sets", type=int, default=1)
    ap.add_argument("--donate-every", type=int, default=0)
    ap.add_argument("--donate-min-n2", type=int, default=0)
    ap.add_argument("--leaf-segments", type=int, default=1)
    ap.add_argument("--poll-max", type=int, default=0)
    ap.add_argument("--leaf-budget", type=int, default=0, help="0: the engine's default (2048)")
    ap.add_argument("--overlap-apply", type=int, default=1,
                    help="1: pipelined batches update a second copy of the graph during the round")
    ap.add_argument("--cluster-filter", type=int, default=0,
                    help="1: the headline engine refines candidates with the cluster encoding")
    ap.add_argument("--no-cluster-leg", action="store_true",
                    help="skip the leg that times the same stream with the cluster-encoding refinement")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-query plain leg (rmat1m: one 6-vertex query, one batch per call)")
    ap.add_argument("--tune-sample", default="edges", choices=["edges", "batches"],
                    help="the autotuner's sample: random edges of the initial graph (--tune-records of them), or "
                         "the records of the first 4 batches")
    ap.add_argument("--tune-records", type=int, default=TUNE_SAMPLE["records"],
                    help="random edges in the autotuner's sample")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="only run one D2D copy of this many bytes")
    args = ap.parse_args()
    TUNE_SAMPLE["kind"] = args.tune_sample
    TUNE_SAMPLE["records"] = args.tune_records
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine is single-GPU (multi-GPU sharding is design-only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    w = make_workload(args.workload)
    log("workload %s generated in %.1f s: %d vertices, %d initial edges, %d batches, %d queries"
        % (w.name, time.perf_counter() - t, w.n, len(w.eu), w.num_batches, len(w.queries)))
    engine_kw = {"waves_per_block": args.waves_per_block, "donate_min": args.donate_min,
                 "blocks_per_cu": args.blocks_per_cu, "lane_work": args.lane_work,
                 "leaf_sets": bool(args.leaf_sets), "donate_every": args.donate_every,
                 "donate_min_n2": args.donate_min_n2, "leaf_segments": bool(args.leaf_segments),
                 "poll_max": args.poll_max, "overlap_apply": bool(args.overlap_apply),
                 "leaf_budget": args.leaf_budget, "cluster_filter": bool(args.cluster_filter)}

    batch = int(w.bstart[1] - w.bstart[0])
    one = None if args.no_one_by_one else run_batched(w, args.warmup, args.steps, engine_kw,
                                                      autotune=not args.no_autotune, pipelined=False)
    bat = run_batched(w, args.warmup, args.steps, engine_kw, autotune=not args.no_autotune, pipelined=True)
    for b, ((p1, n1), (p2, n2)) in enumerate(zip(one["per_batch"] if one else [], bat["per_batch"])):
        assert np.array_equal(p1, p2) and np.array_equal(n1, n2), "pipelined and one-by-one differ at batch %d" % b
    total_s = sum(bat["lat"])
    value = bat["records"] / total_s
    ms_per_step = total_s / len(bat["lat"]) * 1e3
    agg = bat["agg"]
    # roofline of the dominant kernel (the DFS). Algorithmic bytes per launch:
    # the bytes the kernel's search has to read, counted by the kernel itself,
    # i.e. every source entry it tests (4 B id + 8 B candidate-bitmap word) and
    # every 4 B read of its membership searches. The work measure of a
    # merge-based enumerator (4 B x the degree of each backward list of each
    # expanded partial match, + 1/8 B per bit probe) is reported beside it: the
    # label groups, search windows and leaf sets skip most of those bytes, so
    # that figure divided by the launch time can exceed what HBM delivers.
    nl = max(agg["launches"], 1)
    dfs_ms_avg = agg["ms_dfs"] / nl
    bytes_per_launch = (agg["entries"] * 12 + agg["loads"] * 4) / nl
    merge_per_launch = (agg["merge_bytes"] + agg["entries"] / 8.0) / nl
    achieved = bytes_per_launch / (dfs_ms_avg * 1e-3) / 1e9 if dfs_ms_avg > 0 else 0.0
    merge_rate = merge_per_launch / (dfs_ms_avg * 1e-3) / 1e9 if dfs_ms_avg > 0 else 0.0
    # HBM bytes per launch measured by rocprofv3 (FETCH_SIZE / WRITE_SIZE runs
    # scaled by a copy of known size), kept in profiles/
    # The profile was taken on this workload's default run (the same command
    # under rocprofv3, tuned shape passed explicitly); it carries the
    # algorithmic bytes of its own launches, so traffic / algorithmic bytes is
    # a ratio of one run, not of two.
    traffic = prof = None
    tpath = os.path.join(ROOT, "profiles", "k_dfs_%s_traffic.json" % w.name)
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof.get("traffic_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": "k_dfs", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(dfs_ms_avg, 4),
                # what limits the kernel, from its counters (profiles/): waves
                # wait on dependent loads most of the time, far below the
                # bandwidth peak
                "limiter": "memory latency (dependent scattered loads)",
                "merge_equivalent": {"bytes_per_launch": int(merge_per_launch), "GB_per_s": round(merge_rate, 2)}}
    # The whole path per batch, in the survey's terms: B = B_upd + B_filt +
    # B_enum with B_upd = 2 x k x (4 B adjacency entry + 8 B delta record) and
    # B_filt = per re-filtered vertex 4 B offset + deg x (4 B id + 1 B label),
    # + 1/8 B per candidate bit written. B_enum, the merge-equivalent bytes, is
    # a work measure this search does not read (it exceeds HBM's peak when
    # divided by the batch time); the search's own bytes stand in for it in B.
    nbt = max(len(bat["lat"]), 1)
    b_upd = 2.0 * bat["records"] * 12 / nbt
    b_filt = (agg["dirty"] * 4 + agg["filter_entries"] * 5 + agg["filter_bits"] / 8.0) / nbt
    b_search = (agg["entries"] * 12 + agg["loads"] * 4) / nbt
    b_total = b_upd + b_filt + b_search
    t_batch = total_s / nbt
    roofline["per_batch"] = {"B_upd": int(b_upd), "B_filt": int(b_filt), "B_search": int(b_search),
                             "B_enum_merge_equivalent": int((agg["merge_bytes"] + agg["entries"] / 8.0) / nbt),
                             "B": int(b_total), "t_batch_ms": round(t_batch * 1e3, 3),
                             "achieved": round(b_total / t_batch / 1e9, 2), "unit": "GB/s",
                             "frac": round(b_total / t_batch / 1e9 / HBM_PEAK_GBS, 5)}
    if prof and prof.get("algorithmic_bytes_per_launch"):
        roofline["traffic_run"] = {"algorithmic_bytes_per_launch": prof["algorithmic_bytes_per_launch"],
                                   "avg_launch_ms": prof.get("avg_launch_ms"),
                                   "traffic_over_algorithmic": round(traffic / prof["algorithmic_bytes_per_launch"],
                                                                     3),
                                   "dram_GB_per_s": prof.get("dram_GB_per_s"),
                                   "profile": os.path.relpath(tpath, ROOT)}

    out = {"metric": "edge updates/s + p50/p99 batch latency, served/batched vs plain GPU path",
           "value": round(value, 1), "unit": "edge updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": "%s: %s, %d-record batches, %d standing queries, batched (one persistent "
                                  "launch per graph state)" % (w.name, GRAPH_TEXT.get(w.name, ""), batch,
                                                                len(w.queries)),
                      "batch": batch, "queries": len(w.queries),
                      "query_vertices": query_histogram(w),
                      "query_selection": QUERY_RULE.get(w.name, ""),
                      "cluster_filter": bool(args.cluster_filter),
                      "mode": "batched, pipelined (a batch's positives share a launch with the next batch's negatives; the next batch is applied to a second copy of the graph meanwhile)",
                      "parallelism": "single GPU", "tuned": bat["tuned"],
                      "tune_sample": ("%d random edges of the initial graph (seed 777)" % TUNE_SAMPLE["records"]
                                      if TUNE_SAMPLE["kind"] == "edges" else "records of the first 4 batches")},
           # pipelined batches complete one after another: the gap between two
           # batches' counts, not a batch's latency (a batch's positives finish
           # in the next batch's launch); per-batch latency is the one batch per
           # call leg's, and the served leg's per request
           "p50_interval_ms": round(pct(bat["lat"], 50) * 1e3, 3),
           "p99_interval_ms": round(pct(bat["lat"], 99) * 1e3, 3),
           "p50_batch_ms": round(pct(one["lat"], 50) * 1e3, 3) if one else None,
           "p99_batch_ms": round(pct(one["lat"], 99) * 1e3, 3) if one else None,
           "one_batch_per_call": {"value": round(one["records"] / sum(one["lat"]), 1),
                                  "p50_batch_ms": round(pct(one["lat"], 50) * 1e3, 3),
                                  "p99_batch_ms": round(pct(one["lat"], 99) * 1e3, 3),
                                  "ms_dfs_per_batch": round(one["agg"]["ms_dfs"] / len(one["lat"]), 3)}
           if one else None,
           "matches_per_s": round((bat["pos"] + bat["neg"]) / total_s, 1),
           "dfs_query_share": query_share(w, bat["query_ms"]),
           "stages_ms_per_batch": {k: round(agg[k] / len(bat["lat"]), 3) for k in ("ms_apply", "ms_filter",
                                                                               "ms_seed", "ms_dfs")},
           "roofline": roofline,
           # persistent-grid balance of the DFS: share of the waves' time spent on
           # tasks (the rest is waiting for donated work), and the busiest block's
           # task time over the mean block's, per launch
           "dfs_balance": {"wave_utilization": round(agg["busy_ms"] / max(agg["waves"] / nl * agg["ms_dfs"], 1e-9), 4),
                           "block_imbalance": round(agg["imbalance"] / nl, 4),
                           "head_idle": round(agg["head_idle"] / nl, 4),
                           "tail_idle": round(agg["tail_idle"] / nl, 4),
                           "waves_per_launch": int(agg["waves"] / nl)}}

    if not args.no_cluster_leg and w.name == "rmat1m_q64":
        # the same stream with the other candidate filter: counts must agree
        kw2 = dict(engine_kw, cluster_filter=not args.cluster_filter)
        cl = run_batched(w, args.warmup, args.steps, kw2, autotune=not args.no_autotune, pipelined=True)
        for b, ((p1, n1), (p2, n2)) in enumerate(zip(bat["per_batch"], cl["per_batch"])):
            assert np.array_equal(p1, p2) and np.array_equal(n1, n2), "filters disagree at batch %d" % b
        a2 = cl["agg"]
        out["other_filter"] = {"cluster_filter": kw2["cluster_filter"],
                               "value": round(cl["records"] / sum(cl["lat"]), 1),
                               "p50_interval_ms": round(pct(cl["lat"], 50) * 1e3, 3),
                               "p99_interval_ms": round(pct(cl["lat"], 99) * 1e3, 3),
                               "dfs_ms_per_launch": round(a2["ms_dfs"] / max(a2["launches"], 1), 4),
                               "refine_ms_per_batch": round(a2["ms_refine"] / len(cl["lat"]), 3),
                               "dfs_entries_per_launch": int(a2["entries"] / max(a2["launches"], 1))}
        out["dfs_entries_per_launch"] = int(agg["entries"] / nl)
        out["refine_ms_per_batch"] = round(agg["ms_refine"] / len(bat["lat"]), 3)

    if not args.no_single and w.name == "rmat1m_q64":
        # the single-query workload on the same graph and stream, through the
        # plain drop-in path (one dsm_process_batch per batch, default shapes)
        w1 = make_workload("rmat1m")
        sq = run_batched(w1, args.warmup, args.steps, dict(engine_kw, cluster_filter=False), autotune=False,
                         pipelined=False)
        out["single_query_plain"] = {
            "workload": "rmat1m: %s, %d-record batches, one 6-vertex query (5-cycle with a chord plus a pendant), "
                        "one dsm_process_batch per batch" % (GRAPH_TEXT["rmat1m"], batch),
            "value": round(sq["records"] / sum(sq["lat"]), 1), "steps": args.steps,
            "p50_batch_ms": round(pct(sq["lat"], 50) * 1e3, 3), "p99_batch_ms": round(pct(sq["lat"], 99) * 1e3, 3),
            "matches_per_s": round((sq["pos"] + sq["neg"]) / sum(sq["lat"]), 1)}

    if not args.no_plain:
        ps = min(args.plain_steps, args.steps) if args.plain_steps else args.steps
        pl = run_plain(w, args.warmup, ps, engine_kw)
        # the plain path must produce the same counts
        for b in range(args.warmup + ps):
            bp, bn = bat["per_batch"][b]
            assert np.array_equal(pl["counts"][b, :, 0], bp) and np.array_equal(pl["counts"][b, :, 1], bn), \
                "plain and batched counts differ at batch %d" % b
        plain_value = pl["records"] / sum(pl["lat"])
        out["plain"] = {"value": round(plain_value, 1), "steps": ps,
                        "p50_batch_ms": round(pct(pl["lat"], 50) * 1e3, 3),
                        "p99_batch_ms": round(pct(pl["lat"], 99) * 1e3, 3)}
        out["batched_vs_plain"] = round(value / plain_value, 3)

    if not args.no_served:
        rates = [float(x) for x in args.served_rates.split(",") if x]
        sv = run_served(w, rates, args.warmup + args.steps, engine_kw, autotune=not args.no_autotune)
        out["served"] = sv
        best = max(sv, key=lambda r: r["value"])
        out["served_peak"] = {"value": best["value"], "offered": best["offered"], "p50_ms": best["p50_ms"],
                              "p99_ms": best["p99_ms"]}
        if "plain" in out:
            out["served_vs_plain"] = round(best["value"] / out["plain"]["value"], 3)
            # the highest served throughput whose p99 request latency is at most
            # the plain path's p99 batch latency
            fit = [r for r in sv if r["p99_ms"] <= out["plain"]["p99_batch_ms"]]
            if fit:
                b2 = max(fit, key=lambda r: r["value"])
                out["served_at_plain_p99"] = {"value": b2["value"], "offered": b2["offered"], "p99_ms": b2["p99_ms"],
                                              "vs_plain": round(b2["value"] / out["plain"]["value"], 3)}

    if not args.no_cpu:
        # the oracle on the host's cores (at most 16, one GPU's share of the
        # host), and on one core with a time budget
        threads = max(1, min(16, os.cpu_count() or 1))
        cpu_n = run_cpu_threads(w, args.warmup, threads)
        # the oracle's counts of that batch are the GPU's (first timed batch)
        gp, gn = bat["per_batch"][args.warmup]
        oracle_ok = bool(np.array_equal(cpu_n["pos"], gp) and np.array_equal(cpu_n["neg"], gn))
        assert oracle_ok, "GPU counts differ from the oracle's at batch %d" % args.warmup
        cpu_1 = run_cpu(w, args.warmup, args.cpu_budget)
        log("cpu: %d threads %.1f updates/s; 1 thread %.1f updates/s" % (threads, cpu_n["value"], cpu_1["value"]))
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            pass
        out["cpu_baseline"] = {"value": round(cpu_n["value"], 2), "unit": "edge updates/s", "cores": threads,
                               "kind": "port", "sample": cpu_n["sample"],
                               "counts_equal_gpu": oracle_ok,
                               "one_core": {"value": round(cpu_1["value"], 2), "sample": cpu_1["sample"]},
                               "host": {"nproc": os.cpu_count(), "cpu_model": model}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
