"""Parity at the R-MAT workloads' own sizes: the HIP engine (through the C ABI)
against the CPU oracle, bit-exact per batch and query (uint64 positive and
negative counts).

  rmat1m       R-MAT 1M V / 10M E, 16 Zipf(1.0) labels, 4,096-record batches,
               the single 6-vertex query, the plain path (one
               dsm_process_batch per batch)
  rmat1m_q64   the same graph and stream with the 64 standing queries, autotuned,
               pipelined batches (dsm_process_batches)
  served       the 64-query engine behind a server: log-uniform request sizes
               (1..16,384 records) arriving as a Poisson process
  rmat16m      R-MAT 16M V / 256M E, 32 Zipf(0.8) labels, the dense queries:
               a reduced batch (512 records) and a whole 32,768-record batch,
               pipelined and one batch per call

The oracle's counts are sums over records (main.cu:268-329), so it counts
(query, record share) tasks on the host's threads.
"""
import threading
import time

import numpy as np
import pytest

from dsm_amd import engine as E
from dsm_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def qtuples(w):
    return [(q.n, q.labels, q.eu, q.ev) for q in w.queries]


def oracle_counts(w, batches, parts=8):
    t = time.perf_counter()
    g = O.Graph(w.n, w.labels, w.eu, w.ev)
    say("  oracle graph: %.1f s" % (time.perf_counter() - t))
    oq = [O.Query(q.n, q.labels, q.eu, q.ev) for q in w.queries]
    out = O.run_stream_threaded(g, oq, batches, parts=parts, log=say)
    say("  oracle: %d batches, %d records in %.1f s" % (len(batches), sum(len(b[1]) for b in batches),
                                                       time.perf_counter() - t))
    return out


def check(bi, pos, neg, want):
    assert [int(x) for x in pos] == want["pos"], "batch %d positive counts" % bi
    assert [int(x) for x in neg] == want["neg"], "batch %d negative counts" % bi


@pytest.mark.timeout(600)
def test_rmat1m_single_query_plain_path():
    say("case: rmat1m, 1 query, batches 0-1, one dsm_process_batch per batch")
    w = W.rmat()
    assert w.n == 1_000_000 and len(w.queries) == 1 and w.queries[0].n == 6
    batches = [w.batch(0), w.batch(1)]
    assert all(len(b[1]) == 4096 for b in batches)
    want = oracle_counts(w, batches)
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qtuples(w))
    try:
        for bi, b in enumerate(batches):
            pos, neg = eng.process_batch(*b)
            say("  batch %d: pos %d neg %d" % (bi, int(pos[0]), int(neg[0])))
            check(bi, pos, neg, want[bi])
    finally:
        eng.close()
    assert sum(r["pos"][0] + r["neg"][0] for r in want) > 0


@pytest.mark.timeout(600)
def test_rmat1m_q64_autotuned_pipelined():
    say("case: rmat1m_q64, 64 queries, autotuned, batches 0-2 pipelined")
    w = W.rmat_q64()
    assert w.n == 1_000_000 and len(w.queries) == 64
    batches = w.batches(0, 3)
    want = oracle_counts(w, batches)
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qtuples(w))
    try:
        smp = [w.batch(b) for b in range(4)]
        say("  autotune: %s" % eng.autotune(*(np.concatenate([p[i] for p in smp]) for i in range(4))))
        pos, neg, _ = eng.process_batches(batches)
        for bi in range(len(batches)):
            say("  batch %d: pos %d neg %d" % (bi, int(pos[bi].sum()), int(neg[bi].sum())))
            check(bi, pos[bi], neg[bi], want[bi])
    finally:
        eng.close()


@pytest.mark.timeout(600)
def test_rmat1m_q64_served_poisson_requests():
    say("case: rmat1m_q64 served, Poisson arrivals of log-uniform requests")
    w = W.rmat_q64()
    rng = np.random.default_rng(11)
    hi, rate = 16384, 3e5
    sizes = np.exp(rng.uniform(0, np.log(hi + 1), 12)).astype(np.int64).clip(1, hi)
    # requests cut from the stream in order (batch 0 onwards), at most 24k records
    sizes = sizes[np.cumsum(sizes) <= 24_000]
    assert len(sizes) >= 4
    flat = [np.concatenate([w.batch(b)[i] for b in range(8)]) for i in range(4)]
    cuts = np.concatenate([[0], np.cumsum(sizes)])
    reqs = [tuple(a[cuts[i]:cuts[i + 1]] for a in flat) for i in range(len(sizes))]
    mean = (hi - 1) / np.log(hi)
    gaps = rng.exponential(mean / rate, len(reqs))
    want = oracle_counts(w, reqs)
    eng = E.Engine((w.n, w.labels, w.eu, w.ev), qtuples(w))
    srv = E.Server(eng)
    ids = [None] * len(reqs)
    try:
        def submitter():
            for i, r in enumerate(reqs):
                time.sleep(gaps[i])
                ids[i] = srv.submit(*r)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for i in range(len(reqs)):
            st, pos, neg, info = srv.wait(ids[i])
            assert st == E.DSM_OK, srv.error()
            say("  request %d (%d records, fused %d): pos %d neg %d" % (i, len(reqs[i][1]), info["fused"],
                                                                       int(pos.sum()), int(neg.sum())))
            check(i, pos, neg, want[i])
    finally:
        srv.close()
        eng.close()


@pytest.mark.timeout(1200)
def test_rmat16m_reduced_and_full_batch():
    say("case: rmat16m, 3 dense queries, 512 records then a 32,768-record batch")
    t = time.perf_counter()
    w = W.rmat16m()
    say("  workload: %.1f s" % (time.perf_counter() - t))
    assert w.n == 16_000_000 and len(w.eu) == 256_000_000 - 25_600_000
    b0 = tuple(a[:512] for a in w.batch(0))
    b1 = w.batch(1)
    assert len(b1[1]) == 32768
    batches = [b0, b1]
    want = oracle_counts(w, batches, parts=64)
    g = (w.n, w.labels, w.eu, w.ev)
    eng = E.Engine(g, qtuples(w))
    try:
        pos, neg, _ = eng.process_batches(batches)
        for bi in range(2):
            say("  pipelined batch %d: pos %s neg %s" % (bi, pos[bi].tolist(), neg[bi].tolist()))
            check(bi, pos[bi], neg[bi], want[bi])
    finally:
        eng.close()
    eng = E.Engine(g, qtuples(w))
    try:
        for bi, b in enumerate(batches):
            pos, neg = eng.process_batch(*b)
            say("  one-by-one batch %d: pos %s neg %s" % (bi, pos.tolist(), neg.tolist()))
            check(bi, pos, neg, want[bi])
    finally:
        eng.close()
