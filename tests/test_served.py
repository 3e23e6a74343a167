"""Served stream (dsm_server_*): continuous batching of update requests.

Host tests cover the admission rule (dsm_admit) without a GPU. GPU tests
check that every request's counts are the same whether it is served alone
(dsm_process_batch on an engine of its own, in the same order) or in the mix
of a step, bit-exact, and against the CPU oracle on small inputs; and that
admission by estimated memory rejects a request that does not fit.
"""
import threading
import time

import numpy as np
import pytest

from dsm_amd import engine as E
from oracle import oracle as O


def say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------ host (no GPU)

def test_admit_rule():
    assert E.admit(10, 0, 10) == 1
    assert E.admit(0, 0, 0) == 1
    assert E.admit(5, 5, 10) == 1
    # does not fit beside the request in flight: waits
    assert E.admit(5, 6, 10) == 0
    assert E.admit(10, 1, 10) == 0
    # larger than the whole budget: rejected
    assert E.admit(11, 0, 10) == E.DSM_EADMIT
    assert E.admit(2 ** 63, 0, 2 ** 62) == E.DSM_EADMIT
    assert E.admit(2 ** 62, 2 ** 62, 2 ** 63) == 1


# ---------------------------------------------------------------- GPU tests

SHAPES = {
    "triangle": [(0, 1), (1, 2), (0, 2)],
    "square": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "house": [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4)],
    "path5": [(0, 1), (1, 2), (2, 3), (3, 4)],
}


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


def random_request(n, k, present, rng, nl, conflict_free=False):
    """k records, half inserts and half deletes of present edges. Unless
    conflict_free, a pair can be touched twice and self loops occur (the
    update rule decides)."""
    add, src, dst = [], [], []
    pl = sorted(present)
    used = set()
    while len(src) < k:
        if rng.random() < 0.5 or not pl:
            a, b = map(int, rng.integers(0, n, 2))
            key = (min(a, b), max(a, b))
            if conflict_free and (a == b or key in used or key in present):
                continue
            add.append(1)
            present.add(key)
        else:
            a, b = pl[int(rng.integers(0, len(pl)))]
            key = (a, b)
            if conflict_free and (key in used or key not in present):
                continue
            if rng.random() < 0.5:
                a, b = b, a
            add.append(0)
            present.discard(key)
        used.add(key)
        src.append(a)
        dst.append(b)
    return (np.array(add, np.uint8), np.array(src, np.uint32), np.array(dst, np.uint32),
            rng.integers(0, nl, k).astype(np.uint32))


def sizes_log_uniform(rng, count, hi):
    return [int(x) for x in np.exp(rng.uniform(0, np.log(hi + 1), count)).astype(int).clip(1, hi)]


@pytest.mark.gpu
def test_served_equals_alone_in_order():
    """Requests of log-uniform sizes arrive in three waves (the later ones
    while earlier steps run on the GPU); each request's counts equal those of
    the same request served alone, in submission order, by a plain engine."""
    say("case: served mix == alone, 48 requests in 3 waves")
    rng = np.random.default_rng(1201)
    n, m, nl = 2000, 16000, 3
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 2]), qtuple("square", [0, 0, 1, 1]),
               qtuple("house", [0, 1, 2, 0, 1]), qtuple("path5", [1, 1, 0, 2, 2])]
    present = set(zip(eu.tolist(), ev.tolist()))
    reqs = [random_request(n, k, present, rng, nl) for k in sizes_log_uniform(rng, 48, 2000)]
    seq = E.Engine((n, labels, eu, ev), queries)
    srv_eng = E.Engine((n, labels, eu, ev), queries)
    srv = E.Server(srv_eng)
    try:
        ids = []
        for wave in range(3):
            for r in reqs[wave * 16:(wave + 1) * 16]:
                ids.append(srv.submit(*r))
            time.sleep(0.05)
        got = [srv.wait(i) for i in ids]
        say("  served; %d of %d requests fused with their predecessor"
            % (sum(g[3]["fused"] for g in got), len(got)))
        for i, r in enumerate(reqs):
            pa, na = seq.process_batch(*r)
            st, ps, ns, info = got[i]
            assert st == E.DSM_OK, srv.error()
            assert ps.tolist() == pa.tolist(), "request %d positives" % i
            assert ns.tolist() == na.tolist(), "request %d negatives" % i
            assert info["t_submit"] <= info["t_start"] <= info["t_done"]
        assert any(g[3]["fused"] for g in got), "no request shared a launch with its predecessor"
    finally:
        srv.close()
        srv_eng.close()
        seq.close()


@pytest.mark.gpu
def test_served_matches_oracle():
    """Served counts against the CPU oracle on a small stream."""
    say("case: served == oracle, 20 requests")
    rng = np.random.default_rng(1202)
    n, m, nl = 300, 2400, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 0, 1]), qtuple("house", [0, 1, 0, 1, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    reqs = [random_request(n, k, present, rng, nl, conflict_free=True) for k in sizes_log_uniform(rng, 20, 200)]
    g = O.Graph(n, labels, eu, ev)
    want = O.run_stream(g, [O.Query(*q) for q in queries], reqs, relabel=True)
    eng = E.Engine((n, labels, eu, ev), queries)
    srv = E.Server(eng)
    try:
        ids = [srv.submit(*r) for r in reqs]
        for i, rid in enumerate(ids):
            st, ps, ns, _ = srv.wait(rid)
            assert st == E.DSM_OK, srv.error()
            assert ps.tolist() == want[i]["pos"] and ns.tolist() == want[i]["neg"], "request %d" % i
    finally:
        srv.close()
        eng.close()


@pytest.mark.gpu
def test_admission_rejects_oversize_request():
    """With a budget of 300 records' estimated memory, a 1,000-record request is
    rejected (DSM_EADMIT) and leaves the graph alone; a request that does not
    fit beside the one in flight waits for it (never shares its launch); the
    rest equal the plain path without the rejected ones."""
    say("case: admission by estimated memory")
    rng = np.random.default_rng(1203)
    n, m, nl = 500, 4000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 1]), qtuple("square", [0, 1, 0, 1])]
    present = set(zip(eu.tolist(), ev.tolist()))
    sizes = [50, 120, 1000, 80, 250, 100, 1000, 10, 60]
    reqs = [random_request(n, k, present, rng, nl) for k in sizes]
    seq = E.Engine((n, labels, eu, ev), queries)
    eng = E.Engine((n, labels, eu, ev), queries)
    srv = None
    try:
        est = lambda k: int(E.lib().dsm_estimate_request_bytes(eng._h, k))  # noqa: E731
        budget = est(300)
        assert est(1000) > budget and est(80) + est(250) > budget and est(250) + est(100) > budget
        srv = E.Server(eng, budget_bytes=budget)
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i) for i in ids]
        for (st, ps, ns, info), k in zip(got, sizes):
            if k == 1000:
                assert st == E.DSM_EADMIT
                continue
            assert st == E.DSM_OK, srv.error()
            if k in (250, 100):
                assert info["fused"] == 0, "request of %d records shared a launch over budget" % k
        for (st, ps, ns, _), r, k in zip(got, reqs, sizes):
            if k == 1000:
                continue
            pa, na = seq.process_batch(*r)
            assert ps.tolist() == pa.tolist() and ns.tolist() == na.tolist()
    finally:
        if srv is not None:
            srv.close()
        eng.close()
        seq.close()


@pytest.mark.gpu
def test_concurrent_submitters():
    """Two threads submit; the server's order is its submission order, which
    the request ids record. Replaying the requests in id order on a plain
    engine gives the same counts."""
    say("case: two submitting threads")
    rng = np.random.default_rng(1204)
    n, m, nl = 800, 6000, 2
    labels, eu, ev = random_graph(n, m, nl, rng)
    queries = [qtuple("triangle", [0, 1, 0]), qtuple("path5", [0, 1, 0, 1, 0])]
    present = set(zip(eu.tolist(), ev.tolist()))
    reqs = [random_request(n, k, present, rng, nl) for k in sizes_log_uniform(rng, 40, 500)]
    eng = E.Engine((n, labels, eu, ev), queries)
    seq = E.Engine((n, labels, eu, ev), queries)
    srv = E.Server(eng)
    ids = {}

    def submitter(part):
        for i in part:
            ids[i] = srv.submit(*reqs[i])
            time.sleep(0.001)

    try:
        th = [threading.Thread(target=submitter, args=(range(j, 40, 2),)) for j in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        res = {i: srv.wait(rid) for i, rid in ids.items()}
        for i in sorted(ids, key=lambda i: ids[i]):
            st, ps, ns, _ = res[i]
            assert st == E.DSM_OK, srv.error()
            pa, na = seq.process_batch(*reqs[i])
            assert ps.tolist() == pa.tolist() and ns.tolist() == na.tolist(), "request %d" % i
    finally:
        srv.close()
        eng.close()
        seq.close()


@pytest.mark.gpu
def test_server_refuses_tuple_engine():
    rng = np.random.default_rng(1205)
    labels, eu, ev = random_graph(50, 100, 2, rng)
    eng = E.Engine((50, labels, eu, ev), [qtuple("triangle", [0, 0, 1])], tuple_capacity=16)
    try:
        with pytest.raises(E.DsmError) as ei:
            E.Server(eng)
        assert ei.value.code == E.DSM_EINVAL
    finally:
        eng.close()
