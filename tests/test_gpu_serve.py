"""Serving engine: continuous batching over a resident index. Every result equals the
same query searched alone; requests arriving while batches run are admitted into later
batches (several batches, more than one query per batch); the HBM budget bounds the
batch slots; a streamCEOs server applies updates between batches."""
import threading

import numpy as np
import pytest

from tests._parity import bits

pytestmark = pytest.mark.gpu


def _rand(n, d, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return (g * rng.lognormal(0, sigma, (n, 1))).astype(np.float32)


@pytest.fixture(scope="module")
def index(CEOs):
    X = _rand(40000, 64, 51, 0.7)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(256, 4, 150, -1, 42)
    g.build_coCEOs_Est(X)
    return X, g


def _alone(g, q, k, s, c):
    ids, ips, _ = g.search_batched(q[None, :], np.array([k], np.int32), np.array([s], np.int32),
                                   np.array([c], np.int32))
    return ids, ips


def test_serve_large_batches(CEOs, index):
    """3,000 requests at once into slots of 2,048: a backlog shared among the free slots
    gives batches of 512 requests or more. Every sampled result equals the query
    searched alone."""
    X, g = index
    n = 3000
    Q = _rand(n, 64, 58)
    rng = np.random.default_rng(59)
    k = rng.choice(np.array([1, 10, 100], np.int32), n).astype(np.int32)
    s = rng.integers(10, 81, n).astype(np.int32)
    c = rng.integers(50, 3001, n).astype(np.int32)
    srv = CEOs.Server(g, max_batch=2048, k_max=100, s_max=80, c_max=3000, inflight=2)
    t0 = srv.submit(Q, k, s, c)
    got = [srv.result(t0 + i) for i in range(n)]
    trace = srv.batches()
    srv.close()
    assert trace[:, 4].max() >= 512
    for i in rng.choice(n, 200, replace=False):
        ids, ips, _ = got[i]
        a, p = _alone(g, Q[i], int(k[i]), int(s[i]), int(c[i]))
        assert np.array_equal(ids, a), i
        assert np.array_equal(bits(ips), bits(p)), i


def test_serve_backlog_shared_among_free_slots(CEOs, index):
    """4,000 requests arriving at once with four free slots of 4,096: the backlog is
    shared among the slots (a quarter each, at least 256) instead of one slot taking
    all of it while the device waits for its copy."""
    X, g = index
    n = 4000
    Q = _rand(n, 64, 63)
    k = np.full(n, 10, np.int32)
    s = np.full(n, 20, np.int32)
    c = np.full(n, 200, np.int32)
    srv = CEOs.Server(g, max_batch=4096, k_max=10, s_max=20, c_max=200, inflight=4)
    t0 = srv.submit(Q, k, s, c)
    srv.drain()
    trace = srv.batches()
    got = [srv.result(t0 + i) for i in (0, n // 2, n - 1)]
    srv.close()
    assert trace[:, 4].sum() == n
    assert trace[0, 4] <= -(-n // 4)  # the first batch took its share, not the backlog
    assert len(trace) >= 4
    for i, (ids, ips, _) in zip((0, n // 2, n - 1), got):
        a, p = _alone(g, Q[i], 10, 20, 200)
        assert np.array_equal(ids, a), i
        assert np.array_equal(bits(ips), bits(p)), i


def test_serve_open_loop_each_result_equals_alone(CEOs, index):
    X, g = index
    n = 600
    Q = _rand(n, 64, 52)
    rng = np.random.default_rng(53)
    k = rng.choice(np.array([1, 10, 100], np.int32), n).astype(np.int32)
    s = rng.integers(10, 81, n).astype(np.int32)
    c = rng.integers(50, 3001, n).astype(np.int32)
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, n))  # Poisson, 20k queries/s
    srv = CEOs.Server(g, max_batch=64, k_max=100, s_max=80, c_max=3000, inflight=3)
    t0 = srv.submit(Q, k, s, c, arrivals)
    got = [srv.result(t0 + i) for i in range(n)]
    st = srv.stats()
    srv.close()
    assert st["completed"] == n and st["submitted"] == n
    assert st["batches"] > 1 and st["mean_batch"] > 1.0  # continuous batching happened
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    for i in range(n):
        ids, ips, lat = got[i]
        a, p = _alone(g, Q[i], int(k[i]), int(s[i]), int(c[i]))
        assert np.array_equal(ids, a), i
        assert np.array_equal(bits(ips), bits(p)), i
        assert lat > 0


def test_serve_concurrent_clients(CEOs, index):
    X, g = index
    srv = CEOs.Server(g, max_batch=32, k_max=10, s_max=40, c_max=500, inflight=2)
    out = {}

    def client(cid):
        Q = _rand(100, 64, 60 + cid)
        k = np.full(100, 10, np.int32)
        s = np.full(100, 20 + cid, np.int32)
        c = np.full(100, 300, np.int32)
        tickets = []
        for i in range(100):  # one request at a time, closed loop for half of them
            tickets.append(srv.submit(Q[i:i + 1], k[i:i + 1], s[i:i + 1], c[i:i + 1]))
            if i % 2 == 0:
                out[(cid, i)] = (Q[i], int(s[i]), srv.result(tickets[-1]))
        for i in range(100):
            if i % 2 == 1:
                out[(cid, i)] = (Q[i], int(s[i]), srv.result(tickets[i]))

    th = [threading.Thread(target=client, args=(cid,)) for cid in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    assert len(out) == 400
    for (cid, i), (q, s_, (ids, ips, _)) in out.items():
        a, p = _alone(g, q, 10, s_, 300)
        assert np.array_equal(ids, a) and np.array_equal(bits(ips), bits(p)), (cid, i)


def test_serve_hbm_budget_and_bounds(CEOs, index):
    X, g = index
    with pytest.raises(RuntimeError):
        CEOs.Server(g, max_batch=64, k_max=10, s_max=40, c_max=500, inflight=2, hbm_budget=1000)
    big = CEOs.Server(g, max_batch=64, k_max=10, s_max=40, c_max=500, inflight=4)
    per = big.stats()["device_bytes"] // 4
    big.close()
    one = CEOs.Server(g, max_batch=64, k_max=10, s_max=40, c_max=500, inflight=4, hbm_budget=per + per // 2)
    assert one.stats()["inflight_slots"] == 1
    with pytest.raises(ValueError):
        one.submit(_rand(1, 64, 70), np.array([11], np.int32), np.array([5], np.int32), np.array([5], np.int32))
    Q = _rand(50, 64, 71)
    t0 = one.submit(Q, np.full(50, 10, np.int32), np.full(50, 30, np.int32), np.full(50, 200, np.int32))
    for i in range(50):
        ids, ips, _ = one.result(t0 + i)
        a, p = _alone(g, Q[i], 10, 30, 200)
        assert np.array_equal(ids, a)
    one.close()


def test_stream_server_updates_between_batches(CEOs):
    d, D, R, m = 32, 128, 2, 40
    X = _rand(8000, d, 72)
    Q = _rand(40, d, 73)
    st = CEOs.streamCEOs(d)
    st.setIndexParam(D, R, m, -1, 42)
    st.build(X[:4000])
    st.n_probed_vectors, st.n_probed_points, st.n_cand = 16, m, 150
    srv = CEOs.Server(st, max_batch=16, k_max=10, s_max=16, c_max=150, inflight=2)
    k = np.full(40, 10, np.int32)
    s = np.full(40, 16, np.int32)
    c = np.full(40, 150, np.int32)
    pos = 4000
    for step in range(3):
        t0 = srv.submit(Q, k, s, c)
        res = [srv.result(t0 + i) for i in range(40)]
        ei, ep = st.estimate_search(Q, 10)
        for i in range(40):
            assert np.array_equal(res[i][0], ei[i]), (step, i)
            assert np.array_equal(bits(res[i][1]), bits(ep[i])), (step, i)
        srv.update(X[pos:pos + 1000], 700)
        pos += 1000
        srv.delete_ids(np.array([1, 2, 3], np.int32))
        srv.drain()
        assert st.size() == 4000 + (step + 1) * 300
    srv.close()


def test_serve_batch_trace(CEOs, index):
    """The per-batch trace (in the order batches completed) covers every served request,
    with the wait of each batch's oldest request and its service time."""
    X, g = index
    n = 400
    Q = _rand(n, 64, 61)
    k = np.full(n, 10, np.int32)
    s = np.full(n, 20, np.int32)
    c = np.full(n, 500, np.int32)
    srv = CEOs.Server(g, max_batch=64, k_max=10, s_max=20, c_max=500, inflight=2)
    srv.submit(Q, k, s, c, np.cumsum(np.random.default_rng(62).exponential(1.0 / 50000.0, n)))
    srv.drain()
    B = srv.batches()
    st = srv.stats()
    assert B.shape == (st["batches"], 11)
    assert int(B[:, 4].sum()) == n and (B[:, 4] <= 64).all()
    assert np.all(B[:, 0] >= 0) and B[:, 0].max() < st["elapsed_s"] + 1.0   # launched within the run
    assert np.all(B[:, 2] >= B[:, 1]) and np.all(B[:, 3] >= 0)
    assert set(B[:, 5].astype(int).tolist()) <= {0, 1}
    # the enqueue's steps add up to the enqueue time
    assert np.all(B[:, 6:] >= 0) and np.allclose(B[:, 6:].sum(1), B[:, 1], rtol=1e-6, atol=1e-6)
    srv.reset_stats()
    assert srv.batches().shape == (0, 11)
    srv.close()
