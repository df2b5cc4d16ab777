"""Serving at the 10,000,000 x 256 configuration (1,000-centre Gaussian mixture, sigma
0.25, LogNormal(0,0.5) norms; D=1024, R=4, m=2000, n_probed_vectors 40), built on the
device:
  * 32 lists (8 columns in each of two repeats, close and far) equal the
    oracle's build of those columns over all 10,000,000 points;
  * 1,000 requests served under Poisson arrivals (k in {10,100}, n_cand U[500,5000]):
    each result equals the same query searched alone, bit for bit, and the oracle's
    search on the device-built buckets (at least 99.9% identical top-k, the rest only
    reordered among points within 1e-5 of the k-th inner product; logged).
"""
import sys
import time

import numpy as np
import pytest

from tests._parity import bits, compare_list, topk_verdict

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[serve_10m] {msg}", file=sys.stderr, flush=True)


def test_serve_10m(CEOs, oracle):
    from ceos_amd.workloads import WORKLOADS, make_points, make_queries

    w = WORKLOADS["serve_10m"]
    t = time.time()
    X = make_points(w)
    say(f"data {X.shape} in {time.time() - t:.1f}s")
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(w.D, w.R, w.m, -1, w.seed)
    t = time.time()
    g.build_coCEOs_Est(X)
    say(f"device build (with the host copy) in {time.time() - t:.2f}s")
    g.n_probed_points = w.m
    cid, cv, fid, fv = g.get_buckets()

    # lists of 16 columns of two repeats against the oracle's build of those columns
    rng = np.random.default_rng(8)
    cols = np.concatenate([r * w.D + rng.choice(w.D, 8, replace=False) for r in (0, 3)])
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(w.D, w.R, w.m, 16, w.seed)
    t = time.time()
    oc, ocv, of, ofv = o.bucket_columns(X, cols)
    say(f"oracle lists of {len(cols)} columns in {time.time() - t:.1f}s")
    mism = 0
    for j, col in enumerate(cols):
        mism += compare_list(cid[col], cv[col], oc[j], ocv[j])
        mism += compare_list(fid[col], fv[col], of[j], ofv[j])
    say(f"buckets: {2 * len(cols)} lists x {w.m} identical ({mism} ids placed differently in equal-value runs)")
    del oc, ocv, of, ofv

    # served requests
    n = 1000
    Q = make_queries(w, n)
    prng = np.random.default_rng(9)
    k = prng.choice(np.array([10, 100], np.int32), n).astype(np.int32)
    s = np.full(n, 40, np.int32)
    c = prng.integers(500, 5001, n).astype(np.int32)
    arrivals = np.cumsum(prng.exponential(1.0 / 20000.0, n))
    srv = CEOs.Server(g, max_batch=1024, k_max=100, s_max=40, c_max=5000, inflight=2)
    t0 = srv.submit(Q, k, s, c, arrivals)
    got = [srv.result(t0 + i) for i in range(n)]
    st = srv.stats()
    srv.close()
    say(f"served {st['completed']} requests: p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
        f"mean batch {st['mean_batch']:.1f}")
    assert st["completed"] == n and st["batches"] > 1
    for i in range(n):
        a, p, _ = g.search_batched(Q[i:i + 1], k[i:i + 1], s[i:i + 1], c[i:i + 1])
        assert np.array_equal(got[i][0], a), i
        assert np.array_equal(bits(got[i][1]), bits(p)), i

    o.load_index(X, cid, cv, fid, fv)
    o.n_probed_points = w.m
    t = time.time()
    oi, op = o.search_coCEOs_Est(Q, 0, ks=k, ss=s, cs=c)
    say(f"oracle search of {n} requests in {time.time() - t:.1f}s")
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(k)
    verdict = {"same": 0, "tie": 0, "diff": 0}
    for i in range(n):
        v = topk_verdict(Q[i], lambda r: X[r], got[i][0], got[i][1], oi[off[i]:off[i + 1]], op[off[i]:off[i + 1]])
        verdict[v] += 1
        if v != "same":
            say(f"request {i} (k={k[i]}, c={c[i]}): {v}")
    say(f"oracle top-k on {n} served requests: {verdict}")
    assert verdict["diff"] == 0
    assert verdict["same"] >= 0.999 * n
