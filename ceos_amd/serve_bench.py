"""Serving benchmark: open-loop Poisson query arrivals into the continuous-batching
server (ceos_amd.CEOs.Server) over a resident index, reporting achieved QPS and
p50/p99 latency (arrival to results ready) per offered rate.

  python -m ceos_amd.serve_bench --workload serve_10m --rates 20000,50000,100000
  python -m ceos_amd.serve_bench --workload stream_5m --rates 50000
  python -m ceos_amd.serve_bench --workload gist_like --rates 50000,200000,400000

serve_10m: 10,000,000 x 256 points (1,000-centre mixture), D=1024, R=4, m=2000;
requests with k in {10,100}, n_probed_vectors 40, n_cand U[500,5000].
gist_like: 1,000,000 x 960 points, D=1024, R=8, m=500; requests with the uneven batch's
per-query k in {1,10,100}, n_probed_vectors U[10,80] and n_cand U[200,5000].
stream_5m: 5,000,000 x 384 points, D=512, R=8, m=1000 in a streamCEOs index; requests
with k=10, n_probed_vectors 40, n_cand 1000, while every 100 ms a batch of 10,000 new
points is inserted and 0.5% of the live ids are deleted at random.

A sample of the served results is checked against the same queries searched alone.
Prints one JSON line per rate on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import threading
import time

import numpy as np


def log(msg):
    print(f"[serve] {msg}", file=sys.stderr, flush=True)


def slowest(trace, n=3):
    """The n batches that took longest from launch to results, from the server's batch trace
    (launch time s, enqueue ms, done ms, oldest request's wait ms, queries, slot, and the
    enqueue's steps in ms: rows to staging, inputs to device, plan launch, results copy and
    event, server lock)."""
    if len(trace) == 0:
        return []
    return [{"launch_s": round(float(b[0]), 4), "enqueue_ms": round(float(b[1]), 3), "done_ms": round(float(b[2]), 3),
             "oldest_wait_ms": round(float(b[3]), 3), "queries": int(b[4]),
             "enqueue_steps_ms": [round(float(x), 3) for x in b[6:11]]}
            for b in trace[np.argsort(-trace[:, 2])[:n]]]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="serve_10m", choices=["serve_10m", "stream_5m", "gist_like"])
    ap.add_argument("--rates", default="20000,50000,100000")
    ap.add_argument("--seconds", type=float, default=2.0, help="arrival window per rate")
    ap.add_argument("--n", type=int, default=0, help="override number of points (testing only)")
    ap.add_argument("--max-batch", type=int, default=4096)
    ap.add_argument("--inflight", type=int, default=2,
                    help="batch slots (more slots overlap large batches' tails; fewer keep small batches fast)")
    ap.add_argument("--warmup-seconds", type=float, default=0.5,
                    help="untimed arrivals at the first rate before the measured rates")
    ap.add_argument("--check", type=int, default=64, help="served results checked against alone")
    args = ap.parse_args()

    from ceos_amd import CEOs
    from ceos_amd.workloads import WORKLOADS, make_points, make_queries, make_query_params

    w = WORKLOADS[args.workload]
    n = args.n or w.n
    t = time.time()
    X = make_points(w, n)
    log(f"{w.name}: {n} x {w.d} points generated in {time.time() - t:.1f}s")
    stream = args.workload == "stream_5m"
    t = time.time()
    if stream:
        index = CEOs.streamCEOs(w.d)
        index.setIndexParam(w.D, w.R, w.m, -1, w.seed)
        index.build(X)
    else:
        index = CEOs.CEOs(n, w.d)
        index.setIndexParam(w.D, w.R, w.m, -1, w.seed)
        index.build_coCEOs_Est(X)
    log(f"index built in {time.time() - t:.1f}s")
    index.n_probed_vectors, index.n_probed_points, index.n_cand = w.s, w.m, w.c
    uneven = bool(w.uneven)
    if stream:
        k_max, s_max, c_max = 10, w.s, w.c
    elif uneven:
        k_max, s_max, c_max = 100, 80, 5000
    else:
        k_max, s_max, c_max = 100, w.s, 5000
    srv = CEOs.Server(index, max_batch=args.max_batch, k_max=k_max, s_max=s_max, c_max=c_max, inflight=args.inflight)
    rng = np.random.default_rng(3)
    upd_rng = np.random.default_rng(4)
    next_row = [0]

    rates = [int(r) for r in args.rates.split(",")]
    if args.warmup_seconds > 0 and rates:  # the server and the GPU at load first, untimed
        nw = max(1, int(rates[0] * args.warmup_seconds))
        Qw = make_queries(w, nw, seed=7)
        if uneven:
            kw, sw, cw = make_query_params(w, nw, seed=8)
        else:
            kw = np.full(nw, min(10, k_max), np.int32)
            sw = np.full(nw, w.s, np.int32)
            cw = np.full(nw, min(w.c, c_max), np.int32)
        srv.submit(Qw, kw, sw, cw, np.cumsum(rng.exponential(1.0 / rates[0], nw)))
        srv.drain()
        del Qw

    for rate in rates:
        nreq = int(rate * args.seconds)
        Q = make_queries(w, nreq, seed=1 + rate)
        if stream:
            k = np.full(nreq, 10, np.int32)
            c = np.full(nreq, w.c, np.int32)
            s = np.full(nreq, w.s, np.int32)
        elif uneven:
            k, s, c = make_query_params(w, nreq, seed=2 + rate)
        else:
            k = rng.choice(np.array([10, 100], np.int32), nreq).astype(np.int32)
            c = rng.integers(500, 5001, nreq).astype(np.int32)
            s = np.full(nreq, w.s, np.int32)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
        srv.reset_stats()
        stop = threading.Event()
        n_updates = [0]

        def updater():  # streaming: every 100 ms insert 10,000 points and delete 0.5% of live ids
            t0 = time.perf_counter()
            while not stop.is_set():
                time.sleep(max(0.0, 0.1 * (n_updates[0] + 1) - (time.perf_counter() - t0)))
                if stop.is_set():
                    break
                a = next_row[0] % n
                new = X[a:a + 10000]
                next_row[0] += 10000
                srv.update(new, 0)
                live = index.size()  # ids below the current count stay valid after the insert
                dead = upd_rng.choice(live, max(1, live // 200), replace=False).astype(np.int32)
                srv.delete_ids(dead)
                n_updates[0] += 1

        th = threading.Thread(target=updater) if stream else None
        t0 = srv.submit(Q, k, s, c, arrivals)
        if th:
            th.start()
        got = {}
        chk = set(rng.choice(nreq, min(args.check, nreq), replace=False).tolist())
        for i in range(nreq):
            r = srv.result(t0 + i)
            if i in chk:
                got[i] = r
        if th:
            stop.set()
            th.join()
            srv.drain()
        st = srv.stats()
        # served results equal the same query searched alone (static index only: a
        # stream index has moved on since)
        match = None
        if not stream:
            ok = 0
            for i, (ids, ips, _) in got.items():
                a, p, _ = index.search_batched(Q[i:i + 1], k[i:i + 1], s[i:i + 1], c[i:i + 1])
                ok += int(np.array_equal(a, ids) and np.array_equal(p, ips))
            match = ok / max(1, len(got))
        out = {"workload": w.name, "offered_qps": rate, "requests": nreq, "achieved_qps": round(st["qps"], 1),
               "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
               "batches": st["batches"], "mean_batch": round(st["mean_batch"], 1),
               "inflight_slots": st["inflight_slots"], "device_bytes_slots": st["device_bytes"],
               "alone_match": match, "updates": n_updates[0] if stream else None,
               "insert_10k_mean_ms": round(st["insert_mean_ms"], 3) if stream else None,
               "insert_10k_max_ms": round(st["insert_max_ms"], 3) if stream else None,
               "delete_mean_ms": round(st["delete_mean_ms"], 3) if stream else None,
               "delete_max_ms": round(st["delete_max_ms"], 3) if stream else None,
               # where the time goes: a request's wait before its batch is launched, and a
               # batch's time from launch to results in host memory
               "wait_ms": {"p50": round(st["wait_p50_ms"], 3), "p99": round(st["wait_p99_ms"], 3),
                           "max": round(st["wait_max_ms"], 3)},
               "batch_service_ms": {"p50": round(st["service_p50_ms"], 3), "p99": round(st["service_p99_ms"], 3),
                                    "max": round(st["service_max_ms"], 3)},
               "slowest_batches": slowest(srv.batches())}
        log(json.dumps(out))
        print(json.dumps(out), flush=True)
    srv.close()


if __name__ == "__main__":
    main()
