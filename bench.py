#!/usr/bin/env python3
"""coCEOs search benchmark on one MI355X.

A step is one batched search of the workload's whole query set, replayed from a
captured plan (hipGraph of query projection + probe selection + the fused persistent
search, launch shape autotuned on the first run), with the points, buckets, queries
and per-query (k, n_probed_vectors, n_cand) already resident in HBM. Reported beside
it, from the same run:
  * plain per-query GPU path: the same index searched one query per call,
    synchronising after each (the drop-in search used query by query);
  * served: the continuous-batching server over the same index for about a second of
    open-loop Poisson arrivals (60% of the batched rate by default), achieved QPS and
    p50/p99 latency from arrival to results in host memory;
  * build throughput (Mpoints/s) of the device build from HBM-resident points;
  * recall@k against exact MIPS (FP32 matmul on the device);
  * roofline of the fused search kernel: algorithmic HBM bytes per launch over its
    average duration, from HIP events the library records around it on its stream,
    and the measured HBM traffic of that kernel (rocprofv3 counters, from the
    committed profile of the same workload);
  * the CPU baseline: the oracle (a C++/OpenMP restatement of the reference's
    search_coCEOs_Est) timed on the host on the same index and queries.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10>=0.9, served+batched vs plain GPU; p50/p99 ms; build Mpts/s"
HBM_PEAK_GBS = 8000.0
# FP32 adds per second without packed FP32 instructions: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# (the 157.3 TFLOP/s FP32 vector peak counts a packed FMA as 4 flops per lane)
VALU_ADD_PEAK_T = 39.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """nproc, the cores this process may run on, OMP_NUM_THREADS, CPU model and RAM."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    info["omp_num_threads"] = int(omp) if omp.isdigit() else None
    info["cpu_model"] = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    info["ram_gb"] = round(int(line.split()[1]) / 2 ** 20, 1)
                    break
    except OSError:
        pass
    return info


def cpu_threads(args, info):
    """Every core this process may use, capped by OMP_NUM_THREADS when set (the host's
    share of a shared machine)."""
    if args.cpu_threads:
        return args.cpu_threads
    t = info.get("affinity") or 1
    if info.get("omp_num_threads"):
        t = min(t, info["omp_num_threads"])
    return t


def stream_update_baseline(O, timing, threads, n0, reps=3):
    """streamCEOs.update of 10,000 new rows with n_delPoints = 10,000 at the streaming
    configuration's parameters (d=384, D=512, R=8, m=1000), on an index of n0 points: the
    reference's update (src/streamCEOs.cpp:174-369, as the oracle restates it) on the host
    beside the device's, median of `reps` updates each. The cost of an update does not
    grow with the index (new rows projected, each list merged with them)."""
    import ctypes

    from ceos_amd import capi
    from ceos_amd.workloads import WORKLOADS, make_points

    w = WORKLOADS["stream_5m"]
    X0 = make_points(w, n0, seed=0)
    news = [make_points(w, 10000, seed=40 + i) for i in range(reps)]
    o = O.OracleStream(w.d, timing=timing)
    o.setIndexParam(w.D, w.R, w.m, threads, w.seed)
    o.build(X0)
    cpu = []
    for Xn in news:
        t0 = time.perf_counter()
        o.update(Xn, 10000)
        cpu.append((time.perf_counter() - t0) * 1e3)
    del o
    L = capi.lib()
    h = ctypes.c_void_p()
    capi.check(L.ceos_stream_create(w.d, ctypes.byref(h)))
    gpu = []
    try:
        capi.check(L.ceos_stream_set_index_param(h, w.D, w.R, w.m, -1, w.seed))
        capi.check(L.ceos_stream_build(h, X0.ctypes.data, n0, w.d))
        for Xn in news:
            t0 = time.perf_counter()
            capi.check(L.ceos_stream_update(h, Xn.ctypes.data, len(Xn), w.d, 10000))
            gpu.append((time.perf_counter() - t0) * 1e3)
    finally:
        L.ceos_stream_destroy(h)
    out = {"cpu_ms_per_10k": round(float(np.median(cpu)), 3), "gpu_ms_per_10k": round(float(np.median(gpu)), 3),
           "sample": f"index of {n0} points x {w.d}, D={w.D}, R={w.R}, m={w.m}; +10,000 rows, n_delPoints 10,000; "
                     f"median of {reps} (host: {threads} threads; device: host rows copied in)"}
    log(f"stream update per 10K: cpu {out['cpu_ms_per_10k']} ms, gpu {out['gpu_ms_per_10k']} ms")
    return out


def exact_topk(torch, dX, dQ, k, chunk=256):
    out = []
    for a in range(0, dQ.shape[0], chunk):
        s = dQ[a:a + chunk] @ dX.T
        out.append(torch.topk(s, k, dim=1).indices.to(torch.int32).cpu())
    return torch.cat(out).numpy()


def run_mode(args, w, X, Q, dX, dQ, n, d, nq, torch, CEOs):
    """CEOs-Est (`build`/`search`) or CEOs-Hash (`build_CEOs_Hash`/`search_CEOs_Hash`) on
    the workload's points and queries with its n_probed_vectors / n_probed_points / n_cand
    for every query. A step is one search call over all queries (host entry point: the
    queries and results cross PCIe, about 1 MB each way, inside the timed region)."""
    from oracle import oracle as O

    mode = args.mode
    k = 10
    s, mp, c = w.s, (w.m if mode == "est" else args.hash_points), w.c
    g = CEOs.CEOs(n, d)
    g.setIndexParam(w.D, w.R, w.m, -1, w.seed)
    g.rerank_bounds = not args.no_rerank_bounds
    build = g.build if mode == "est" else g.build_CEOs_Hash
    build(X)
    t0 = time.perf_counter()
    build(X)  # the second build (host points, copied in as the reference does)
    build_s = time.perf_counter() - t0
    g.n_probed_vectors, g.n_probed_points, g.n_cand = s, mp, c
    search = g.search if mode == "est" else g.search_CEOs_Hash
    for _ in range(args.warmup):
        ids, _ = search(Q, k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ids, _ = search(Q, k)
    step_s = (time.perf_counter() - t0) / args.steps
    qps = nq / step_s
    log(f"{mode}: build {build_s * 1e3:.1f} ms, search {step_s * 1e3:.2f} ms/step, {qps:,.0f} QPS")
    exact = exact_topk(torch, dX, dQ, k)
    rec = float(np.mean([len(np.intersect1d(exact[q], ids[q])) / k for q in range(nq)]))
    # CPU baseline: the oracle's search on a sample of the queries, same parameters
    o = O.OracleCEOs(n, d, timing=O.lib(timing=True) is not None)
    info = host_info()
    threads = cpu_threads(args, info)
    o.setIndexParam(w.D, w.R, w.m, threads, w.seed)
    (o.build if mode == "est" else o.build_CEOs_Hash)(X)
    o.n_probed_vectors, o.n_probed_points, o.n_cand = s, mp, c
    S = min(nq, 64 if mode == "est" else 1000)
    if mode == "est":
        oi, _ = o.search(Q[:S], k)
        distinct = None
    else:
        oi, _, distinct = o.search_CEOs_Hash(Q[:S], k, stats=True)
    reps, t0 = 1, time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        (o.search if mode == "est" else o.search_CEOs_Hash)(Q[:S], k)
        reps += 1
    cpu_qps = S * (reps - 1) / max(1e-9, time.perf_counter() - t0) if reps > 1 else None
    match = float(np.mean([np.array_equal(oi[q], ids[q]) for q in range(S)]))
    # algorithmic bytes per query: the query, the probed columns (CEOs-Est: s columns of n
    # estimates) or lists (CEOs-Hash: s x n_probed_points ids), the reranked rows, the output
    if mode == "est":
        per_q = 4.0 * d + 4.0 * s * n + 4.0 * d * min(c, n) + 8 * k
    else:
        # with rerank bounds (4 k <= distinct candidates): (d + 16) per candidate plus the
        # FP32 rows of the k best
        nd = float(np.mean(distinct))
        rows_b = (d + 16) * nd + 4.0 * d * k if (g.rerank_bounds and 4 * k <= nd) else 4.0 * d * nd
        per_q = 4.0 * d + 8.0 * s * mp + rows_b + 8 * k
    achieved = per_q * nq / step_s / 1e9
    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (numpy PCG64 seeds 0/1; index sign seed 42)",
        "config": {"workload": w.name, "mode": "CEOs-Est" if mode == "est" else "CEOs-Hash", "n": n, "d": d,
                   "queries": nq, "n_proj": w.D, "n_repeats": w.R, "top_m": w.m, "n_probed_vectors": s,
                   "n_probed_points": mp, "n_cand": c, "k": k},
        "recall_at_10": round(rec, 4), "build_ms": round(build_s * 1e3, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_est_search" if mode == "est" else "k_search_fused (hash mode)",
                     "bytes_per_query": round(per_q), "timed": "host entry point (PCIe copies of Q and results included)"},
        "cpu_baseline": {"value": round(cpu_qps, 1) if cpu_qps else None, "unit": "queries/s", "cores": threads,
                         "kind": "port", "sample": f"{S} queries, {reps - 1} passes", "topk_match_vs_gpu": match,
                         "host": info},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    # default: the 1,000,000 x 960 uneven batch, the configuration the headline target
    # (>= 20x the plain per-query path at >= 60% of the HBM roofline) is stated on
    ap.add_argument("--workload", default="gist_like")
    ap.add_argument("--n", type=int, default=0, help="override number of points (testing only)")
    ap.add_argument("--nq", type=int, default=0, help="override number of queries (testing only)")
    ap.add_argument("--plain-queries", type=int, default=300)
    ap.add_argument("--cpu-queries", type=int, default=0, help="CPU baseline sample (default: all queries)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="repeat the CPU sample for this long")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-build-points", type=int, default=100_000, help="CPU build baseline sample")
    ap.add_argument("--stream-points", type=int, default=200_000,
                    help="index size of the streamCEOs update baseline (host and device)")
    ap.add_argument("--no-stream-baseline", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-qps", type=float, default=-1.0,
                    help="offered rate of the served run (default: 60%% of the batched rate; 0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-inflight", type=int, default=4, help="batch slots of the served run")
    ap.add_argument("--serve-warmup-seconds", type=float, default=0.5,
                    help="untimed served run at the same rate before the measured one")
    ap.add_argument("--serve-max-batch", type=int, default=4096, help="queries per batch slot of the served run")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-rerank-bounds", action="store_true",
                    help="keep no int8 row copies: every candidate gets the exact FP32 rerank")
    ap.add_argument("--mode", default="coceos", choices=["coceos", "est", "hash"],
                    help="coceos (default): batched coCEOs; est: CEOs-Est build/search; hash: CEOs-Hash")
    ap.add_argument("--hash-points", type=int, default=100, help="n_probed_points of the CEOs-Hash mode")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (counter calibration under a profiler)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU (got --gpus {args.gpus}): the multi-GPU sharding is a design, not code",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from ceos_amd import CEOs, capi
    from ceos_amd.workloads import WORKLOADS, make_points, make_queries, make_query_params

    w = WORKLOADS[args.workload]
    n = args.n or w.n
    nq = args.nq or w.nq
    d = w.d
    dev = torch.device("cuda:0")
    torch.cuda.init()

    ks, ss, cs = make_query_params(w, nq)
    k_max, s_max, c_max = int(ks.max()), int(ss.max()), int(cs.max())
    log(f"workload {w.name}: n={n} d={d} nq={nq} D={w.D} R={w.R} m={w.m} "
        f"k in {sorted(set(ks.tolist()))[:3]} s in [{ss.min()},{s_max}] c in [{cs.min()},{c_max}]")
    t = time.time()
    X = make_points(w, n)
    Q = make_queries(w, nq)
    log(f"generated data in {time.time() - t:.1f}s")
    dX = torch.from_numpy(X).to(dev)
    dQ = torch.from_numpy(Q).to(dev)
    if args.mode != "coceos":
        run_mode(args, w, X, Q, dX, dQ, n, d, nq, torch, CEOs)
        return
    # a real (non-null) stream: the library launches on it and the events below bracket them
    ts = torch.cuda.Stream()
    torch.cuda.set_stream(ts)
    stream = ts.cuda_stream
    L = capi.lib()

    index = CEOs.CEOs(n, d)
    index.setIndexParam(w.D, w.R, w.m, -1, w.seed)
    index.rerank_bounds = not args.no_rerank_bounds
    h = index._handle
    torch.cuda.synchronize()
    # build: the first build also pays one-time allocations; time the second
    builds = []
    for _ in range(2):
        t0 = time.perf_counter()
        capi.check(L.ceos_build_coceos_device(h, dX.data_ptr(), n, d, stream))
        torch.cuda.synchronize()
        builds.append(time.perf_counter() - t0)
    build_s = builds[-1]
    pt = (ctypes.c_double * 5)()  # ceos_phase_times: copy, project, select, total (ms), index GB
    capi.check(L.ceos_get_build_times(h, ctypes.cast(pt, ctypes.c_void_p)))
    rs = (ctypes.c_int * 3)()
    capi.check(L.ceos_get_build_rescans(h, ctypes.byref(rs, 0), ctypes.byref(rs, 4), ctypes.byref(rs, 8)))
    build_phases = {"copy_ms": round(pt[0], 3), "project_ms": round(pt[1], 3), "select_ms": round(pt[2], 3),
                    "rescans": {"lists_under": rs[0], "lists_over": rs[1], "passes": rs[2]}}
    log(f"build {build_s * 1e3:.1f} ms ({n / build_s / 1e6:.1f} Mpts/s): copy {pt[0]:.3f} ms, "
        f"sample + scan {pt[1]:.3f} ms, list select {pt[2]:.3f} ms; rescans {list(rs)}")
    index.n_probed_vectors = w.s
    index.n_probed_points = w.m
    index.n_cand = w.c

    # ---- batched: a captured plan. The first search copies the queries and per-query
    # parameters into the plan's device buffers and autotunes the launch shape; the timed
    # steps replay the plan on those resident buffers.
    plan = index.make_plan(nq, k_max, s_max, c_max, autotune=not args.no_autotune)
    t0 = time.perf_counter()
    ids, _ = plan.search(Q, ks, ss, cs)  # rows of k_max: k[q] results then zeros
    log(f"plan first run (with autotune) {time.perf_counter() - t0:.2f}s; "
        f"(wg/CU, tile log2, tuned ms) = {plan.config()}, projection waves/WG {plan.proj_waves()}")
    dk = torch.from_numpy(ks).to(dev)
    ds = torch.from_numpy(ss).to(dev)
    dc = torch.from_numpy(cs).to(dev)

    def step():
        plan.run(nq, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    qps = nq / (ms_step / 1e3)
    log(f"batched (graph): {ms_step:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f} ms/step, "
        f"{qps:,.0f} QPS")
    wg, tl, _ = plan.config()

    # ---- the fused kernel alone, same launch shape, direct launches with kernel events
    index.set_search_config(wg, tl)
    off = np.zeros(nq, np.int64)
    off[1:] = np.cumsum(ks)[:-1]
    doff = torch.from_numpy(off).to(dev)
    tot = int(ks.sum())
    dI2 = torch.empty(max(1, tot), dtype=torch.int32, device=dev)
    dP2 = torch.empty(max(1, tot), dtype=torch.float32, device=dev)

    def direct():
        capi.check(L.ceos_search_batched_device(h, dQ.data_ptr(), nq, d, dk.data_ptr(), ds.data_ptr(), dc.data_ptr(),
                                                doff.data_ptr(), k_max, s_max, c_max, dI2.data_ptr(),
                                                dP2.data_ptr(), stream))

    direct()
    torch.cuda.synchronize()
    capi.check(L.ceos_set_kernel_timing(h, 1))
    for _ in range(args.steps):
        direct()
    front = ctypes.c_double()
    fused = ctypes.c_double()
    nl = ctypes.c_int()
    capi.check(L.ceos_get_kernel_timing(h, ctypes.byref(front), ctypes.byref(fused), ctypes.byref(nl)))
    capi.check(L.ceos_set_kernel_timing(h, 0))
    fused_ms = fused.value / max(1, nl.value)
    front_ms = front.value / max(1, nl.value)
    log(f"kernel events: fused search {fused_ms:.3f} ms/launch, projection+probe select {front_ms:.3f} ms/launch")
    _, tile_log2, tile_lds = index.get_search_config()
    index.set_search_config(0, 0)

    # ---- the practical streaming-read rate of this GPU: a plain grid-stride read of 4 GiB
    # (16-byte non-temporal loads, 8 in flight per thread), timed with HIP events on the
    # same stream, at 2048 to 16384 workgroups of 256; the best of the four
    xr = torch.ones(1 << 30, dtype=torch.float32, device=dev)
    orr = torch.empty(16384 * 256, dtype=torch.float32, device=dev)
    read_by_grid = {}
    for g in (2048, 4096, 8192, 16384):
        r = ctypes.c_double()
        capi.check(L.ceos_read_bandwidth(xr.data_ptr(), xr.numel(), orr.data_ptr(), g * 256, 10, stream,
                                         ctypes.byref(r)))
        read_by_grid[g] = round(r.value, 1)
    del xr, orr
    read_gbs = ctypes.c_double(max(read_by_grid.values()))
    log(f"streaming read of 4 GiB by grid: {read_by_grid} GB/s")

    if args.copy_mib:
        # a plain copy kernel of known size: under rocprofv3 --pmc its FETCH_SIZE and
        # WRITE_SIZE give the scale of the counters against known bytes
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"device copy of {args.copy_mib} MiB done")
        del src, dst

    # ---- recall against exact MIPS at each query's own k (and @10 where k >= 10)
    exact = exact_topk(torch, dX, dQ, k_max)
    rec_k = float(np.mean([len(np.intersect1d(exact[q, :ks[q]], ids[q, :ks[q]])) / ks[q]
                           for q in range(nq) if ks[q] > 0]))
    sel10 = np.nonzero(ks >= 10)[0]
    rec10 = (float(np.mean([len(np.intersect1d(exact[q, :10], ids[q, :10])) / 10.0 for q in sel10]))
             if len(sel10) else None)
    log(f"recall@k (own k) = {rec_k:.4f}; recall@10 over queries with k>=10 = {rec10}")

    # ---- plain per-query path: one call per query (its own parameters), synchronised
    P = min(args.plain_queries, nq)
    plain_qps, lat = None, np.zeros(1)
    if P > 0:
        dI1 = torch.empty(k_max, dtype=torch.int32, device=dev)
        dP1 = torch.empty(k_max, dtype=torch.float32, device=dev)

        def one(i):
            capi.check(L.ceos_set_query_param(h, int(ss[i]), w.m, int(cs[i])))
            capi.check(L.ceos_search_coceos_device(h, dQ[i].data_ptr(), 1, d, int(ks[i]), dI1.data_ptr(),
                                                   dP1.data_ptr(), stream))
            torch.cuda.synchronize()

        for i in range(3):
            one(i)
        lat = []
        t0 = time.perf_counter()
        for i in range(P):
            a = time.perf_counter()
            one(i)
            lat.append(time.perf_counter() - a)
        plain_qps = P / (time.perf_counter() - t0)
        lat = np.array(lat) * 1e3
        log(f"plain per-query: {plain_qps:,.0f} QPS, p50 {np.percentile(lat, 50):.3f} ms "
            f"p99 {np.percentile(lat, 99):.3f} ms")
        capi.check(L.ceos_set_query_param(h, w.s, w.m, w.c))

    # ---- served: the continuous-batching server (four batch slots of up to 4096 queries)
    # over the same index, open-loop Poisson arrivals of the workload's queries with their
    # per-query parameters; latency from arrival to results in host memory
    served = None
    rate = args.serve_qps if args.serve_qps >= 0 else round(0.6 * qps, -3)
    if rate > 0:
        srv = CEOs.Server(index, max_batch=args.serve_max_batch, k_max=k_max, s_max=s_max, c_max=c_max,
                          inflight=args.serve_inflight)

        def offer(seconds, seed):
            n_r = max(1, int(rate * seconds))
            idx = np.arange(n_r) % nq
            arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n_r))
            # the request arrays stay alive until the run ends: freeing gigabytes while
            # requests are being served stalls the process (the server copies them at submit)
            reqs = (Q[idx], ks[idx], ss[idx], cs[idx])
            srv.submit(*reqs, arrivals)
            srv.drain()
            del reqs
            return n_r

        offer(args.serve_warmup_seconds, 4)  # untimed: the server at the same load first
        srv.reset_stats()
        nreq = offer(args.serve_seconds, 3)
        st = srv.stats()
        trace = srv.batches()
        srv.close()
        slow = [{"launch_s": round(float(b[0]), 4), "enqueue_ms": round(float(b[1]), 3), "done_ms": round(float(b[2]), 3),
                 "oldest_wait_ms": round(float(b[3]), 3), "queries": int(b[4]),
                 "enqueue_steps_ms": [round(float(x), 3) for x in b[6:11]]}
                for b in trace[np.argsort(-trace[:, 2])[:3]]] if len(trace) else []
        served = {"offered_qps": rate, "achieved_qps": round(st["qps"], 1), "requests": nreq,
                  "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
                  "mean_batch": round(st["mean_batch"], 1), "inflight_slots": st["inflight_slots"],
                  "wait_ms": {"p50": round(st["wait_p50_ms"], 3), "p99": round(st["wait_p99_ms"], 3),
                              "max": round(st["wait_max_ms"], 3)},
                  "batch_service_ms": {"p50": round(st["service_p50_ms"], 3), "p99": round(st["service_p99_ms"], 3),
                                       "max": round(st["service_max_ms"], 3)},
                  "slowest_batches": slow, "warmup_s": args.serve_warmup_seconds,
                  "vs_plain": round(st["qps"] / plain_qps, 1) if plain_qps else None}
        log(f"served: {st['qps']:,.0f} QPS offered {rate:,.0f}, p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms, "
            f"mean batch {st['mean_batch']:.1f}; slowest batches {slow}")

    # ---- CPU baseline: the oracle's restatement of the reference's search and build,
    # timed on this host's cores with the reference's FFHT, tsl::robin_map and an FP32
    # SIMD dot (the timing build under oracle/_ref), on the same index and queries;
    # median of 3 runs
    cpu = None
    c_eff = cs.astype(np.float64)
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        info = host_info()
        threads = cpu_threads(args, info)
        timing = O.lib(timing=True) is not None
        cid, cv, fid, fv = index.get_buckets()
        o = O.OracleCEOs(n, d, timing=timing)
        o.setIndexParam(w.D, w.R, w.m, threads, w.seed)
        o.load_index(X, cid, cv, fid, fv)
        o.n_probed_vectors, o.n_probed_points, o.n_cand = w.s, w.m, w.c
        S = min(args.cpu_queries or nq, nq)
        o.search_coCEOs_Est(Q[:64], 0, ks=ks[:64], ss=ss[:64], cs=cs[:64])  # warm caches / thread pool
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            oi, op, distinct, _ = o.search_coCEOs_Est(Q[:S], 0, ks=ks[:S], ss=ss[:S], cs=cs[:S], stats=True)
            runs.append(time.perf_counter() - t0)
        cpu_s = float(np.median(runs))
        cpu_qps = S / cpu_s
        ooff = np.zeros(S + 1, np.int64)
        ooff[1:] = np.cumsum(ks[:S])
        match = float(np.mean([np.array_equal(oi[ooff[q]:ooff[q + 1]], ids[q, :ks[q]]) for q in range(S)]))
        c_eff[:S] = np.minimum(cs[:S], distinct)
        log(f"cpu baseline: {cpu_qps:,.0f} QPS on {S} queries (median of 3: {[round(r, 2) for r in runs]} s), "
            f"{threads} threads; top-k match {match:.4f}; mean distinct candidates {distinct.mean():.0f}")
        # one query per call on one thread: the per-query latency of the reference's search
        o.set_threads(1)
        lat1 = []
        for q in range(min(200, nq)):
            t0 = time.perf_counter()
            o.search_coCEOs_Est(Q[q:q + 1], 0, ks=ks[q:q + 1], ss=ss[q:q + 1], cs=cs[q:q + 1])
            lat1.append((time.perf_counter() - t0) * 1e3)
        del o
        # the build (the reference's: all projections, then a heap per list) on a bounded
        # sample of the same points and parameters
        nb = min(n, max(w.m, args.cpu_build_points))
        bt = []
        for _ in range(3):
            ob = O.OracleCEOs(nb, d, timing=timing)
            ob.setIndexParam(w.D, w.R, w.m, threads, w.seed)
            t0 = time.perf_counter()
            ob.build_coCEOs_Est(X[:nb])
            bt.append(time.perf_counter() - t0)
            del ob
        cpu_build_s = float(np.median(bt))
        cpu_build = {"mpts_s": round(nb / cpu_build_s / 1e6, 4), "sample": f"first {nb} points, median of 3",
                     "cores": threads, "seconds": round(cpu_build_s, 3)}
        log(f"cpu baseline build: {nb / cpu_build_s / 1e6:.4f} Mpts/s on {nb} points, {threads} threads")
        stream_upd = None if args.no_stream_baseline else stream_update_baseline(O, timing, threads, args.stream_points)
        what = ("the oracle restatement compiled with the reference's FFHT (src/fht.c), tsl::robin_map (src/tsl) "
                "and an FP32 AVX2 dot" if timing else "the oracle restatement (FP64 dot, textbook FWHT)")
        cpu = {"value": round(cpu_qps, 1), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"{S} of the {nq} queries on the same index (buckets from the device build, identical to "
                         f"the oracle's build in tests), median of 3 runs ({cpu_s:.2f} s); {what}",
               "topk_match_vs_gpu": match,
               "latency_ms_1thread": {"p50": round(float(np.percentile(lat1, 50)), 3),
                                      "p99": round(float(np.percentile(lat1, 99)), 3), "queries": len(lat1)},
               "build": cpu_build, "stream_update": stream_upd, "host": info}

    # ---- roofline of the fused search kernel. Algorithmic bytes per query: the query row,
    # the probed bucket entries (8 B each), the candidates' rows and the top-k output;
    # one launch covers all nq queries. With rerank bounds (k <= 256, 4 k <= candidates)
    # a candidate costs its int8 row and 16 B of row parameters, plus the FP32 rows of
    # the k best (the candidates past k that survive the bound are not counted).
    mp = w.m
    kf = ks.astype(np.float64)
    pruned = bool(index.rerank_bounds) & (kf >= 1) & (kf <= 256) & (4 * kf <= c_eff)
    rerank_bytes = np.where(pruned, (d + 16) * c_eff + 4 * d * kf, 4 * d * c_eff)
    bytes_launch = float(np.sum(4 * d + 8 * ss.astype(np.float64) * mp + rerank_bytes + 8 * ks))
    # the byte model of the reference's algorithm (every candidate's FP32 row read), for
    # comparison with the rate the search would need to match without the bound pass
    full_bytes = float(np.sum(4 * d + 8 * ss.astype(np.float64) * mp + 4 * d * c_eff + 8 * ks))
    achieved = bytes_launch / (fused_ms / 1e3) / 1e9
    # measured HBM bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on
    # this workload, scaled by a 1 GiB copy (profiles/k_search_fused_<workload>_traffic.json)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"k_search_fused_{w.name}_traffic.json")
    if os.path.exists(tpath) and n == w.n and nq == w.nq:
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_launch),
            "kernel": "k_search_fused", "kernel_ms": round(fused_ms, 4),
            "bytes_per_query": round(bytes_launch / nq), "front_kernels_ms": round(front_ms, 4),
            "stream_read_gbs": round(read_gbs.value, 1), "stream_read_by_grid_gbs": read_by_grid,
            "frac_of_stream_read": round(achieved / read_gbs.value, 4),
            "full_rerank_bytes_per_launch": round(full_bytes),
            "full_rerank_equivalent_gbs": round(full_bytes / (fused_ms * 1e-3) / 1e9, 1)}

    # ---- build roofline: the FWHT adds of every point (R repeats x 3 rounds x F log2 F) against
    # the non-packed FP32 add peak; HBM bytes = the points read once (4d per point) plus the
    # buckets written (2 lists x D*R columns x top_m entries x 8 B), plus with rerank bounds
    # the points read again and their int8 copies and row parameters written (5d + 16)
    F = 1 << (w.D - 1).bit_length() if w.D >= d else 1 << (d - 1).bit_length()
    adds_pt = w.R * 3 * F * (F.bit_length() - 1)
    build_bytes = 4.0 * d * n + 8.0 * 2 * w.D * w.R * w.m
    if index.rerank_bounds:
        build_bytes += (5.0 * d + 16) * n
    build = {"mpts_s": round(n / build_s / 1e6, 2), "ms": round(build_s * 1e3, 2),
             "phases": build_phases, "adds_per_point": adds_pt,
             "valu": {"achieved": round(adds_pt * n / build_s / 1e12, 3), "peak": VALU_ADD_PEAK_T,
                      "unit": "T add/s", "frac": round(adds_pt * n / build_s / 1e12 / VALU_ADD_PEAK_T, 4)},
             "hbm": {"achieved": round(build_bytes / build_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(build_bytes / build_s / 1e9 / HBM_PEAK_GBS, 4), "bytes": round(build_bytes)}}
    if cpu:
        build["cpu_oracle_mpts_s"] = cpu["build"]["mpts_s"]
        build["vs_cpu_oracle"] = round(build["mpts_s"] / cpu["build"]["mpts_s"], 1)

    uneven = bool(w.uneven)
    out = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (numpy PCG64 seeds 0/1/2; index sign seed 42)",
        "config": {"workload": w.name, "n": n, "d": d, "queries": nq, "n_proj": w.D, "n_repeats": w.R,
                   "top_m": w.m, "n_probed_points": mp,
                   "k": "per query {1,10,100}" if uneven else w.k,
                   "n_probed_vectors": f"per query U[{ss.min()},{s_max}]" if uneven else w.s,
                   "n_cand": f"per query U[{cs.min()},{c_max}]" if uneven else w.c,
                   "batch": "all queries in one hipGraph-captured plan (projection, probe select, fused search)",
                   "launch": {"wg_per_cu": wg, "table_log2": tl, "tile": "automatic (most buckets that fit)" if tl == 0
                              else "2^table_log2 slots", "tile_log2_floor": tile_log2, "lds_bytes": tile_lds,
                              "proj_waves_per_wg": plan.proj_waves(),
                              "autotune": [{"wg_per_cu": a, "table_log2": b, "proj_waves": c_, "median_ms": round(t, 3)}
                                           for a, b, c_, t in plan.tuning()]}},
        "recall_at_10": round(rec10, 4) if rec10 is not None else None,
        "recall_at_k": round(rec_k, 4),
        "plain_qps": round(plain_qps, 1) if plain_qps else None,
        "batched_vs_plain": round(qps / plain_qps, 1) if plain_qps else None,
        "served": served,
        "plain_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "build_mpts_s": round(n / build_s / 1e6, 2),
        "build_ms": round(build_s * 1e3, 2),
        "build": build,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
