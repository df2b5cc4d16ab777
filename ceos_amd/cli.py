"""Command line: the reference's test driver (src/main.cpp) over the MI355X library.

  python -m ceos_amd.cli --X data.bin --Q queries.bin --n_points N --n_features D
      --n_queries Q [--n_proj P --n_repeats R --top_m M --n_threads T --random_seed S
       --centering --n_neighbors K --n_probed_vectors V --n_probed_points PP --n_cand C
       --verbose] [--method est,stream,hash,coceos] [--out results.npz]

Flags and defaults follow readIndexParam / readQueryParam (src/Utilities.cpp:221-496):
n_proj = max(256, 2^ceil(log2 n_features)), n_repeats 1, top_m 1000, n_threads -1,
random_seed -1, n_neighbors 1, n_probed_vectors 10, n_probed_points 10,
n_cand = n_neighbors. The default sequence is main.cpp's: CEOs-Est build and search,
streamCEOs build and estimate_search, update(queries, 1000), hash_search. --method picks
other sequences (coceos: build_coCEOs_Est + search_coCEOs_Est); --out saves every
result as ids_<step> / ips_<step>.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np


def parse(argv=None):
    ap = argparse.ArgumentParser(prog="ceos_amd.cli", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--X", required=True)
    ap.add_argument("--Q", required=True)
    ap.add_argument("--n_points", type=int, required=True)
    ap.add_argument("--n_features", type=int, required=True)
    ap.add_argument("--n_queries", type=int, required=True)
    ap.add_argument("--n_proj", type=int, default=0)
    ap.add_argument("--n_repeats", type=int, default=1)
    ap.add_argument("--top_m", type=int, default=1000)
    ap.add_argument("--n_threads", type=int, default=-1)
    ap.add_argument("--random_seed", type=int, default=-1)
    ap.add_argument("--centering", action="store_true")
    ap.add_argument("--n_neighbors", type=int, default=1)
    ap.add_argument("--n_probed_vectors", type=int, default=10)
    ap.add_argument("--n_probed_points", type=int, default=10)
    ap.add_argument("--n_cand", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--method", default="est,stream,hash")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    if a.n_proj <= 0:  # Utilities.cpp:278-281
        a.n_proj = max(256, 1 << int(math.ceil(math.log2(a.n_features))))
    if a.n_cand <= 0:
        a.n_cand = a.n_neighbors
    return a


def run(a) -> dict:
    from ceos_amd import CEOs
    from ceos_amd.io import load_bin

    X = np.ascontiguousarray(load_bin(a.X, a.n_points, a.n_features))
    Q = np.ascontiguousarray(load_bin(a.Q, a.n_queries, a.n_features))
    out = {}
    k = a.n_neighbors

    def timed(name, fn):
        t0 = time.perf_counter()
        r = fn()
        print(f"{name}: {(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
        if r is not None:
            out[f"ids_{name}"], out[f"ips_{name}"] = r
        return r

    steps = [m.strip() for m in a.method.split(",") if m.strip()]
    for m in steps:
        if m in ("est", "coceos", "hash"):
            g = CEOs.CEOs(a.n_points, a.n_features)
            g.setIndexParam(a.n_proj, a.n_repeats, a.top_m, a.n_threads, a.random_seed)
            g.centering = a.centering
            g.n_probed_vectors, g.n_probed_points, g.n_cand = a.n_probed_vectors, min(a.n_probed_points,
                                                                                     a.top_m), a.n_cand
            if m == "est":
                timed("build_CEOs", lambda: g.build(X))
                timed("search_CEOs", lambda: g.search(Q, k, a.verbose))
            elif m == "coceos":
                timed("build_coCEOs_Est", lambda: g.build_coCEOs_Est(X))
                timed("search_coCEOs_Est", lambda: g.search_coCEOs_Est(Q, k, a.verbose))
            else:
                timed("build_CEOs_Hash", lambda: g.build_CEOs_Hash(X))
                timed("search_CEOs_Hash", lambda: g.search_CEOs_Hash(Q, k, a.verbose))
        elif m == "stream":  # main.cpp: build, estimate_search, update(Q, 1000), hash_search
            s = CEOs.streamCEOs(a.n_features)
            s.setIndexParam(a.n_proj, a.n_repeats, a.top_m, a.n_threads, a.random_seed)
            s.n_probed_vectors, s.n_cand = a.n_probed_vectors, a.n_cand
            s.n_probed_points = min(a.n_probed_points, a.top_m)
            timed("stream_build", lambda: s.build(X))
            timed("estimate_search", lambda: s.estimate_search(Q, k, a.verbose))
            timed("update", lambda: s.update(Q, min(1000, s.size())))
            timed("hash_search", lambda: s.hash_search(Q, k, a.verbose))
        else:
            raise SystemExit(f"unknown method {m!r}")
    if a.out:
        np.savez(a.out, **out)
    return out


def main(argv=None) -> int:
    a = parse(argv)
    try:
        run(a)
    except (ValueError, RuntimeError) as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
