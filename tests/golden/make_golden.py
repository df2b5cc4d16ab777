"""Writes the committed golden fixtures under tests/golden/.

Run from the repository root with the reference tree present (it compiles the
reference's own FFHT, src/fht.c, into oracle/_ref/ via oracle/Makefile):

    python tests/golden/make_golden.py

Fixtures (data only):
  sign_bits.json     bitHD3Generator bits (src/Utilities.cpp:21-41) for a few seeds,
                     drawn with the host's std::default_random_engine +
                     uniform_int_distribution<uint32_t>(0,1) through the oracle;
  ffht_vectors.npz   inputs and outputs of the reference's fht_float
                     (src/fht_avx.c) for log_n = 1..12, as raw float32;
  tiny_index.npz     a small coCEOs index and search computed by the oracle
                     (n=600, d=20, n_proj=32, n_repeats=2, top_m=12, seed=42).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    O.build()
    bits = {}
    for seed, nbits in ((42, 6144), (0, 2048), (7, 2048), (123456789, 2048)):
        b = O.sign_bits(nbits, seed)
        bits[str(seed)] = "".join(map(str, b.tolist()))
    with open(os.path.join(HERE, "sign_bits.json"), "w") as f:
        json.dump(bits, f, indent=0)

    ff = O.ffht_ref()
    if ff is None:
        raise SystemExit("reference FFHT not built (needs the reference tree)")
    rng = np.random.default_rng(2024)
    arrs = {}
    for L in range(1, 13):
        x = (rng.standard_normal(1 << L) * rng.lognormal(0, 1.0)).astype(np.float32)
        y = x.copy()
        ff.fht_float(y, L)
        arrs[f"x{L}"] = x
        arrs[f"y{L}"] = y
    np.savez_compressed(os.path.join(HERE, "ffht_vectors.npz"), **arrs)

    rng = np.random.default_rng(5)
    n, d = 600, 20
    X = rng.standard_normal((n, d)).astype(np.float32)
    X *= rng.lognormal(0, 0.5, (n, 1)).astype(np.float32)
    Q = rng.standard_normal((40, d)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    o = O.OracleCEOs(n, d)
    o.setIndexParam(32, 2, 12, 1, 42)
    o.build_coCEOs_Est(X)
    cid, cv, fid, fv = o.buckets()
    o.n_probed_vectors, o.n_cand = 6, 30
    ids, ips = o.search_coCEOs_Est(Q, 5)
    np.savez_compressed(os.path.join(HERE, "tiny_index.npz"), X=X, Q=Q, cid=cid, cval=cv, fid=fid, fval=fv,
                        ids=ids, ips=ips, params=np.array([32, 2, 12, 42, 6, 12, 30, 5], np.int32))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
