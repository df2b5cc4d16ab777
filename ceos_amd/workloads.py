"""Synthetic workloads for the coCEOs benchmarks and parity tests.

Shapes and distributions of the five benchmark configurations (all FP32, seeds:
data = 0, queries = 1, per-query parameters = 2, arrivals = 3, index sign seed 42):

  netflix_like  17,770 x 300 rows N(0,I) rescaled to LogNormal(0,0.5) norms;
                1,000 unit queries; D=1024, R=2, m=100, s=20, c=500, k=10
  mixture_1m    1,000,000 x 128 from a 100-centre Gaussian mixture (centres N(0,I),
                sigma 0.3), rows rescaled to LogNormal(0,0.5) norms; 10,000 unit
                queries from the same mixture; D=1024, R=4, m=1000, s=40, c=1000, k=10
  gist_like     1,000,000 x 960 rows N(0,I) rescaled to LogNormal(0,0.7) norms;
                20,000 queries with k in {1,10,100}, s ~ U[10,80], c ~ U[200,5000];
                D=1024, R=8, m=500
  serve_10m     10,000,000 x 256, 1,000-centre mixture (sigma 0.25); D=1024, R=4, m=2000
  stream_5m     5,000,000 x 384 N(0,I) rows, LogNormal(0,0.5) norms; D=512, R=8, m=1000
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    n: int
    d: int
    nq: int
    D: int
    R: int
    m: int
    s: int
    c: int
    k: int
    seed: int = 42
    sigma_norm: float = 0.5
    mixture: int = 0          # number of centres (0 = isotropic Gaussian rows)
    mixture_sigma: float = 0.3
    uneven: bool = False      # per-query k / s / c
    extra: dict = field(default_factory=dict)


WORKLOADS = {
    "netflix_like": Workload("netflix_like", 17770, 300, 1000, 1024, 2, 100, 20, 500, 10),
    "mixture_1m": Workload("mixture_1m", 1_000_000, 128, 10_000, 1024, 4, 1000, 40, 1000, 10, mixture=100,
                           mixture_sigma=0.3),
    "gist_like": Workload("gist_like", 1_000_000, 960, 20_000, 1024, 8, 500, 40, 1000, 10, sigma_norm=0.7,
                          uneven=True),
    "serve_10m": Workload("serve_10m", 10_000_000, 256, 100_000, 1024, 4, 2000, 40, 1000, 10, mixture=1000,
                          mixture_sigma=0.25),
    "stream_5m": Workload("stream_5m", 5_000_000, 384, 50_000, 512, 8, 1000, 40, 1000, 10),
}


def _rescale(rows: np.ndarray, rng: np.random.Generator, sigma: float) -> np.ndarray:
    norms = np.linalg.norm(rows, axis=1, keepdims=True)
    norms[norms == 0] = 1.0
    target = rng.lognormal(0.0, sigma, size=(rows.shape[0], 1)).astype(np.float32)
    return (rows / norms * target).astype(np.float32)


def _unit(rows: np.ndarray) -> np.ndarray:
    norms = np.linalg.norm(rows, axis=1, keepdims=True)
    norms[norms == 0] = 1.0
    return (rows / norms).astype(np.float32)


def make_points(w: Workload, n: int | None = None, seed: int = 0, chunk: int = 1 << 18) -> np.ndarray:
    """Data rows (n x d FP32), generated in chunks to bound host memory."""
    n = w.n if n is None else n
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((w.mixture, w.d)).astype(np.float32) if w.mixture else None
    out = np.empty((n, w.d), np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        g = rng.standard_normal((b - a, w.d), dtype=np.float32)
        if centres is not None:
            g = centres[rng.integers(0, w.mixture, b - a)] + w.mixture_sigma * g
        out[a:b] = _rescale(g, rng, w.sigma_norm)
    return out


def make_queries(w: Workload, nq: int | None = None, seed: int = 1) -> np.ndarray:
    """Unit-norm queries (same mixture as the data when the workload has one)."""
    nq = w.nq if nq is None else nq
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nq, w.d), dtype=np.float32)
    if w.mixture:
        crng = np.random.default_rng(0)  # the data's centres
        centres = crng.standard_normal((w.mixture, w.d)).astype(np.float32)
        g = centres[rng.integers(0, w.mixture, nq)] + w.mixture_sigma * g
    return _unit(g)


def make_query_params(w: Workload, nq: int | None = None, seed: int = 2):
    """Per-query (k, s, c) arrays: uneven ranges for the uneven-batching workload."""
    nq = w.nq if nq is None else nq
    if not w.uneven:
        return (np.full(nq, w.k, np.int32), np.full(nq, w.s, np.int32), np.full(nq, w.c, np.int32))
    rng = np.random.default_rng(seed)
    k = rng.choice(np.array([1, 10, 100], np.int32), nq).astype(np.int32)
    s = rng.integers(10, 81, nq).astype(np.int32)
    c = rng.integers(200, 5001, nq).astype(np.int32)
    return k, s, c


def recall(exact: np.ndarray, approx: np.ndarray) -> float:
    """Mean |exact ∩ approx| / k over queries (test/utils.py:27-32)."""
    n, k = exact.shape
    tot = 0.0
    for i in range(n):
        tot += len(np.intersect1d(exact[i], approx[i])) / k
    return tot / n
