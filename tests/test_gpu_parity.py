"""GPU parity: the HIP path against the CPU oracle on the same inputs and seed.

Bars: projections bit-exact; buckets identical (ids and value bits) except at
logged exact ties; top-k ids identical and inner products within 1e-5 relative.
"""
import numpy as np
import pytest

from tests._parity import compare_buckets
from ceos_amd.workloads import WORKLOADS, make_points, make_queries, recall

pytestmark = pytest.mark.gpu

IP_RTOL = 1e-5


def _rand(n, d, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return (g * rng.lognormal(0, sigma, (n, 1))).astype(np.float32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("d,D", [(10, 16), (40, 64), (100, 128), (200, 256), (300, 512), (300, 1024),
                                 (960, 1024), (1500, 2048), (3000, 4096), (129, 100)])
def test_projection_bitexact(CEOs, oracle, d, D):
    n = 300
    X = _rand(n, d, 7)
    R = 2
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, 5, -1, 42)
    g.build_coCEOs_Est(X)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, 5, 1, 42)
    assert g.fhtDim == o.F
    assert np.array_equal(g.get_sign_bits(), oracle.sign_bits(o.F * 3 * R, 42))
    rows = _rand(64, d, 8)
    pg = g.project(rows)
    po = o.project(rows)
    assert np.array_equal(_bits(pg), _bits(po)), f"max abs diff {np.abs(pg - po).max()}"


def _build_pair(CEOs, oracle, X, D, R, m, seed=42):
    n, d = X.shape
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, seed)
    g.build_coCEOs_Est(X)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 8, seed)
    o.build_coCEOs_Est(X)
    return g, o


def _compare_buckets(g, o):
    return compare_buckets(g.get_buckets(), o.buckets())


def test_buckets_small(CEOs, oracle):
    X = _rand(3000, 60, 1)
    g, o = _build_pair(CEOs, oracle, X, 128, 2, 50)
    _compare_buckets(g, o)


def _compare_search(gi, gp, oi, op, tol_frac=0.001):
    assert gi.shape == oi.shape
    same = np.all(gi == oi, axis=1)
    assert same.mean() >= 1 - tol_frac, f"{(~same).sum()} of {len(same)} queries differ"
    np.testing.assert_allclose(gp[same], op[same], rtol=IP_RTOL, atol=1e-6)


def test_search_small(CEOs, oracle):
    X = _rand(3000, 60, 1)
    Q = _rand(200, 60, 2)
    g, o = _build_pair(CEOs, oracle, X, 128, 2, 50)
    for obj in (g, o):
        obj.n_probed_vectors = 10
        obj.n_cand = 100
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


def test_netflix_like_config(CEOs, oracle):
    w = WORKLOADS["netflix_like"]
    X = make_points(w)
    Q = make_queries(w)
    g, o = _build_pair(CEOs, oracle, X, w.D, w.R, w.m)
    _compare_buckets(g, o)
    for obj in (g, o):
        obj.n_probed_vectors = w.s
        obj.n_cand = w.c
    gi, gp = g.search_coCEOs_Est(Q, w.k)
    oi, op = o.search_coCEOs_Est(Q, w.k)
    _compare_search(gi, gp, oi, op)
    ex, _ = oracle.exact_mips(X, Q, w.k)
    rg, ro = recall(ex, gi), recall(ex, oi)
    print(f"recall@10 gpu {rg:.4f} oracle {ro:.4f}")
    assert abs(rg - ro) <= 0.001


def test_ceos_est_small(CEOs, oracle):
    X = _rand(2500, 50, 3)
    Q = _rand(100, 50, 4)
    g = CEOs.CEOs(2500, 50)
    g.setIndexParam(128, 2, 10, -1, 42)
    g.build(X)
    o = oracle.OracleCEOs(2500, 50)
    o.setIndexParam(128, 2, 10, 8, 42)
    o.build(X)
    for obj in (g, o):
        obj.n_probed_vectors = 12
        obj.n_cand = 80
    gi, gp = g.search(Q, 10)
    oi, op = o.search(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


def test_buckets_sampled_thresholds(CEOs, oracle):
    """150,000 points: the build samples every 4th point for its per-list thresholds
    (rank about 83), so the exact per-list selection runs on the candidates that pass
    them. Buckets must equal the oracle's, which scans every point."""
    X = _rand(150_000, 24, 11)
    g, o = _build_pair(CEOs, oracle, X, 64, 2, 120)
    _compare_buckets(g, o)


def test_buckets_sampled_thresholds_duplicates(CEOs, oracle):
    """A third of the points are one repeated row, so in every column tens of thousands of
    sample values tie at the threshold bound: the threshold kernel's collected tail
    overflows and it falls back to a radix select over the whole column. Buckets must
    still equal the oracle's (ties compared as sets)."""
    X = _rand(120_000, 16, 12)
    X[::3] = X[5]
    g, o = _build_pair(CEOs, oracle, X, 32, 2, 100)
    _compare_buckets(g, o)


@pytest.mark.parametrize("scale", [100.0, 0.01])
def test_buckets_misjudging_sample(CEOs, oracle, scale):
    """The build's strided sample (every `stride`-th point, stride = n // max(32768,
    16 n / top_m), 4 here) is made unrepresentative: the sampled rows are scaled up (every
    list's threshold is too high and under top_m points pass it) or down (nearly every
    point passes: every list overflows its space). Every list is then rescanned; with a
    budget of 4 n entries the rescans run over many small ranges of lists. Buckets must
    still equal the oracle's."""
    from ceos_amd import capi

    n, d, D, R, m = 150_000, 24, 64, 2, 120
    stride = n // max(32768, -(-16 * n // m))
    assert stride == 4
    X = _rand(n, d, 14)
    X[::stride] *= np.float32(scale)
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, 42)
    capi.check(capi.lib().ceos_set_build_budget(g._handle, 4 * n))
    g.build_coCEOs_Est(X)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 8, 42)
    o.build_coCEOs_Est(X)
    _compare_buckets(g, o)


@pytest.mark.parametrize("misaligned", [False, True])
def test_device_source_build(CEOs, oracle, misaligned):
    """A build from device memory: from a 16-byte aligned source the bucket scan reads
    the caller's rows and stores the index's copy on the way; a misaligned source is
    copied first. Buckets must equal the oracle's, and searches (exact rerank, and with
    the int8 bound pass) must give the same bits as an index built from host memory."""
    from ceos_amd import capi
    from tests._hip import DeviceArray

    n, d, D, R, m = 50_000, 96, 128, 3, 200
    X = _rand(n, d, 15)
    Q = _rand(100, d, 16, sigma=0.0)
    flat = np.concatenate([np.zeros(1, np.float32), X.ravel()]) if misaligned else X.ravel()
    dX = DeviceArray(flat)
    src = dX.ptr + (4 if misaligned else 0)
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, 42)
    capi.check(capi.lib().ceos_build_coceos_device(g._handle, src, n, d, None))
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 8, 42)
    o.build_coCEOs_Est(X)
    _compare_buckets(g, o)
    h = CEOs.CEOs(n, d)
    h.setIndexParam(D, R, m, -1, 42)
    h.build_coCEOs_Est(X)
    for bounds in (False, True):
        for idx in (g, h):
            idx.rerank_bounds = bounds
            idx.n_probed_vectors = 20
            idx.n_cand = 2000
        gi, gp = g.search_coCEOs_Est(Q, 10)
        hi, hp = h.search_coCEOs_Est(Q, 10)
        assert np.array_equal(gi, hi)
        assert np.array_equal(_bits(gp), _bits(hp))


def test_build_sampled_thresholds_no_rescans(CEOs):
    """600,000 points, D = 1024, R = 4, top_m = 300: the threshold sample has 33,334
    points (130 per thread of the per-column kernel) and rank 49, so in about 1% of the
    8,192 lists the 49 largest sample values fall to 49 different threads and the
    collected tail holds exactly rank values. Every threshold must still be the sample's
    rank-th value: no list may end under top_m or over its space (below 1e-9 per list by
    the sample's statistics), which a wrong threshold shows as rescans."""
    from ceos_amd import capi

    n, d = 600_000, 32
    X = _rand(n, d, 17)
    g = CEOs.CEOs(n, d)
    g.setIndexParam(1024, 4, 300, -1, 42)
    g.build_coCEOs_Est(X)
    import ctypes

    r = (ctypes.c_int * 3)()
    capi.check(capi.lib().ceos_get_build_rescans(g._handle, ctypes.byref(r, 0), ctypes.byref(r, 4),
                                                 ctypes.byref(r, 8)))
    assert list(r) == [0, 0, 0]


def test_buckets_large_top_m_exact_sample(CEOs, oracle):
    """n below the sample size: every point is sampled and the threshold rank equals
    top_m (300, above the per-column kernel's 256 per-thread bounds), which takes the
    one-list-per-workgroup radix path."""
    X = _rand(6000, 40, 13)
    g, o = _build_pair(CEOs, oracle, X, 64, 2, 300)
    _compare_buckets(g, o)


@pytest.mark.parametrize("n_cand", [200, 1000, 5000])
def test_search_candidate_select_paths(CEOs, oracle, n_cand):
    """About 12,700 distinct candidates per query in one tile pass. n_cand 200 and 1000
    take the sampled-bound compaction (a lower bound of the n_cand-th score from 512
    sampled tile slots, then an exact select among the sums above it); 5000 takes the
    value-histogram select. Every query's top-10 must equal the oracle's."""
    rng = np.random.default_rng(21)
    n, d = 200_000, 32
    X = (rng.standard_normal((n, d)) * rng.lognormal(0, 0.5, (n, 1))).astype(np.float32)
    Q = rng.standard_normal((300, d)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    g, o = _build_pair(CEOs, oracle, X, 256, 2, 2000)
    for obj in (g, o):
        obj.n_probed_vectors = 10
        obj.n_cand = n_cand
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


@pytest.mark.parametrize("n_points", [200_000, 200_001])
def test_ceos_est_sampled_bound(CEOs, oracle, n_points):
    """CEOs-Est over enough points that the device bounds the n_cand-th estimate from a
    sample and keeps only the points at or above it; an odd point count takes the
    scalar column path. Ids and inner products must equal the oracle's dense search."""
    rng = np.random.default_rng(33)
    d = 24
    X = (rng.standard_normal((n_points, d)) * rng.lognormal(0, 0.5, (n_points, 1))).astype(np.float32)
    Q = rng.standard_normal((100, d)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    g = CEOs.CEOs(n_points, d)
    g.setIndexParam(128, 2, 10, -1, 42)
    g.build(X)
    o = oracle.OracleCEOs(n_points, d)
    o.setIndexParam(128, 2, 10, 8, 42)
    o.build(X)
    for obj in (g, o):
        obj.n_probed_vectors = 12
        obj.n_cand = 300
    gi, gp = g.search(Q, 10)
    oi, op = o.search(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


@pytest.mark.parametrize("s", [8, 64, 256])
def test_probe_select_sampled_bound(CEOs, oracle, s):
    """D*R = 4096 projections per query: the probe select gathers the columns above a
    bound sampled from 256 strided |p| values and selects the top-s from that list
    (16 s <= D*R); s = 256 is also the largest s that takes the sampled path. The
    probe order decides the summation order, so ids must equal the oracle's."""
    X = _rand(20_000, 64, 41)
    Q = _rand(120, 64, 42)
    g, o = _build_pair(CEOs, oracle, X, 1024, 4, 100)
    for obj in (g, o):
        obj.n_probed_vectors = s
        obj.n_cand = 300
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


@pytest.mark.parametrize("wg,tl", [(1, 0), (2, 0), (1, 12), (1, 9)])
def test_search_tile_sizes(CEOs, oracle, wg, tl):
    """The automatic tile takes the most hashed slots that fit the LDS share (not a
    power of two in general: buckets are chosen by multiply-high of the id hash); small
    explicit tiles force several hash partitions. Every shape gives the oracle's ids."""
    X = _rand(60_000, 96, 51)
    Q = _rand(150, 96, 52)
    g, o = _build_pair(CEOs, oracle, X, 256, 2, 1500)
    g.set_search_config(wg, tl)
    for obj in (g, o):
        obj.n_probed_vectors = 60
        obj.n_cand = 2000
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)


@pytest.mark.parametrize("s", [3, 5, 7, 8, 9])
def test_search_single_bucket_tile_load(CEOs, oracle, s):
    """A tile of 2^14 slots and about 1,500 distinct ids per probe: the passes planned at
    most 70% full place each id in one bucket, so the fuller ones send the ids of full
    buckets to the stash; fuller passes take both buckets or several partitions. Every
    load gives the oracle's ids."""
    X = _rand(60_000, 96, 61)
    Q = _rand(150, 96, 62)
    g, o = _build_pair(CEOs, oracle, X, 256, 2, 1500)
    g.set_search_config(1, 14)
    for obj in (g, o):
        obj.n_probed_vectors = s
        obj.n_cand = 500
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    _compare_search(gi, gp, oi, op, tol_frac=0.0)
