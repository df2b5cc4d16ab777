"""The rerank's bound pass: with int8 row copies on, a search gives exactly the ids and
inner products of the exact rerank over every candidate (copies off), and matches the
oracle. Cases cover both lane-load widths (d % 16 == 0 and d % 4 == 0), the size
limits of the pass (k <= 256, 4 k <= n_cand), exact ties, rows the int8 copy
represents badly, survivor lists past the in-LDS capacity (exact fallback), and
stream indexes whose ring grows and wraps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(n, d, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return (g * rng.lognormal(0, sigma, (n, 1))).astype(np.float32)


def _index(CEOs, X, D=256, R=2, m=400):
    n, d = X.shape
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, 42)
    g.build_coCEOs_Est(X)
    return g


def _both(g, Q, k, s, c, search=None):
    """(ids, ips) with the bound pass on and off; the index is left with it on."""
    g.n_probed_vectors = s
    if c > 1:
        g.n_cand = c
    run = search or (lambda: g.search_coCEOs_Est(Q, k))
    assert g.rerank_bounds
    on = run()
    g.rerank_bounds = False
    assert not g.rerank_bounds
    off = run()
    g.rerank_bounds = True
    again = run()
    for a, b in ((on, off), (again, off)):
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    return on


@pytest.mark.parametrize("d", [960, 128, 60, 300])
@pytest.mark.parametrize("k,c", [(1, 200), (10, 1000), (100, 1000), (100, 300), (256, 2000)])
def test_bounds_equal_exact_rerank(CEOs, d, k, c):
    X = _rand(20_000, d, 3, sigma=0.7)
    Q = _rand(64, d, 4, sigma=0.0)
    g = _index(CEOs, X)
    _both(g, Q, k, 40, c)


def test_bounds_match_oracle(CEOs, oracle):
    X = _rand(20_000, 96, 5)
    Q = _rand(100, 96, 6, sigma=0.0)
    g = _index(CEOs, X, 256, 2, 300)
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(256, 2, 300, 8, 42)
    o.build_coCEOs_Est(X)
    for obj in (g, o):
        obj.n_probed_vectors, obj.n_cand = 20, 800
    gi, gp = _both(g, Q, 10, 20, 800)
    oi, op = o.search_coCEOs_Est(Q, 10)
    assert np.array_equal(gi, oi)
    np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-6)


def test_bounds_adversarial_rows(CEOs):
    """Exact duplicates (tied inner products, broken by id), rows with one large entry
    (coarse int8 steps for the rest), tiny-scale rows, zero rows and near-duplicates
    that the int8 copy cannot tell apart (every one survives)."""
    rng = np.random.default_rng(7)
    n, d = 12_000, 64
    X = _rand(n, d, 8)
    X[1000:1400] = X[1000]                                   # 400 exact duplicates
    X[2000:2300, 0] = 1e4                                    # one huge entry
    X[3000:3100] *= 1e-30                                    # tiny scale
    X[4000:4050] = 0.0                                       # zero rows
    base = X[5000].copy()
    X[5000:8000] = base + rng.standard_normal((3000, d)).astype(np.float32) * 1e-6  # near-duplicates
    Q = np.concatenate([X[[1000, 2000, 5000, 5001, 7]] / np.linalg.norm(X[[1000, 2000, 5000, 5001, 7]], axis=1,
                                                                          keepdims=True),
                        _rand(59, d, 9, sigma=0.0)]).astype(np.float32)
    g = _index(CEOs, X, 256, 2, 1000)
    for k, c in ((10, 200), (10, 3000), (50, 5000), (1, 4000)):
        _both(g, Q, k, 30, c)


def test_bounds_batched_uneven(CEOs):
    """Per-query k / n_cand in one batched launch: queries on both sides of the
    4 k <= n_cand and k <= 256 limits share the launch."""
    X = _rand(30_000, 256, 10, sigma=0.7)
    Q = _rand(200, 256, 11, sigma=0.0)
    g = _index(CEOs, X, 512, 2, 500)
    rng = np.random.default_rng(12)
    k = rng.choice(np.array([1, 10, 100, 300], np.int32), len(Q)).astype(np.int32)
    s = rng.integers(5, 41, len(Q)).astype(np.int32)
    c = rng.integers(100, 3001, len(Q)).astype(np.int32)
    _both(g, Q, 0, 40, 3000, search=lambda: g.search_batched(Q, k, s, c))


def test_bounds_stream_ring(CEOs):
    """A stream index keeps its int8 copies in step through inserts (ring growth and
    wrap) and deletes."""
    d = 128
    g = CEOs.streamCEOs(d)
    g.setIndexParam(256, 2, 300, -1, 42)
    g.build(_rand(5000, d, 13))
    Q = _rand(50, d, 14, sigma=0.0)

    def run():
        return g.estimate_search(Q, 10)

    for step in range(4):
        g.update(_rand(3000 + 1000 * step, d, 20 + step), 2500)
        g.n_probed_vectors, g.n_cand = 30, 600
        on = run()
        g.rerank_bounds = False
        off = run()
        g.rerank_bounds = True
        assert np.array_equal(on[0], off[0])
        assert np.array_equal(on[1].view(np.uint32), off[1].view(np.uint32))


@pytest.mark.parametrize("d,k", [(128, 10), (960, 10), (300, 1), (64, 100)])
def test_bounds_hash_mode(CEOs, d, k):
    """CEOs-Hash reranks every distinct candidate from global scratch: the bound pass
    there too gives the exact rerank's ids and bits."""
    X = _rand(15_000, d, 30, sigma=0.7)
    Q = _rand(48, d, 31, sigma=0.0)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(256, 2, 300, -1, 42)
    g.build_CEOs_Hash(X)
    g.n_probed_vectors, g.n_probed_points = 20, 200
    _both(g, Q, k, 20, 1, search=lambda: g.search_CEOs_Hash(Q, k))
