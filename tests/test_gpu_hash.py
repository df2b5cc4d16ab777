"""CEOs-Hash (build_CEOs_Hash / search_CEOs_Hash, streamCEOs.hash_search) on the GPU
against the oracle: every distinct id of the probed lists is reranked."""
import numpy as np
import pytest

from tests._parity import bits

pytestmark = pytest.mark.gpu


def _rand(n, d, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    return (g * rng.lognormal(0, sigma, (n, 1))).astype(np.float32)


def _same(gi, gp, oi, op):
    """ids identical except inside runs of exactly equal inner products."""
    np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-6)
    for q in range(len(gi)):
        if np.array_equal(gi[q], oi[q]):
            continue
        for t in np.nonzero(gi[q] != oi[q])[0]:
            assert gp[q, t] == op[q, t]  # a tie


@pytest.mark.parametrize("n,d,D,R,m,s,mp,k", [(5000, 48, 128, 2, 80, 10, 80, 10),
                                               (20000, 100, 256, 4, 200, 30, 150, 25),
                                               (3000, 33, 64, 1, 3000, 4, 3000, 5)])
def test_hash_search_matches_oracle(CEOs, oracle, n, d, D, R, m, s, mp, k):
    X = _rand(n, d, 81)
    Q = _rand(60, d, 82)
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, 42)
    g.build_CEOs_Hash(X)
    g.n_probed_vectors, g.n_probed_points = s, mp
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 8, 42)
    o.build_CEOs_Hash(X)
    o.n_probed_vectors, o.n_probed_points = s, mp
    gi, gp = g.search_CEOs_Hash(Q, k)
    oi, op = o.search_CEOs_Hash(Q, k)
    _same(gi, gp, oi, op)
    if m == n:  # every point is a candidate: exact MIPS
        ex, _ = oracle.exact_mips(X, Q, k)
        assert np.array_equal(gi, ex)


def test_hash_equals_coceos_with_every_candidate(CEOs):
    """coCEOs with n_cand above the number of distinct candidates reranks the same set."""
    X = _rand(10000, 40, 83)
    Q = _rand(40, 40, 84)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(128, 2, 100, -1, 42)
    g.build_coCEOs_Est(X)
    g.n_probed_vectors, g.n_probed_points, g.n_cand = 12, 100, 100000
    a, pa = g.search_coCEOs_Est(Q, 10)
    b, pb = g.search_CEOs_Hash(Q, 10)
    assert np.array_equal(a, b) and np.array_equal(bits(pa), bits(pb))


def test_hash_centering(CEOs, oracle):
    X = _rand(4000, 24, 85) + 3.0
    Q = _rand(30, 24, 86)
    g = CEOs.CEOs(*X.shape)
    g.centering = True
    g.setIndexParam(64, 2, 60, -1, 42)
    g.build_CEOs_Hash(X)
    g.n_probed_vectors, g.n_probed_points = 8, 60
    Xc = (X - X.astype(np.float64).mean(axis=0)).astype(np.float32)
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(64, 2, 60, 8, 42)
    o.build_CEOs_Hash(Xc)
    o.n_probed_vectors, o.n_probed_points = 8, 60
    gi, gp = g.search_CEOs_Hash(Q, 5)
    oi, op = o.search_CEOs_Hash(Q, 5)
    _same(gi, gp, oi, op)


def test_stream_hash_search(CEOs, oracle):
    d, D, R, m = 32, 128, 2, 50
    X = _rand(9000, d, 87)
    Q = _rand(40, d, 88)
    g = CEOs.streamCEOs(d)
    g.setIndexParam(D, R, m, -1, 42)
    o = oracle.OracleStream(d)
    o.setIndexParam(D, R, m, 8, 42)
    g.build(X[:4000]); o.build(X[:4000])
    for obj in (g, o):
        obj.n_probed_vectors, obj.n_probed_points = 16, m
    for step in range(3):
        gi, gp = g.hash_search(Q, 10)
        oi, op = o.hash_search(Q, 10)
        _same(gi, gp, oi, op)
        g.update(X[4000 + 1000 * step:5000 + 1000 * step], 600)
        o.update(X[4000 + 1000 * step:5000 + 1000 * step], 600)
    g.n_probed_vectors = D + 1  # s <= n_proj (src/streamCEOs.cpp:630)
    with pytest.raises(ValueError):
        g.hash_search(Q, 10)


def test_hash_search_multi_pass_small_tile(CEOs, oracle):
    """A 256-slot tile forces several hash partitions per query (and stash use)."""
    X = _rand(8000, 32, 89)
    Q = _rand(40, 32, 90)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(128, 2, 200, -1, 42)
    g.build_CEOs_Hash(X)
    g.n_probed_vectors, g.n_probed_points = 20, 200
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(128, 2, 200, 8, 42)
    o.build_CEOs_Hash(X)
    o.n_probed_vectors, o.n_probed_points = 20, 200
    oi, op = o.search_CEOs_Hash(Q, 10)
    for cfg in ((1, 8), (2, 9), (0, 0)):
        g.set_search_config(*cfg)
        gi, gp = g.search_CEOs_Hash(Q, 10)
        _same(gi, gp, oi, op)
