"""Centering and non-finite rows against the oracle, and the reference's printed lines.

Centering (src/streamCEOs.cpp:22-34, 198-203; CEOs-Hash src/CEOs.cpp:721-728): the device
computes the centre as the FP64 column mean rounded once to FP32. The reference takes
Eigen's FP32 colwise().mean(), whose summation order depends on the Eigen version (not
in the reference tree), so the centre is compared with the oracle's Eigen-order FP32
mean within that sum's rounding bound, (n/4 + 4) u mean|x| + u |mean|, u = 2^-24. Given
the device's centre, everything after it is exact: the oracle built over the same
centred rows has identical lists, and searches return identical top-k.

Non-finite rows: NaN projections never enter a list (rows with a NaN or an infinite
entry project to NaN); lists equal the oracle's, and searches over such an index equal
the oracle's.
"""
import numpy as np
import pytest

from tests._parity import compare_list

pytestmark = pytest.mark.gpu


def _data(n, d, seed, shift=2.0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, d)) * rng.lognormal(0, 0.5, (n, 1)) + shift).astype(np.float32)
    Q = rng.standard_normal((64, d)).astype(np.float32)
    return X, Q / np.linalg.norm(Q, axis=1, keepdims=True)


def _centre_bound(X):
    n = X.shape[0]
    m64 = X.astype(np.float64).mean(0)
    return m64, (n / 4 + 4) * 2.0 ** -24 * np.abs(X).astype(np.float64).mean(0) + 2.0 ** -24 * np.abs(m64)


def test_stream_centering_against_oracle(CEOs, oracle):
    d, D, R, m = 96, 256, 3, 80
    X, Q = _data(60000, d, 1)
    Xn, _ = _data(9000, d, 2)
    s = CEOs.streamCEOs(d)
    s.setIndexParam(D, R, m, -1, 42)
    s.centering = True
    s.build(X)
    c = s.center
    m64, bound = _centre_bound(X)
    assert np.all(np.abs(c - m64) <= 2.0 ** -24 * np.abs(m64) + 1e-30)      # one rounding of the FP64 mean
    assert np.all(np.abs(c - oracle.column_mean(X)) <= 2 * bound)           # the reference's FP32 mean
    o = oracle.OracleStream(d)
    o.setIndexParam(D, R, m, 8, 42)
    o.set_centering(True, c)
    o.build(X)
    np.testing.assert_array_equal(o.center(), c)
    for step in range(2):
        g = s.get_buckets(D * R, m)
        b = o.buckets()
        for col in range(D * R):
            compare_list(g[0][col, :g[4][col]], g[1][col, :g[4][col]], b[0][col, :b[4][col]], b[1][col, :b[4][col]])
            compare_list(g[2][col, :g[5][col]], g[3][col, :g[5][col]], b[2][col, :b[5][col]], b[3][col, :b[5][col]])
        for h in (s, o):
            h.n_probed_vectors, h.n_probed_points, h.n_cand = 30, m, 400
        gi, gp = s.estimate_search(Q, 10)
        oi, op = o.estimate_search(Q, 10)
        assert np.array_equal(gi, oi), step
        np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-5)
        s.update(Xn[step * 4500:(step + 1) * 4500], 3000)   # inserts are centred with the build's centre
        o.update(Xn[step * 4500:(step + 1) * 4500], 3000)
    np.testing.assert_array_equal(s.center, c)


def test_hash_centering_against_oracle(CEOs, oracle):
    d, D, R, m = 64, 256, 2, 100
    X, Q = _data(40000, d, 3, shift=-1.5)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(D, R, m, -1, 42)
    g.centering = True
    g.build_CEOs_Hash(X)
    c = g.center
    m64, bound = _centre_bound(X)
    assert np.all(np.abs(c - oracle.column_mean(X)) <= 2 * bound)
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(D, R, m, 8, 42)
    o.centering = True
    o.build_CEOs_Hash(X, center=c)
    for a, b in zip(g.get_buckets(), o.buckets()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for h in (g, o):
        h.n_probed_vectors, h.n_probed_points = 20, 60
    gi, gp = g.search_CEOs_Hash(Q, 10)
    oi, op = o.search_CEOs_Hash(Q, 10)
    assert np.array_equal(gi, oi)
    np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-5)
    with pytest.raises(RuntimeError):
        CEOs.CEOs(10, d).center  # no centre before a centred build


def test_nan_and_inf_rows_against_oracle(CEOs, oracle):
    n, d, D, R, m = 30000, 48, 256, 2, 60
    X, Q = _data(n, d, 4, shift=0.0)
    rng = np.random.default_rng(5)
    bad = rng.choice(n, 40, replace=False)
    X[bad[:20], 3] = np.nan
    X[bad[20:30], 7] = np.inf
    X[bad[30:], 1] = -np.inf
    X[bad[:3] * 0 + np.arange(3), 0] = np.nan  # NaN rows among the first ones, while lists fill
    g = CEOs.CEOs(n, d)
    g.setIndexParam(D, R, m, -1, 42)
    g.rerank_bounds = False  # the int8 copies of non-finite rows are not meaningful either way
    g.build_coCEOs_Est(X)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 8, 42)
    o.build_coCEOs_Est(X)
    gb, ob = g.get_buckets(), o.buckets()
    for col in range(D * R):
        compare_list(gb[0][col], gb[1][col], ob[0][col], ob[1][col])
        compare_list(gb[2][col], gb[3][col], ob[2][col], ob[3][col])
    nonfinite = np.unique(np.concatenate([bad, np.arange(3)]))
    assert not np.isin(gb[0], nonfinite).any() and not np.isin(gb[2], nonfinite).any()
    for h in (g, o):
        h.n_probed_vectors, h.n_cand = 30, 500
    gi, gp = g.search_coCEOs_Est(Q, 10)
    oi, op = o.search_coCEOs_Est(Q, 10)
    assert np.array_equal(gi, oi)
    np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-5)


def test_printed_lines_follow_the_reference(CEOs, capfd):
    """Build, update and verbose search print the reference's lines (src/CEOs.cpp:313-328,
    434-440, 466-470, 676-680; src/streamCEOs.cpp:11-37, 154-160, 212-215, 362-368)."""
    X, Q = _data(20000, 32, 6)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(128, 2, 50, -1, 42)
    g.build_coCEOs_Est(X)
    g.search_coCEOs_Est(Q, 5, verbose=True)
    out = capfd.readouterr().out
    for line in ("Building coCEOs-Estimate index...", "n_points: 20000", "fhtDim: 128", "Copying data time (in seconds)",
                 "Size of coCEOs-Est index in GB: ", "Projecting time: ", "Extracting top-m points time: ",
                 "Constructing time (in seconds)", "n_queries: 64", "n_threads: ",
                 "Projecting and extracting top-vectors time: ", "Querying time: "):
        assert line in out, line
    gb = 1.0 * (4 * 50 * 256 + 20000 * 32) * 4 / 2 ** 30
    assert f"Size of coCEOs-Est index in GB: {gb:g}" in out
    s = CEOs.streamCEOs(32)
    s.setIndexParam(128, 2, 50, -1, 42)
    s.build(X)
    s.update(X[:500], 300)
    s.n_probed_vectors = 10
    s.estimate_search(Q, 5, verbose=True)
    out = capfd.readouterr().out
    for line in ("n_proj: 128", "top_m: 50", "centering: 0", "n_points: 20000", "Copying data time (in ms)",
                 "Extracting top-points time (in ms)", "Constructing time (in seconds)",
                 "n_points (after update): 20200", "Updating data time (in ms)", "Updating time (in ms)",
                 "n_probedVectors: 10", "Projecting time: ", "Querying time: "):
        assert line in out, line
