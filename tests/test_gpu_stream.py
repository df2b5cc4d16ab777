"""GPU parity of streamCEOs (build / update / estimate_search) against the oracle,
and the random-delete extension against its closed form."""
import numpy as np
import pytest

from tests._parity import compare_list

pytestmark = pytest.mark.gpu


def _rand(n, d, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    return (g * rng.lognormal(0, 0.5, (n, 1))).astype(np.float32)


def _pair(CEOs, oracle, d, D, R, m):
    g = CEOs.streamCEOs(d)
    g.setIndexParam(D, R, m, -1, 42)
    o = oracle.OracleStream(d)
    o.setIndexParam(D, R, m, 4, 42)
    return g, o


def _same_buckets(g, o, D, R, m):
    gcid, gcv, gfid, gfv, gcs, gfs = g.get_buckets(D * R, m)
    ocid, ocv, ofid, ofv, ocs, ofs = o.buckets()
    assert np.array_equal(gcs, ocs) and np.array_equal(gfs, ofs)
    for col in range(D * R):
        a, b = gcs[col], gfs[col]
        compare_list(gcid[col, :a], gcv[col, :a], ocid[col, :a], ocv[col, :a])
        compare_list(gfid[col, :b], gfv[col, :b], ofid[col, :b], ofv[col, :b])


def _same_search(g, o, Q, k, s, c):
    for obj in (g, o):
        obj.n_probed_vectors, obj.n_cand = s, c
    gi, gp = g.estimate_search(Q, k)
    oi, op = o.estimate_search(Q, k)
    assert np.array_equal(gi, oi)
    np.testing.assert_allclose(gp, op, rtol=1e-5, atol=1e-6)


def test_stream_sliding_window(CEOs, oracle):
    """The reference's own stream benchmark pattern (test/utils.py:325-375): build on a
    window, then update(new window, n_del = window) repeatedly."""
    d, D, R, m = 48, 128, 2, 40
    win = 2500
    X = _rand(win * 5, d, 11)
    Q = _rand(60, d, 12)
    g, o = _pair(CEOs, oracle, d, D, R, m)
    g.build(X[:win]); o.build(X[:win])
    g.n_probed_points = o.n_probed_points = m
    _same_buckets(g, o, D, R, m)
    _same_search(g, o, Q, 10, 12, 150)
    for i in range(1, 5):
        new = X[i * win:(i + 1) * win]
        g.update(new, win); o.update(new, win)
        assert g.size() == o.size() == win
        _same_buckets(g, o, D, R, m)
        _same_search(g, o, Q, 10, 12, 150)


def test_stream_partial_deletes_ring_growth(CEOs, oracle):
    d, D, R, m = 33, 64, 3, 25   # d not a multiple of 4
    X = _rand(9000, d, 13)
    Q = _rand(40, d, 14)
    g, o = _pair(CEOs, oracle, d, D, R, m)
    g.build(X[:1000]); o.build(X[:1000])
    g.n_probed_points = o.n_probed_points = m
    pos = 1000
    for n_new, n_del in ((300, 100), (2000, 0), (50, 900), (3000, 1500), (10, 0), (700, 2000)):
        g.update(X[pos:pos + n_new], n_del); o.update(X[pos:pos + n_new], n_del)
        pos += n_new
        assert g.size() == o.size()
        _same_buckets(g, o, D, R, m)
        _same_search(g, o, Q, 7, 9, 80)


def test_stream_errors(CEOs):
    g = CEOs.streamCEOs(8)
    g.setIndexParam(16, 1, 10, -1, 42)
    g.build(_rand(20, 8, 1))
    with pytest.raises(ValueError):
        g.update(_rand(1, 8, 2), 25)
    with pytest.raises(ValueError):
        g.update(_rand(1, 8, 2), 15)
    g.n_probed_vectors = 17  # > n_proj (src/streamCEOs.cpp:386)
    with pytest.raises(ValueError):
        g.estimate_search(_rand(2, 8, 3), 3)


def test_stream_random_delete(CEOs, oracle):
    """delete_ids removes the points from every bucket at once (order kept) and they are
    never returned; other ids do not move."""
    d, D, R, m = 24, 64, 2, 30
    X = _rand(3000, d, 15)
    Q = _rand(50, d, 16)
    g = CEOs.streamCEOs(d)
    g.setIndexParam(D, R, m, -1, 42)
    g.build(X)
    before = g.get_buckets(D * R, m)
    rng = np.random.default_rng(17)
    dead = rng.choice(3000, 150, replace=False).astype(np.int32)
    g.delete_ids(dead)
    after = g.get_buckets(D * R, m)
    dset = set(dead.tolist())
    for (ids0, v0), (ids1, v1), sz in (((before[0], before[1]), (after[0], after[1]), after[4]),
                                       ((before[2], before[3]), (after[2], after[3]), after[5])):
        for col in range(D * R):
            keep = [t for t in range(m) if ids0[col, t] not in dset]
            assert sz[col] == len(keep)
            assert np.array_equal(ids1[col, :sz[col]], ids0[col, keep])
            assert np.array_equal(v1[col, :sz[col]], v0[col, keep])
    g.n_probed_vectors, g.n_probed_points, g.n_cand = 16, m, 300
    ids, _ = g.estimate_search(Q, 10)
    assert not (set(ids[ids != 0].ravel().tolist()) & dset)
    # a FIFO update afterwards still agrees with the closed form of the survivors
    g.update(_rand(500, d, 18), 100)
    assert g.size() == 3400


def test_stream_centering_equals_precentered_data(CEOs):
    """centering stores rows minus the build rows' column mean, fixed for updates
    (src/streamCEOs.cpp:22-34, 198-203): the same as feeding pre-centered rows."""
    d, D, R, m = 24, 64, 2, 30
    X = _rand(6000, d, 19) + 2.5
    Q = _rand(30, d, 20)
    c = X[:3000].astype(np.float64).mean(axis=0).astype(np.float32)
    a = CEOs.streamCEOs(d)
    a.centering = True
    assert a.centering
    b = CEOs.streamCEOs(d)
    for g in (a, b):
        g.setIndexParam(D, R, m, -1, 42)
    a.build(X[:3000])
    b.build((X[:3000] - c).astype(np.float32))
    for g in (a, b):
        g.n_probed_vectors, g.n_probed_points, g.n_cand = 10, m, 100
    ia, pa = a.estimate_search(Q, 8)
    ib, pb = b.estimate_search(Q, 8)
    assert np.array_equal(ia, ib) and np.array_equal(pa, pb)
    a.update(X[3000:4500], 1000)
    b.update((X[3000:4500] - c).astype(np.float32), 1000)
    ia, pa = a.estimate_search(Q, 8)
    ib, pb = b.estimate_search(Q, 8)
    assert np.array_equal(ia, ib) and np.array_equal(pa, pb)
