"""Captured search plans (hipGraph replay of the batched search): every replay equals the
direct batched search, with eager work between replays; autotuning keeps results; a
plan re-captures itself when its index changes; stream plans follow updates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(n, d, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, d)).astype(np.float32)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return (g * rng.lognormal(0, sigma, (n, 1))).astype(np.float32)


def _params(nq, seed, k_choices=(1, 10, 100), s_rng=(10, 81), c_rng=(50, 2001)):
    rng = np.random.default_rng(seed)
    k = rng.choice(np.array(k_choices, np.int32), nq).astype(np.int32)
    s = rng.integers(*s_rng, nq).astype(np.int32)
    c = rng.integers(*c_rng, nq).astype(np.int32)
    return k, s, c


def _rows(ids, ips, off, nq, k_max):
    """batched (flat, offsets) -> fixed-stride rows zero padded to k_max."""
    I = np.zeros((nq, k_max), np.int32)
    P = np.zeros((nq, k_max), np.float32)
    for q in range(nq):
        a, b = off[q], off[q + 1]
        I[q, : b - a] = ids[a:b]
        P[q, : b - a] = ips[a:b]
    return I, P


@pytest.fixture(scope="module")
def index(CEOs):
    X = _rand(30000, 64, 31, 0.7)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(256, 4, 150, -1, 42)
    g.build_coCEOs_Est(X)
    return X, g


def test_plan_replays_match_direct_batched(index):
    X, g = index
    Q = _rand(500, 64, 32)
    k, s, c = _params(len(Q), 33)
    plan = g.make_plan(500, 100, 80, 2000, autotune=False)
    ids, ips, off = g.search_batched(Q, k, s, c)
    I0, P0 = _rows(ids, ips, off, len(Q), 100)
    for rep in range(4):
        pi, pp = plan.search(Q, k, s, c)
        assert np.array_equal(pi, I0), rep
        assert np.array_equal(pp.view(np.uint32), P0.view(np.uint32)), rep
        # eager work between replays
        g.search_coCEOs_Est(Q[:7], 5)
    # a smaller batch through the same graph, then the full one again
    pi, pp = plan.search(Q[:123], k[:123], s[:123], c[:123])
    assert np.array_equal(pi, I0[:123]) and np.array_equal(pp, P0[:123])
    pi, pp = plan.search(Q, k, s, c)
    assert np.array_equal(pi, I0)


def test_plan_autotune_keeps_results(index):
    X, g = index
    Q = _rand(400, 64, 34)
    k, s, c = _params(len(Q), 35)
    plan = g.make_plan(400, 100, 80, 2000, autotune=True)
    ids, ips, off = g.search_batched(Q, k, s, c)
    I0, P0 = _rows(ids, ips, off, len(Q), 100)
    pi, pp = plan.search(Q, k, s, c)  # first run tunes
    wg, tl, ms = plan.config()
    assert ms > 0 and wg >= 1 and (tl == 0 or 8 <= tl <= 15)  # 0: the automatic tile
    # the tuner's record: every shape it timed (median of 3 launches); the pick is the
    # fastest of them, and both projection widths were tried with the picked shape
    rec = plan.tuning()
    assert len(rec) >= 4 and all(t > 0 for *_, t in rec)
    assert ms == min(t for *_, t in rec)
    assert {pw for *_, pw, _ in rec} >= {2, 4, 8}
    assert (wg, tl, plan.proj_waves()) in {(a, b, c_) for a, b, c_, _ in rec}
    assert np.array_equal(pi, I0) and np.array_equal(pp, P0)
    pi, pp = plan.search(Q, k, s, c)
    assert np.array_equal(pi, I0) and np.array_equal(pp, P0)


def test_plan_recaptures_after_rebuild(CEOs):
    X1 = _rand(8000, 48, 36)
    X2 = _rand(9000, 48, 37)
    Q = _rand(64, 48, 38)
    k, s, c = _params(len(Q), 39, k_choices=(10,), s_rng=(5, 30), c_rng=(20, 400))
    g = CEOs.CEOs(*X1.shape)
    g.setIndexParam(128, 2, 60, -1, 42)
    g.build_coCEOs_Est(X1)
    plan = g.make_plan(64, 10, 30, 400, autotune=False)
    a1, _ = plan.search(Q, k, s, c)
    g.build_coCEOs_Est(X2)  # new points and buckets: the captured pointers are stale
    a2, p2 = plan.search(Q, k, s, c)
    ids, ips, off = g.search_batched(Q, k, s, c)
    I2, P2 = _rows(ids, ips, off, len(Q), 10)
    assert np.array_equal(a2, I2) and np.array_equal(p2, P2)
    assert not np.array_equal(a1, a2)


def test_plan_bounds_and_errors(index):
    X, g = index
    plan = g.make_plan(10, 5, 20, 100, autotune=False)
    Q = _rand(11, 64, 40)
    k = np.full(11, 5, np.int32)
    s = np.full(11, 10, np.int32)
    c = np.full(11, 50, np.int32)
    with pytest.raises(ValueError):
        plan.search(Q, k, s, c)  # 11 > nq_max
    k[0] = 6
    with pytest.raises(ValueError):
        plan.search(Q[:10], k[:10], s[:10], c[:10])  # k above k_max
    with pytest.raises(ValueError):
        g.make_plan(0, 5, 20, 100)


def test_stream_plan_follows_updates(CEOs):
    rng = np.random.default_rng(41)
    d, D, R, m = 40, 128, 2, 50
    X = _rand(6000, d, 42)
    Q = _rand(50, d, 43)
    st = CEOs.streamCEOs(d)
    st.setIndexParam(D, R, m, -1, 42)
    st.build(X[:3000])
    st.n_probed_vectors, st.n_probed_points, st.n_cand = 20, m, 200
    plan = st.make_plan(50, 10, 20, 200, autotune=False)
    k = np.full(50, 10, np.int32)
    s = np.full(50, 20, np.int32)
    c = np.full(50, 200, np.int32)
    for step in range(3):
        pi, pp = plan.search(Q, k, s, c)
        ei, ep = st.estimate_search(Q, 10)
        assert np.array_equal(pi, ei), step
        assert np.array_equal(pp, ep), step
        st.update(X[3000 + step * 1000:4000 + step * 1000], int(rng.integers(0, 800)))
