"""Index state and stream ordering: searches of the coCEOs lists on a handle that holds a
CEOs-Est index (dense projections, no lists) are refused on the host; device searches
queued on different caller streams without synchronisation, and an update queued
behind them, give the results of the same calls made one at a time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(n, d, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, d)) * rng.lognormal(0, 0.5, (n, 1))).astype(np.float32)


def test_list_searches_refused_on_a_ceos_est_index(CEOs):
    X = _rand(3000, 32, 1)
    Q = _rand(8, 32, 2)
    g = CEOs.CEOs(*X.shape)
    g.setIndexParam(128, 2, 40, -1, 42)
    g.build_coCEOs_Est(X)
    plan = g.make_plan(16, 10, 20, 100, autotune=False)
    g.build(X)  # CEOs-Est: the lists are released
    g.n_probed_vectors, g.n_cand = 10, 50
    ids, _ = g.search(Q, 5)  # the dense search works
    assert ids.shape == (8, 5)
    with pytest.raises(RuntimeError, match="not built"):
        g.search_coCEOs_Est(Q, 5)
    with pytest.raises(RuntimeError, match="not built"):
        g.search_CEOs_Hash(Q, 5)
    with pytest.raises(RuntimeError, match="not built"):
        g.get_buckets()
    k = np.full(8, 5, np.int32)
    s = np.full(8, 10, np.int32)
    c = np.full(8, 50, np.int32)
    with pytest.raises(RuntimeError, match="not built"):
        plan.search(Q, k, s, c)  # a plan made on the coCEOs index
    with pytest.raises(RuntimeError, match="not built"):
        g.make_plan(16, 10, 20, 100, autotune=False).search(Q, k, s, c)
    with pytest.raises(RuntimeError, match="not built"):
        CEOs.Server(g, max_batch=16, k_max=10, s_max=20, c_max=100, inflight=1)
    with pytest.raises(RuntimeError, match="not built"):
        g.search_batched(Q, k, s, c)
    g.build_coCEOs_Est(X)  # lists back: the old plan works again
    pi, _ = plan.search(Q, k, s, c)
    assert np.array_equal(pi[:, :5], g.search_coCEOs_Est(Q, 5)[0])


def test_device_searches_on_two_streams_match_one_at_a_time():
    from ceos_amd import capi
    from tests._hip import DeviceArray, Stream, synchronize

    n, d, nq, k = 40000, 64, 600, 10
    X = _rand(n, d, 3)
    Q = _rand(nq, d, 4)
    idx = capi.DeviceIndex(n, d)
    idx.set_index_param(256, 4, 200, -1, 42)
    dX = DeviceArray(X)
    dQ = DeviceArray(Q)
    idx.build_device(dX.ptr, n)
    idx.set_query_param(40, 200, 1000)

    def run(stream, q0, q1, out_i, out_p):
        idx.search_device(dQ.at(q0), q1 - q0, k, out_i.at(q0), out_p.at(q0), stream)

    # one at a time, synchronised
    ri = DeviceArray(shape=(nq, k), dtype=np.int32)
    rp = DeviceArray(shape=(nq, k), dtype=np.float32)
    for q0 in range(0, nq, 100):
        run(0, q0, q0 + 100, ri, rp)
        synchronize()
    ref_i, ref_p = ri.numpy(), rp.numpy()
    # alternating between two caller streams and the index's own, no synchronisation
    sa, sb = Stream(), Stream()
    for rep in range(3):
        gi = DeviceArray(shape=(nq, k), dtype=np.int32)
        gp = DeviceArray(shape=(nq, k), dtype=np.float32)
        gi.fill(0)
        gp.fill(0)
        for j, q0 in enumerate(range(0, nq, 100)):
            st = (sa.handle, sb.handle, 0)[j % 3]
            run(st, q0, q0 + 100, gi, gp)
        synchronize()
        assert np.array_equal(gi.numpy(), ref_i), rep
        assert np.array_equal(gp.numpy().view(np.uint32), ref_p.view(np.uint32)), rep


def test_update_waits_for_a_queued_plan_run(CEOs):
    """A plan replay queued on a caller stream and an update made right after it: the
    replay reads the index as it was, so its results equal those of the synchronised
    search before the update."""
    import ctypes

    from tests._hip import Stream, hip, synchronize

    X = _rand(50000, 48, 5)
    Q = _rand(2000, 48, 6)
    s = CEOs.streamCEOs(48)
    s.setIndexParam(256, 4, 300, -1, 42)
    s.build(X)
    nq = len(Q)
    k = np.full(nq, 10, np.int32)
    sv = np.full(nq, 60, np.int32)
    c = np.full(nq, 3000, np.int32)
    plan = s.make_plan(nq, 10, 60, 3000, autotune=False)
    ref_i, ref_p = plan.search(Q, k, sv, c)  # also leaves Q, k, s, c in the plan's buffers
    buf = plan.buffers()
    L = hip()
    st = Stream()
    for rep in range(2):
        # clear the outputs, queue the replay, then update at once (no synchronisation)
        zero = np.zeros((nq, 10), np.int32)
        assert L.hipMemcpy(buf["ids"], zero.ctypes.data, zero.nbytes, 1) == 0
        plan.run(nq, st.handle)
        s.update(_rand(5000, 48, 10 + rep), 5000)
        synchronize()
        gi = np.zeros((nq, 10), np.int32)
        gp = np.zeros((nq, 10), np.float32)
        assert L.hipMemcpy(gi.ctypes.data, buf["ids"], gi.nbytes, 2) == 0
        assert L.hipMemcpy(gp.ctypes.data, buf["ips"], gp.nbytes, 2) == 0
        assert np.array_equal(gi, ref_i), rep
        assert np.array_equal(gp.view(np.uint32), ref_p.view(np.uint32)), rep
        ref_i, ref_p = plan.search(Q, k, sv, c)  # the updated index, for the next round
