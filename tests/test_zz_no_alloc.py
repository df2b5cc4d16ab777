"""The allocation-free device search (ceos_search_coceos_device_noalloc, static device
scratch) returns what the allocating device search returns, bit for bit, on two calls in
a row (the second with fewer queries), and rejects a search whose scratch share does not
fit the static array before launching anything."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(n, d, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, d)) * rng.lognormal(0, 0.5, (n, 1))).astype(np.float32)


def _search(L, idx, fn, dQ, nq, d, k):
    from tests._hip import DeviceArray, synchronize

    ids = DeviceArray(np.full((nq, k), -7, np.int32))
    ips = DeviceArray(np.full((nq, k), -7.0, np.float32))
    rc = getattr(L, fn)(idx.h, dQ.ptr, nq, d, k, ids.ptr, ips.ptr, None)
    assert rc == 0, L.ceos_last_error()
    synchronize()
    return ids.numpy(), ips.numpy()


@pytest.mark.parametrize("n,d,D,R,m,s,c,k,nq", [
    (60000, 128, 1024, 4, 1000, 40, 1000, 10, 3000),   # chunks of queries, narrowed grid
    (20000, 300, 1024, 2, 100, 20, 500, 10, 700),
    (30000, 960, 1024, 8, 500, 80, 5000, 100, 400),    # the largest per-workgroup share
])
def test_noalloc_equals_allocating_search(n, d, D, R, m, s, c, k, nq):
    from ceos_amd import capi
    from tests._hip import DeviceArray

    L = capi.lib()
    X = _rand(n, d, 1)
    Q = _rand(nq, d, 2)
    idx = capi.DeviceIndex(n, d)
    idx.set_index_param(D, R, m, -1, 42)
    dX = DeviceArray(X)
    dQ = DeviceArray(Q)
    idx.build_device(dX.ptr, n)
    idx.set_query_param(s, m, c)
    for nqi in (nq, nq // 3 + 1):  # two calls in a row, the second smaller
        print(f"noalloc case n={n} d={d} nq={nqi}", flush=True)
        ri, rp = _search(L, idx, "ceos_search_coceos_device", dQ, nqi, d, k)
        gi, gp = _search(L, idx, "ceos_search_coceos_device_noalloc", dQ, nqi, d, k)
        assert np.array_equal(gi, ri)
        assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))


def test_noalloc_rejects_oversized_scratch():
    from ceos_amd import capi
    from tests._hip import DeviceArray

    L = capi.lib()
    n, d = 5000, 16
    X = _rand(n, d, 3)
    idx = capi.DeviceIndex(n, d)
    idx.set_index_param(1024, 4, 4096, -1, 42)
    dX = DeviceArray(X)
    idx.build_device(dX.ptr, n)
    idx.set_query_param(4096, 4096, 100)  # 2^24 probed positions: 256 MiB for one workgroup
    dQ = DeviceArray(X[:4])
    ids = DeviceArray(shape=(4, 10), dtype=np.int32)
    ips = DeviceArray(shape=(4, 10), dtype=np.float32)
    rc = L.ceos_search_coceos_device_noalloc(idx.h, dQ.ptr, 4, d, 10, ids.ptr, ips.ptr, None)
    assert rc == 4 and b"static scratch" in L.ceos_last_error()
