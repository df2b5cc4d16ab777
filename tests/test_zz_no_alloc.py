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


def _search(L, idx, fn, dQ, nq, d, k, torch, dev):
    ids = torch.full((nq, k), -7, dtype=torch.int32, device=dev)
    ips = torch.full((nq, k), -7.0, dtype=torch.float32, device=dev)
    rc = getattr(L, fn)(idx.h, dQ.data_ptr(), nq, d, k, ids.data_ptr(), ips.data_ptr(), None)
    assert rc == 0, L.ceos_last_error()
    torch.cuda.synchronize()
    return ids.cpu().numpy(), ips.cpu().numpy()


@pytest.mark.parametrize("n,d,D,R,m,s,c,k,nq", [
    (60000, 128, 1024, 4, 1000, 40, 1000, 10, 3000),   # chunks of queries, narrowed grid
    (20000, 300, 1024, 2, 100, 20, 500, 10, 700),
    (30000, 960, 1024, 8, 500, 80, 5000, 100, 400),    # the largest per-workgroup share
])
def test_noalloc_equals_allocating_search(n, d, D, R, m, s, c, k, nq):
    import torch

    from ceos_amd import capi

    dev = torch.device("cuda:0")
    L = capi.lib()
    X = _rand(n, d, 1)
    Q = _rand(nq, d, 2)
    idx = capi.DeviceIndex(n, d)
    idx.set_index_param(D, R, m, -1, 42)
    dX = torch.from_numpy(X).to(dev)
    dQ = torch.from_numpy(Q).to(dev)
    idx.build_device(dX.data_ptr(), n)
    idx.set_query_param(s, m, c)
    for nqi in (nq, nq // 3 + 1):  # two calls in a row, the second smaller
        print(f"noalloc case n={n} d={d} nq={nqi}", flush=True)
        ri, rp = _search(L, idx, "ceos_search_coceos_device", dQ, nqi, d, k, torch, dev)
        gi, gp = _search(L, idx, "ceos_search_coceos_device_noalloc", dQ, nqi, d, k, torch, dev)
        assert np.array_equal(gi, ri)
        assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))


def test_noalloc_rejects_oversized_scratch():
    import torch

    from ceos_amd import capi

    dev = torch.device("cuda:0")
    L = capi.lib()
    n, d = 5000, 16
    X = _rand(n, d, 3)
    idx = capi.DeviceIndex(n, d)
    idx.set_index_param(1024, 4, 4096, -1, 42)
    dX = torch.from_numpy(X).to(dev)
    idx.build_device(dX.data_ptr(), n)
    idx.set_query_param(4096, 4096, 100)  # 2^24 probed positions: 256 MiB for one workgroup
    dQ = torch.from_numpy(X[:4]).to(dev)
    ids = torch.zeros((4, 10), dtype=torch.int32, device=dev)
    ips = torch.zeros((4, 10), dtype=torch.float32, device=dev)
    rc = L.ceos_search_coceos_device_noalloc(idx.h, dQ.data_ptr(), 4, d, 10, ids.data_ptr(), ips.data_ptr(), None)
    assert rc == 4 and b"static scratch" in L.ceos_last_error()
