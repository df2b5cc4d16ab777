"""Shared comparison helpers for the GPU parity tests."""
import numpy as np


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def compare_list(ids_g, v_g, ids_o, v_o):
    """One bucket list: values must agree bit for bit position by position; ids may differ
    only inside runs of exactly equal values (the reference's heap orders ties
    arbitrarily). Returns the number of positions whose id differs."""
    assert np.array_equal(bits(v_g), bits(v_o)), "bucket values differ"
    if np.array_equal(ids_g, ids_o):
        return 0
    n = len(v_o)
    t = 0
    while t < n:
        u = t
        while u + 1 < n and v_o[u + 1] == v_o[t]:
            u += 1
        if not np.array_equal(ids_g[t:u + 1], ids_o[t:u + 1]):
            assert u > t, f"id mismatch at slot {t} without a value tie"
            if u + 1 < n:  # a tie run inside the list holds the same ids
                assert set(ids_g[t:u + 1].tolist()) == set(ids_o[t:u + 1].tolist())
        t = u + 1
    return int((ids_g != ids_o).sum())


def compare_buckets(gb, ob, sizes=None):
    """gb/ob: (close_ids, close_vals, far_ids, far_vals) arrays [lists][m]."""
    mism = 0
    for ids_g, v_g, ids_o, v_o in ((gb[0], gb[1], ob[0], ob[1]), (gb[2], gb[3], ob[2], ob[3])):
        for col in range(ids_o.shape[0]):
            lim = ids_o.shape[1] if sizes is None else None
            mism += compare_list(ids_g[col, :lim], v_g[col, :lim], ids_o[col, :lim], v_o[col, :lim])
    return mism


def tie_runs(v):
    """Number of positions of a sorted list that sit in a run of equal values."""
    v = np.asarray(v)
    if len(v) < 2:
        return 0
    eq = v[1:] == v[:-1]
    mark = np.zeros(len(v), bool)
    mark[1:] |= eq
    mark[:-1] |= eq
    return int(mark.sum())


def topk_verdict(q, x_rows, ids_g, ips_g, ids_o, ips_o, rtol=1e-5):
    """Compares one query's top-k from the device with the oracle's.

    'same': ids identical. 'tie': the ids differ only among points whose exact inner
    products (FP64) sit within rtol of the k-th value, so the FP32 rerank and the oracle's
    FP64-rounded dot may order them either way; the value sequences agree within rtol.
    'diff' otherwise. x_rows(ids) returns the rows of the given ids."""
    ids_g = np.asarray(ids_g)
    ids_o = np.asarray(ids_o)
    if np.array_equal(ids_g, ids_o):
        return "same"
    if len(ids_g) == 0 or not np.allclose(ips_g, ips_o, rtol=rtol, atol=1e-6):
        return "diff"
    sym = np.setxor1d(ids_g, ids_o)
    if len(sym) == 0:  # same set, order differs: equal values swapped
        return "tie"
    exact = x_rows(sym).astype(np.float64) @ np.asarray(q, np.float64)
    kth = float(min(ips_o[-1], ips_g[-1]))
    tol = rtol * abs(kth) + 1e-6
    return "tie" if np.all(np.abs(exact - kth) <= tol + rtol * np.abs(exact)) else "diff"
