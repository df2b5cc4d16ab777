"""streamCEOs at the 5,000,000 x 384 configuration (N(0,I) rows, LogNormal(0,0.5) norms;
D=512, R=8, m=1000) with device-resident buckets under inserts, FIFO deletes, random
deletes and serving:
  * build: 32 lists (16 columns of two repeats, close and far) equal the oracle's build;
  * update(10,000 new rows, n_delPoints=10,000): 512 sampled lists equal the closed form
    of src/streamCEOs.cpp:278-351, top_m(old entries with id >= n_del, renumbered by
    -n_del, U the new rows projected by the oracle);
  * random delete of 0.5% of the live ids: every list equals the old list without them;
  * a second FIFO update on the shortened lists: the closed form again;
  * estimate_search on 300 queries against the oracle's search over the live rows and the
    device lists (at least 99.9% identical top-10, the rest reordered only among points
    within 1e-5 of the 10th inner product; logged);
  * a server across updates: every served result equals estimate_search on the index
    state it was served from.
"""
import sys
import time

import numpy as np
import pytest

from tests._parity import bits, compare_list, topk_verdict

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[stream_5m] {msg}", file=sys.stderr, flush=True)


def _closed_form_update(old, sizes, newproj, n_cur, n_del, m, sign):
    """top_m of the valid old entries with id >= n_del (renumbered) and the new rows."""
    ids, vals = old
    keep = ids[:sizes] >= n_del
    ci = np.concatenate([ids[:sizes][keep] - n_del, n_cur - n_del + np.arange(len(newproj), dtype=np.int64)])
    cvv = np.concatenate([vals[:sizes][keep], sign * newproj]).astype(np.float32)
    order = np.lexsort((ci, -cvv.astype(np.float64)))[:m]
    return ci[order].astype(np.int32), cvv[order]


def test_stream_5m(CEOs, oracle):
    from ceos_amd.workloads import WORKLOADS, make_points, make_queries

    w = WORKLOADS["stream_5m"]
    d, DR, m = w.d, w.D * w.R, w.m
    t = time.time()
    X0 = make_points(w)
    say(f"data {X0.shape} in {time.time() - t:.1f}s")
    s = CEOs.streamCEOs(d)
    s.setIndexParam(w.D, w.R, m, -1, w.seed)
    t = time.time()
    s.build(X0)
    say(f"device build (with the host copy) in {time.time() - t:.2f}s")
    o = oracle.OracleCEOs(*X0.shape)
    o.setIndexParam(w.D, w.R, m, 16, w.seed)
    rng = np.random.default_rng(12)

    # build: lists of 16 columns in each of two repeats
    cid, cv, fid, fv, cs, fs = s.get_buckets(DR, m)
    assert (cs == m).all() and (fs == m).all()
    cols = np.concatenate([r * w.D + rng.choice(w.D, 16, replace=False) for r in (1, 6)])
    t = time.time()
    oc, ocv, of, ofv = o.bucket_columns(X0, cols)
    say(f"oracle lists of {len(cols)} columns in {time.time() - t:.1f}s")
    for j, col in enumerate(cols):
        compare_list(cid[col], cv[col], oc[j], ocv[j])
        compare_list(fid[col], fv[col], of[j], ofv[j])
    say(f"build: {2 * len(cols)} lists identical to the oracle's")

    live = [X0]  # live rows in id order = concatenation minus the FIFO-dropped prefix
    dropped = 0

    def fifo_update(seed):
        nonlocal cid, cv, fid, fv, cs, fs, dropped
        Xn = make_points(w, 10000, seed=seed)
        n_cur = sum(len(a) for a in live) - dropped
        t = time.time()
        s.update(Xn, 10000)
        say(f"update (+10,000 rows, n_delPoints 10,000) in {(time.time() - t) * 1e3:.1f} ms")
        live.append(Xn)
        dropped += 10000
        P = o.project(Xn)
        ncid, ncv, nfid, nfv, ncs, nfs = s.get_buckets(DR, m)
        for col in rng.choice(DR, 256, replace=False):
            for side in (0, 1):
                oid, ov, osz = (cid, cv, cs) if side == 0 else (fid, fv, fs)
                gid, gv, gsz = (ncid, ncv, ncs) if side == 0 else (nfid, nfv, nfs)
                ei, ev = _closed_form_update((oid[col], ov[col]), int(osz[col]), P[:, col], n_cur, 10000, m,
                                             1.0 if side == 0 else -1.0)
                assert gsz[col] == len(ei), (col, side)
                compare_list(gid[col, :len(ei)], gv[col, :len(ei)], ei, ev)
        cid, cv, fid, fv, cs, fs = ncid, ncv, nfid, nfv, ncs, nfs
        say("update: 512 sampled lists equal the closed form")

    fifo_update(21)

    # random delete of 0.5% of the live ids: each list loses exactly those ids
    n_live = s.size()
    dels = np.sort(rng.choice(n_live, n_live // 200, replace=False)).astype(np.int32)
    t = time.time()
    s.delete_ids(dels)
    say(f"random delete of {len(dels)} ids in {(time.time() - t) * 1e3:.1f} ms")
    ncid, ncv, nfid, nfv, ncs, nfs = s.get_buckets(DR, m)
    for oid, ov, osz, gid, gv, gsz in ((cid, cv, cs, ncid, ncv, ncs), (fid, fv, fs, nfid, nfv, nfs)):
        for col in range(DR):
            a, b = oid[col, :osz[col]], ov[col, :osz[col]]
            keep = ~np.isin(a, dels)
            assert gsz[col] == keep.sum(), col
            assert np.array_equal(gid[col, :gsz[col]], a[keep]), col
            assert np.array_equal(bits(gv[col, :gsz[col]]), bits(b[keep])), col
    cid, cv, fid, fv, cs, fs = ncid, ncv, nfid, nfv, ncs, nfs
    say(f"random delete: all {2 * DR} lists equal the old lists without the deleted ids")

    fifo_update(22)

    # estimate_search against the oracle over the live rows and the device lists
    Xl = np.concatenate(live)[dropped:]
    assert len(Xl) == s.size()
    Q = make_queries(w, 300)
    s.n_probed_vectors, s.n_probed_points, s.n_cand = 40, m, 1000
    gi, gp = s.estimate_search(Q, 10)
    ol = oracle.OracleCEOs(*Xl.shape)
    ol.setIndexParam(w.D, w.R, m, 16, w.seed)
    ol.load_index(Xl, cid, cv, fid, fv)
    ol.set_sizes(cs, fs)
    ol.n_probed_vectors, ol.n_probed_points, ol.n_cand = 40, m, 1000
    oi, op = ol.search_coCEOs_Est(Q, 10)
    verdict = {"same": 0, "tie": 0, "diff": 0}
    for q in range(len(Q)):
        v = topk_verdict(Q[q], lambda r: Xl[r], gi[q], gp[q], oi[q], op[q])
        verdict[v] += 1
        if v != "same":
            say(f"query {q}: {v}")
    say(f"estimate_search vs oracle on {len(Q)} queries: {verdict}")
    assert verdict["diff"] == 0 and verdict["same"] >= 0.999 * len(Q)
    assert not np.isin(gi, dels[dels >= 10000] - 10000).any()  # deleted points never return (ids shifted by the FIFO update)
    del ol, Xl

    # serving across updates: each wave equals estimate_search on the state it saw
    srv = CEOs.Server(s, max_batch=256, k_max=10, s_max=40, c_max=1000, inflight=2)
    n = len(Q)
    k = np.full(n, 10, np.int32)
    sv = np.full(n, 40, np.int32)
    c = np.full(n, 1000, np.int32)
    for wave in range(2):
        t0 = srv.submit(Q, k, sv, c, np.cumsum(rng.exponential(1.0 / 50000.0, n)))
        res = [srv.result(t0 + i) for i in range(n)]
        ei, ep = s.estimate_search(Q, 10)
        for i in range(n):
            assert np.array_equal(res[i][0], ei[i]), (wave, i)
            assert np.array_equal(bits(res[i][1]), bits(ep[i])), (wave, i)
        srv.update(make_points(w, 10000, seed=30 + wave), 10000)
        srv.delete_ids(rng.choice(s.size() - 10000, 1000, replace=False).astype(np.int32))
        srv.drain()
    st = srv.stats()
    srv.close()
    say(f"served = alone over 2 waves with updates between; inserts {st['inserts']} "
        f"(mean {st['insert_mean_ms']:.2f} ms), deletes {st['deletes']} (mean {st['delete_mean_ms']:.2f} ms)")
