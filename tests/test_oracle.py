"""CPU tests that pin the oracle before it is trusted as the reference answer.

Each oracle function is checked against something other than its own formula:
the reference's own FFHT (compiled from the reference sources, or its committed
outputs), the host C++ library's RNG, an FP64 Hadamard matrix, closed forms,
invariants, and brute force on tiny inputs. A planted mistake in any stage (a
dropped sign, a transposed stage order, an off-by-one probe) fails one of them.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def minstd_bits(nbits, seed):
    """std::default_random_engine (minstd_rand0) + libstdc++ uniform_int_distribution(0,1)
    downscale path (/usr/include/c++/11/bits/uniform_int_dist.h:317-326), in Python."""
    x = seed % 2147483647 or 1
    out = []
    while len(out) < nbits:
        x = (16807 * x) % 2147483647
        r = x - 1
        if r >= 2147483644:
            continue
        out.append(r // 1073741822)
    return np.array(out, np.uint8)


def hadamard(L):
    n = 1 << L
    i = np.arange(n)
    pc = np.vectorize(lambda v: bin(v).count("1"))(i[:, None] & i[None, :])
    return np.where(pc % 2 == 0, 1.0, -1.0)


# ------------------------------------------------------------------ sign bits
def test_sign_bits_match_library_and_closed_form(oracle):
    gold = json.load(open(os.path.join(GOLD, "sign_bits.json")))
    for seed, s in gold.items():
        ref = np.frombuffer(s.encode(), np.uint8) - ord("0")
        got = oracle.sign_bits(len(ref), int(seed))
        assert np.array_equal(got, ref), seed
        assert np.array_equal(minstd_bits(len(ref), int(seed)), ref), seed


def test_sign_bits_kat_seed42(oracle):
    # first 64 bits for seed 42 (std::default_random_engine, libstdc++)
    kat = "0110001110011000000111101010000001111110000100100000110101011010"
    assert "".join(map(str, oracle.sign_bits(64, 42))) == kat


# ------------------------------------------------------------------ FWHT
@pytest.mark.parametrize("L", range(1, 13))
def test_fwht_matches_reference_ffht(oracle, L):
    z = np.load(os.path.join(GOLD, "ffht_vectors.npz"))
    x, y = z[f"x{L}"], z[f"y{L}"]
    assert np.array_equal(oracle.fwht_f32(x).view(np.uint32), y.view(np.uint32))
    ff = oracle.ffht_ref()
    if ff is not None:  # the reference's fht_float compiled from its own sources
        rng = np.random.default_rng(L)
        for _ in range(4):
            v = rng.standard_normal(1 << L).astype(np.float32)
            b = v.copy()
            ff.fht_float(b, L)
            assert np.array_equal(oracle.fwht_f32(v).view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("L", [1, 3, 6, 9])
def test_fwht_f64_equals_hadamard_matrix(oracle, L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(1 << L)
    np.testing.assert_allclose(oracle.fwht_f64(x), hadamard(L) @ x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("L", [4, 10, 12])
def test_fwht_invariants(oracle, L):
    n = 1 << L
    rng = np.random.default_rng(L)
    x = rng.standard_normal(n)
    hx = oracle.fwht_f64(x)
    np.testing.assert_allclose(oracle.fwht_f64(hx), n * x, rtol=1e-10, atol=1e-9)      # H H = n I
    np.testing.assert_allclose((hx ** 2).sum(), n * (x ** 2).sum(), rtol=1e-10)         # Parseval
    e0 = np.zeros(n); e0[0] = 1
    np.testing.assert_array_equal(oracle.fwht_f64(e0), np.ones(n))                     # H e0 = 1
    y = rng.standard_normal(n)
    np.testing.assert_allclose(oracle.fwht_f64(2 * x - 3 * y), 2 * hx - 3 * oracle.fwht_f64(y), rtol=1e-9,
                               atol=1e-9)  # linearity
    # FP32 ascending-stage transform stays within the FP64 error bound L * u * ||x||_1-ish
    x32 = x.astype(np.float32)
    err = np.abs(oracle.fwht_f32(x32).astype(np.float64) - oracle.fwht_f64(x32.astype(np.float64)))
    assert err.max() <= L * 2.0 ** -24 * np.abs(x32).sum() * 2


# ------------------------------------------------------------------ projection
def test_projection_is_sign_fwht_composition(oracle):
    """P = first D of (H S3 H S2 H S1 [x;0]) per repeat, checked in FP64 with the fixture bits."""
    n, d, D, R = 50, 37, 64, 3
    rng = np.random.default_rng(3)
    X = rng.standard_normal((n, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, 4, 1, 42)
    F = o.F
    assert F == 64
    bits = oracle.sign_bits(F * 3 * R, 42).astype(np.float64)
    H = hadamard(6)
    P = o.project(X).astype(np.float64)
    for r in range(R):
        M = np.eye(F)
        for i in range(3):
            s = 2 * bits[r * 3 * F + i * F: r * 3 * F + (i + 1) * F] - 1
            M = H @ (s[:, None] * M)
        xp = np.zeros((n, F)); xp[:, :d] = X
        exp = (xp @ M.T)[:, :D]
        err = np.abs(P[:, r * D:(r + 1) * D] - exp)
        bound = 3 * 6 * 2.0 ** -24 * np.abs(xp).sum(1, keepdims=True) * F * 2
        assert (err <= bound).all()
    # a planted mistake (the second sign-flip-and-FWHT step loses its signs) is caught
    s_wrong = bits.copy(); s_wrong[F:2 * F] = 1
    M = np.eye(F)
    for i in range(3):
        s = 2 * s_wrong[i * F:(i + 1) * F] - 1
        M = H @ (s[:, None] * M)
    xp = np.zeros((n, F)); xp[:, :d] = X
    assert np.abs(P[:, :D] - (xp @ M.T)[:, :D]).max() > 1e-2


def test_fht_dim_rule(oracle):
    # src/CEOs.h:76-79
    assert oracle.fht_dim(1024, 300) == 1024
    assert oracle.fht_dim(1024, 960) == 1024
    assert oracle.fht_dim(512, 384) == 512
    assert oracle.fht_dim(100, 129) == 256
    assert oracle.fht_dim(256, 256) == 256
    assert oracle.fht_dim(257, 10) == 512


# ------------------------------------------------------------------ buckets
def test_buckets_equal_bruteforce_sort(oracle):
    n, d, D, R, m = 400, 16, 32, 2, 15
    rng = np.random.default_rng(4)
    X = rng.standard_normal((n, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 2, 42)
    o.build_coCEOs_Est(X)
    P = o.project(X)
    cid, cv, fid, fv = o.buckets()
    for c in range(D * R):
        col = P[:, c]
        order = np.lexsort((np.arange(n), -col))[:m]       # value desc, id asc
        assert np.array_equal(cv[c], col[order])
        assert np.array_equal(cid[c], order)
        order = np.lexsort((np.arange(n), col))[:m]        # far: -value desc
        assert np.array_equal(fv[c], -col[order])
        assert np.array_equal(fid[c], order)


def test_tiny_index_fixture(oracle):
    z = np.load(os.path.join(GOLD, "tiny_index.npz"))
    D, R, m, seed, s, mp, c, k = z["params"].tolist()
    X, Q = z["X"], z["Q"]
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(D, R, m, 1, seed)
    o.build_coCEOs_Est(X)
    cid, cv, fid, fv = o.buckets()
    assert np.array_equal(cid, z["cid"]) and np.array_equal(fid, z["fid"])
    assert np.array_equal(cv.view(np.uint32), z["cval"].view(np.uint32))
    o.n_probed_vectors, o.n_cand = s, c
    ids, ips = o.search_coCEOs_Est(Q, k)
    assert np.array_equal(ids, z["ids"])
    assert np.array_equal(ips.view(np.uint32), z["ips"].view(np.uint32))


# ------------------------------------------------------------------ search
def test_full_budget_search_is_exact_mips(oracle):
    """With every point in every bucket and every point reranked, coCEOs returns exact MIPS."""
    n, d = 300, 24
    rng = np.random.default_rng(6)
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((20, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(32, 1, n, 1, 42)
    o.build_coCEOs_Est(X)
    o.n_probed_vectors, o.n_cand = 5, n
    ids, ips = o.search_coCEOs_Est(Q, 10)
    ex, exip = oracle.exact_mips(X, Q, 10)
    assert np.array_equal(ids, ex)
    np.testing.assert_allclose(ips, (Q.astype(np.float64) @ X.T.astype(np.float64))[np.arange(20)[:, None], ex],
                               rtol=1e-6)


def test_coceos_with_full_buckets_equals_ceos_est(oracle):
    """m = n: every bucket holds every point, so coCEOs' estimates are CEOs-Est's (same probes)."""
    n, d = 500, 30
    rng = np.random.default_rng(7)
    X = rng.standard_normal((n, d)).astype(np.float32) * rng.lognormal(0, .5, (n, 1)).astype(np.float32)
    Q = rng.standard_normal((30, d)).astype(np.float32)
    a = oracle.OracleCEOs(n, d); a.setIndexParam(64, 2, n, 1, 42); a.build_coCEOs_Est(X)
    b = oracle.OracleCEOs(n, d); b.setIndexParam(64, 2, n, 1, 42); b.build(X)
    for o in (a, b):
        o.n_probed_vectors, o.n_cand = 7, 40
    ia, pa = a.search_coCEOs_Est(Q, 10)
    ib, pb = b.search(Q, 10)
    assert np.array_equal(ia, ib)
    assert np.array_equal(pa, pb)


def test_hash_search_full_buckets_is_exact_mips(oracle):
    """CEOs-Hash with every point in every bucket reranks every point: exact MIPS."""
    n, d = 400, 20
    rng = np.random.default_rng(61)
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((15, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(32, 2, n, 1, 42)
    o.build_CEOs_Hash(X)
    o.n_probed_vectors, o.n_probed_points = 3, n
    ids, ips, cnt = o.search_CEOs_Hash(Q, 8, stats=True)
    assert np.all(cnt == n)
    ex, _ = oracle.exact_mips(X, Q, 8)
    assert np.array_equal(ids, ex)


def test_hash_candidates_are_union_of_probed_lists(oracle):
    """The candidate set is the union of the first m_p ids of the s probed lists (probes
    recomputed from the oracle's projection), and the result is the exact top-k of it."""
    n, d, D, R, m = 2000, 32, 64, 2, 60
    rng = np.random.default_rng(62)
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((12, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 1, 42)
    o.build_CEOs_Hash(X)
    s, mp, k = 9, 40, 7
    o.n_probed_vectors, o.n_probed_points = s, mp
    ids, ips, cnt = o.search_CEOs_Hash(Q, k, stats=True)
    cid, _, fid, _ = o.buckets()
    P = o.project(Q)
    for q in range(len(Q)):
        top = np.argsort(-np.abs(P[q]), kind="stable")[:s]
        union = set()
        for col in top:
            lst = fid[col] if P[q, col] < 0 else cid[col]
            union.update(lst[:mp].tolist())
        assert cnt[q] == len(union)
        cand = np.array(sorted(union))
        ip = (X[cand].astype(np.float64) @ Q[q].astype(np.float64))
        order = cand[np.lexsort((cand, -ip))][:k]
        assert np.array_equal(ids[q], order)


def test_probe_order_and_scores_bruteforce(oracle):
    """Estimates recomputed in numpy from projections, probes and buckets pick the same candidates."""
    n, d, D, R, m, s, c, k = 800, 20, 64, 2, 40, 9, 60, 10
    rng = np.random.default_rng(8)
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((25, d)).astype(np.float32)
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 1, 42)
    o.build_coCEOs_Est(X)
    o.n_probed_vectors, o.n_cand = s, c
    ids, _ = o.search_coCEOs_Est(Q, k)
    cid, cv, fid, fv = o.buckets()
    pq = o.project(Q)
    for q in range(len(Q)):
        p = pq[q]
        cols = np.lexsort((np.arange(D * R), -np.abs(p)))[:s]
        est = {}
        for col in cols[::-1]:  # ascending |p|: the reference's pop order
            lids, lv = (cid[col], cv[col]) if p[col] >= 0 else (fid[col], fv[col])
            for i, v in zip(lids, lv):
                est[i] = np.float32(est.get(i, np.float32(0)) + v)
        cand = sorted(est, key=lambda i: (-est[i], i))[:c]
        ip = X[cand].astype(np.float64) @ Q[q].astype(np.float64)
        top = [cand[j] for j in np.lexsort((np.array(cand), -ip))[:k]]
        assert list(ids[q]) == top


# ------------------------------------------------------------------ streamCEOs
def test_stream_update_closed_form(oracle):
    """bucket' = top_m({old entries with id >= n_del} U {new points}), ids shifted by -n_del."""
    d, D, R, m = 12, 32, 2, 20
    rng = np.random.default_rng(9)
    X0 = rng.standard_normal((300, d)).astype(np.float32)
    X1 = rng.standard_normal((120, d)).astype(np.float32)
    n_del = 90
    st = oracle.OracleStream(d)
    st.setIndexParam(D, R, m, 1, 42)
    st.build(X0)
    cid0, cv0, fid0, fv0, _, _ = st.buckets()
    st.update(X1, n_del)
    assert st.size() == 300 + 120 - 90
    cid, cv, fid, fv, cs, fs = st.buckets()
    o = oracle.OracleCEOs(120, d)
    o.setIndexParam(D, R, m, 1, 42)
    Pn = o.project(X1)
    for c in range(D * R):
        for (ids0, v0, ids, v, sgn) in ((cid0, cv0, cid, cv, 1), (fid0, fv0, fid, fv, -1)):
            keep = ids0[c] >= n_del
            cand_id = np.concatenate([ids0[c][keep], 300 + np.arange(120)]) - n_del
            cand_v = np.concatenate([v0[c][keep], sgn * Pn[:, c]])
            order = np.lexsort((cand_id, -cand_v))[:m]
            assert np.array_equal(v[c], cand_v[order])
            assert np.array_equal(ids[c], cand_id[order])
    assert (cs == m).all() and (fs == m).all()


def test_stream_reference_errors(oracle):
    st = oracle.OracleStream(8)
    st.setIndexParam(16, 1, 10, 1, 42)
    st.build(np.ones((20, 8), np.float32))
    with pytest.raises(ValueError):
        st.update(np.ones((1, 8), np.float32), 25)      # n_del > size
    with pytest.raises(ValueError):
        st.update(np.ones((1, 8), np.float32), 15)      # fewer than top_m left


def test_bucket_columns_equal_full_build(oracle):
    """The chunked per-column build (rows projected 32,768 at a time, only the repeats
    the columns fall in) gives the full build's lists."""
    rng = np.random.default_rng(41)
    X = (rng.standard_normal((70000, 40)) * rng.lognormal(0, 0.5, (70000, 1))).astype(np.float32)
    o = oracle.OracleCEOs(*X.shape)
    o.setIndexParam(128, 3, 50, 4, 42)
    o.build_coCEOs_Est(X)
    full = o.buckets()
    cols = np.array([0, 5, 127, 128, 200, 383])
    part = o.bucket_columns(X, cols)
    for a, b in zip(part, full):
        assert np.array_equal(a, b[cols])


# ------------------------------------------------------------------ non-finite rows
def test_nan_never_enters_a_list(oracle):
    """Rows with a NaN project to NaN everywhere; they never enter a list (the strict '>'
    of src/CEOs.cpp:395,406 never admits one; the heap-filling case is defined the same
    way). A row with an infinite entry projects to +-Inf after the first transform and to
    NaN after the second (Inf - Inf), so it is left out as well; infinite projections of
    finite rows (overflow) would be ordinary values. Checked against a brute-force sort of
    the non-NaN projections."""
    n, d, D, R, m = 300, 16, 32, 2, 12
    rng = np.random.default_rng(14)
    X = rng.standard_normal((n, d)).astype(np.float32)
    X[[0, 3, 150, 299], 5] = np.nan     # NaN rows, two of them while the heaps are filling
    X[[7, 200], 2] = np.inf
    X[[8], 9] = -np.inf
    o = oracle.OracleCEOs(n, d)
    o.setIndexParam(D, R, m, 2, 42)
    o.build_coCEOs_Est(X)
    P = o.project(X)
    cid, cv, fid, fv = o.buckets()
    ids = np.arange(n)
    for c in range(D * R):
        col = P[:, c]
        ok = ~np.isnan(col)
        order = ids[ok][np.lexsort((ids[ok], -col[ok].astype(np.float64)))][:m]
        assert np.array_equal(cv[c], col[order]) and np.array_equal(cid[c], order), c
        order = ids[ok][np.lexsort((ids[ok], col[ok].astype(np.float64)))][:m]
        assert np.array_equal(fv[c], -col[order]) and np.array_equal(fid[c], order), c
        assert not np.isin(cid[c], [0, 3, 7, 8, 150, 200, 299]).any()
        assert not np.isin(fid[c], [0, 3, 7, 8, 150, 200, 299]).any()


# ------------------------------------------------------------------ centering
def test_column_mean_within_fp32_bound(oracle):
    """The Eigen-order FP32 column mean is within the rounding bound of its summation
    ((n/4 + 4) u sum|x| / n, u = 2^-24) of the FP64 mean, and exact on data whose sums
    are exact in FP32."""
    rng = np.random.default_rng(15)
    for n, d in ((1, 3), (2, 5), (9, 7), (10, 7), (4097, 33), (200001, 16)):
        X = (rng.standard_normal((n, d)) * 3 + 5).astype(np.float32)
        c = oracle.column_mean(X)
        m64 = X.astype(np.float64).mean(0)
        bound = (n / 4 + 4) * 2.0 ** -24 * np.abs(X).astype(np.float64).mean(0) + 2.0 ** -24 * np.abs(m64)
        assert np.all(np.abs(c - m64) <= bound), (n, d)
    Xi = rng.integers(-1000, 1000, (1001, 6)).astype(np.float32)   # integer sums: exact
    np.testing.assert_array_equal(oracle.column_mean(Xi), (Xi.astype(np.float64).sum(0) / 1001).astype(np.float32))


def test_stream_centering_subtracts_the_build_mean(oracle):
    """streamCEOs with centering stores rows minus the build rows' mean and subtracts the
    same centre from later inserts (src/streamCEOs.cpp:22-34, 198-203): its lists equal
    those of an uncentred index over the pre-centred rows."""
    rng = np.random.default_rng(16)
    X = (rng.standard_normal((3000, 24)) + 2.0).astype(np.float32)
    Xn = (rng.standard_normal((500, 24)) + 2.0).astype(np.float32)
    a = oracle.OracleStream(24)
    a.setIndexParam(64, 2, 30, 2, 42)
    a.set_centering(True)
    a.build(X)
    a.update(Xn, 200)
    c = a.center()
    np.testing.assert_array_equal(c, oracle.column_mean(X))
    b = oracle.OracleStream(24)
    b.setIndexParam(64, 2, 30, 2, 42)
    b.build(X - c)
    b.update(Xn - c, 200)
    for x, y in zip(a.buckets(), b.buckets()):
        assert np.array_equal(x, y)


# ------------------------------------------------------------------ projection error
def _proj64(X, bits, F, D, R):
    """Sign + FWHT projections in FP64 (numpy, ascending stages), first D per repeat."""
    n, d = X.shape
    out = np.zeros((n, D * R))
    for r in range(R):
        y = np.zeros((n, F))
        y[:, :d] = X
        for i in range(3):
            y *= 2 * bits[r * 3 * F + i * F: r * 3 * F + (i + 1) * F].astype(np.float64) - 1
            h = 1
            while h < F:
                y = y.reshape(n, F // (2 * h), 2, h)
                y = np.concatenate([y[:, :, :1] + y[:, :, 1:], y[:, :, :1] - y[:, :, 1:]], axis=2).reshape(n, F)
                h *= 2
        out[:, r * D:(r + 1) * D] = y[:, :D]
    return out


@pytest.mark.parametrize("F", [512, 1024, 4096])
def test_projection_error_against_fp64(oracle, F):
    """States the FP32 projection's error against FP64 (the device projection is bit-equal
    to the oracle's, checked on the GPU): normwise relative error and the largest
    coordinate error over the largest |coordinate|, on N(0,I) rows with d = F."""
    rng = np.random.default_rng(17)
    n, R = 64, 2
    X = rng.standard_normal((n, F)).astype(np.float32)
    o = oracle.OracleCEOs(n, F)
    o.setIndexParam(F, R, 4, 2, 42)
    assert o.F == F
    P32 = o.project(X).astype(np.float64)
    P64 = _proj64(X.astype(np.float64), oracle.sign_bits(F * 3 * R, 42), F, F, R)
    normwise = np.linalg.norm(P32 - P64, axis=1) / np.linalg.norm(P64, axis=1)
    coord = np.abs(P32 - P64).max(1) / np.abs(P64).max(1)
    print(f"F={F}: normwise max {normwise.max():.3e} mean {normwise.mean():.3e}; "
          f"max coordinate error / max |p| {coord.max():.3e}")
    L = F.bit_length() - 1
    assert normwise.max() <= 3 * L * 2.0 ** -24 * 4  # 3 rounds of L stages
    assert coord.max() < 1e-5


def test_timing_build_computes_what_the_oracle_computes(oracle):
    """The timed CPU baseline (reference FFHT, tsl::robin_map, FP32 SIMD dot) builds the
    same lists and returns the same top-k as the parity oracle."""
    if oracle.lib(timing=True) is None:
        pytest.skip("timing build absent (the reference tree was not present at build time)")
    rng = np.random.default_rng(18)
    X = (rng.standard_normal((5000, 64)) * rng.lognormal(0, 0.5, (5000, 1))).astype(np.float32)
    Q = rng.standard_normal((100, 64)).astype(np.float32)
    out = []
    for timing in (False, True):
        o = oracle.OracleCEOs(*X.shape, timing=timing)
        o.setIndexParam(256, 2, 60, 2, 42)
        o.build_coCEOs_Est(X)
        o.n_probed_vectors, o.n_cand = 20, 300
        out.append((o.buckets(), o.search_coCEOs_Est(Q, 10)))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1][0], out[1][1][0])
    np.testing.assert_allclose(out[0][1][1], out[1][1][1], rtol=1e-5, atol=1e-6)
