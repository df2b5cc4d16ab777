// C-ABI entry points of libceos_amd (include/ceos_amd.h): device-resident coCEOs,
// CEOs-Est and streamCEOs indexes, captured search plans.
#include "ceos_core.h"

extern "C" {

const char* ceos_last_error(void) { return g_err.c_str(); }

int ceos_create(int n_points, int n_features, ceos_index** out) {
    return guarded([&] {
        if (!out) throw ArgError("null output handle");
        if (n_points < 0 || n_features <= 0) throw ArgError("n_points must be >= 0 and n_features > 0");
        if (n_features > 4096) throw ArgError("n_features above 4096 is not supported");
        auto* h = new ceos_index();
        h->n_points = n_points;
        h->core.d = n_features;
        *out = h;
    });
}

void ceos_destroy(ceos_index* h) { delete h; }

int ceos_set_index_param(ceos_index* h, int n_proj, int n_repeats, int top_m, int n_threads, int seed) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.set_index_param(n_proj, n_repeats, top_m, n_threads, seed);
        h->core.mp = top_m;  // src/CEOs.h:69
    });
}

int ceos_set_query_param(ceos_index* h, int s, int mp, int c) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.s = s; h->core.mp = mp; h->core.c = c;
    });
}

int ceos_get_query_param(const ceos_index* h, int* s, int* mp, int* c) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (s) *s = h->core.s;
        if (mp) *mp = h->core.mp;
        if (c) *c = h->core.c;
    });
}

int ceos_set_centering(ceos_index* h, int centering) {
    return guarded([&] { if (!h) throw ArgError("null handle"); h->core.centering = centering != 0; });
}
int ceos_get_centering(const ceos_index* h, int* centering) {
    return guarded([&] { if (!h || !centering) throw ArgError("null handle"); *centering = h->core.centering; });
}
int ceos_set_threads(ceos_index* h, int n_threads) {
    return guarded([&] { if (!h) throw ArgError("null handle"); h->core.threads = n_threads; });
}
int ceos_get_threads(const ceos_index* h, int* n_threads) {
    return guarded([&] { if (!h || !n_threads) throw ArgError("null handle"); *n_threads = h->core.threads; });
}
int ceos_clear(ceos_index* h) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.clear();
    });
}

int ceos_get_dims(const ceos_index* h, int* fht, int* n_proj, int* n_repeats, int* top_m, int* seed) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        const Core& c = h->core;
        if (fht) *fht = c.F;
        if (n_proj) *n_proj = c.D;
        if (n_repeats) *n_repeats = c.R;
        if (top_m) *top_m = c.m;
        if (seed) *seed = (int)c.seed_used;
    });
}

// center: subtract the points' column mean on the device after the copy (CEOs-Hash).
static void build_coceos_impl(ceos_index* h, const float* X, int64_t n, int d, bool dev, hipStream_t st,
                              bool center = false) {
    Core& c = h->core;
    if (!c.params_set) throw StateError("setIndexParam must be called before build");
    check_matrix(X, n, d, c.d);
    if (n <= 0) throw ArgError("empty dataset");
    if (n >= (int64_t(1) << 31)) throw ArgError("more than 2^31-1 points");
    c.wait_reads();
    const auto t0 = std::chrono::steady_clock::now();
    c.built = false;
    h->n_points = (int)n;
    c.estP.release();
    c.dcenter.release();
    c.center.clear();
    c.phase_mark(0, st);
    // a device source without centering: the bucket scan reads it and stores the rows,
    // one pass over the points instead of a copy and then the scan
    const bool direct = dev && !center && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
    if (direct)
        c.reserve_points(n, n);
    else
        c.load_points(X, n, n, dev, st);
    if (center) c.center_rows(n, st);
    c.phase_mark(1, st);
    c.make_signs();
    if (direct) {
        // the rerank bounds from the caller's rows on the second stream, beside the scan
        c.quantize_rows_side(X, n, st);
        try {
            c.build_buckets(n, st, X);
        } catch (...) {
            c.drain_aux();
            throw;
        }
        c.join_aux(st);
    } else {
        c.build_buckets(n, st);
        c.quantize_rows(0, n, n, st);
    }
    HIP_CHECK(hipStreamSynchronize(st));
    c.finish_build_times(t0);
    c.built = true;
}

int ceos_build_coceos(ceos_index* h, const float* X, int64_t n, int d) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        build_coceos_impl(h, X, n, d, false, h->core.st());
    });
}

int ceos_build_coceos_device(ceos_index* h, const float* dX, int64_t n, int d, void* stream) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        build_coceos_impl(h, dX, n, d, true, h->core.stream(stream));
    });
}

// build_CEOs_Hash: the same projections and per-column top-m lists as coCEOs (the
// reference keeps only their ids, src/CEOs.cpp:709-902). With centering the stored
// points are X minus its column mean; the reference's expression (CEOs.cpp:721-728,
// a row-wise mean subtracted from every row) only type-checks when n equals d.
int ceos_build_hash(ceos_index* h, const float* X, int64_t n, int d) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(X, n, d, c.d);
        build_coceos_impl(h, X, n, d, false, c.st(), c.centering);
    });
}

// search_CEOs_Hash(queries, n_neighbors): the union of the first n_probed_points ids of
// the n_probed_vectors lists, every candidate reranked (src/CEOs.cpp:904-1066).
int ceos_search_hash(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        c.validate_query(c.s, c.mp, 1, k, c.DR(), "n_proj * n_repeats");  // src/CEOs.cpp:906-915
        c.search_host(Q, nq, k, c.s, 1, ids, ips, true);
    });
}

int ceos_save(const ceos_index* h, const char* path) {
    return guarded([&] {
        if (!h || !path) throw ArgError("null argument");
        auto& c = const_cast<Core&>(h->core);
        std::lock_guard<std::mutex> g(c.mu);
        if (c.estP.p && !c.bkt.p) throw StateError("CEOs-Est indexes are not saved; build coCEOs or CEOs-Hash");
        c.save(path, 0);
    });
}

int ceos_load(ceos_index* h, const char* path) {
    return guarded([&] {
        if (!h || !path) throw ArgError("null argument");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->n_points = (int)h->core.load(path, 0);
    });
}

int ceos_search_coceos(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        c.validate_query(c.s, c.mp, c.c, k, c.DR(), "n_proj * n_repeats");
        c.search_host(Q, nq, k, c.s, c.c, ids, ips);
    });
}

int ceos_search_coceos_device(ceos_index* h, const float* dQ, int nq, int d, int k, int32_t* d_ids, float* d_ips,
                              void* stream) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(dQ, nq, d, c.d);
        c.validate_query(c.s, c.mp, c.c, k, c.DR(), "n_proj * n_repeats");
        c.search_device(dQ, nq, k, c.s, c.c, d_ids, d_ips, c.stream(stream));
    });
}

int ceos_search_batched_device(ceos_index* h, const float* dQ, int nq, int d, const int32_t* d_k,
                               const int32_t* d_s, const int32_t* d_c, const int64_t* d_off, int k_max,
                               int s_max, int c_max, int32_t* d_ids, float* d_ips, void* stream) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(dQ, nq, d, c.d);
        c.validate_query(s_max, c.mp, c_max, k_max, c.DR(), "n_proj * n_repeats");
        if (nq <= 0) return;
        if (!d_k || !d_s || !d_c || !d_off) throw ArgError("per-query arrays must not be null");
        const ceos::QueryConfig cfg = c.search_config(s_max, c_max, std::max(1, k_max));
        hipStream_t st = c.stream(stream);
        c.before_scratch(st);
        const ceos::QueryScratch scr = c.scratch(cfg, nq);
        ceos::QueryBatch qb{dQ, nq, d_k, d_s, d_c, 0, 0, 0, c.mp, d_off, d_ids, d_ips};
        HIP_CHECK(ceos::launch_coceos_search(c.dims(), c.signs.p, c.bkt.p, c.bsize.p, c.rows(), qb, cfg, scr, st));
        c.after_scratch(st);
    });
}

int ceos_search_batched(ceos_index* h, const float* Q, int nq, int d, const int32_t* k, const int32_t* s,
                        const int32_t* c_, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        if (nq <= 0) return;
        if (!k || !s || !c_) throw ArgError("per-query arrays must not be null");
        int k_max = 0, s_max = 1, c_max = 1;
        std::vector<int64_t> off(nq);
        int64_t tot = 0;
        for (int q = 0; q < nq; ++q) {
            if (k[q] < 0 || s[q] <= 0 || c_[q] <= 0) throw ArgError("per-query k must be >= 0 and s, c > 0");
            off[q] = tot;
            tot += k[q];
            k_max = std::max(k_max, k[q]);
            s_max = std::max(s_max, s[q]);
            c_max = std::max(c_max, c_[q]);
        }
        c.validate_query(s_max, c.mp, c_max, k_max, c.DR(), "n_proj * n_repeats");
        if (tot == 0) return;
        DBuf<float> dQ((size_t)nq * d);
        DBuf<int32_t> dk(nq), ds(nq), dc(nq), dI(tot);
        DBuf<int64_t> doff(nq);
        DBuf<float> dP(tot);
        HIP_CHECK(hipMemcpyAsync(dQ.p, Q, (size_t)nq * d * 4, hipMemcpyHostToDevice, c.st()));
        HIP_CHECK(hipMemcpyAsync(dk.p, k, nq * 4, hipMemcpyHostToDevice, c.st()));
        HIP_CHECK(hipMemcpyAsync(ds.p, s, nq * 4, hipMemcpyHostToDevice, c.st()));
        HIP_CHECK(hipMemcpyAsync(dc.p, c_, nq * 4, hipMemcpyHostToDevice, c.st()));
        HIP_CHECK(hipMemcpyAsync(doff.p, off.data(), nq * 8, hipMemcpyHostToDevice, c.st()));
        const ceos::QueryConfig cfg = c.search_config(s_max, c_max, std::max(1, k_max));
        c.before_scratch(c.st());
        const ceos::QueryScratch scr = c.scratch(cfg, nq);
        ceos::QueryBatch qb{dQ.p, nq, dk.p, ds.p, dc.p, 0, 0, 0, c.mp, doff.p, dI.p, dP.p};
        HIP_CHECK(ceos::launch_coceos_search(c.dims(), c.signs.p, c.bkt.p, c.bsize.p, c.rows(), qb, cfg, scr, c.st()));
        c.after_scratch(c.st());
        HIP_CHECK(hipMemcpyAsync(ids, dI.p, tot * 4, hipMemcpyDeviceToHost, c.st()));
        HIP_CHECK(hipMemcpyAsync(ips, dP.p, tot * 4, hipMemcpyDeviceToHost, c.st()));
        HIP_CHECK(hipStreamSynchronize(c.st()));
        c.check_status();
    });
}

int ceos_build_est(ceos_index* h, const float* X, int64_t n, int d) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        if (!c.params_set) throw StateError("setIndexParam must be called before build");
        check_matrix(X, n, d, c.d);
        if (n <= 0) throw ArgError("empty dataset");
        c.wait_reads();
        c.built = false;
        h->n_points = (int)n;
        c.bkt.release();
        c.bsize.release();
        c.dcenter.release();
        c.center.clear();
        c.phase_mark(0, c.st());
        c.load_points(X, n, n, false, c.st());
        c.phase_mark(1, c.st());
        c.make_signs();
        c.estP.alloc((size_t)c.DR() * n);
        c.est_n = n;
        // column-major [DR][n]: chunks of rows projected row-major (coalesced stores) into a
        // workspace of at most 1 GiB, then transposed into place through LDS tiles
        const int64_t DR_ = c.DR();
        const int64_t chunk = std::min<int64_t>(
            {n, int64_t(65535) * 64, std::max<int64_t>(64, ((int64_t(1) << 28) / DR_) & ~int64_t(63))});
        DBuf<float> tmp((size_t)chunk * DR_);
        for (int64_t r0 = 0; r0 < n; r0 += chunk) {
            const int64_t rows = std::min(chunk, n - r0);
            HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p + r0 * c.d, 0, 1, rows, c.signs.p, tmp.p, DR_, false,
                                                c.st()));
            HIP_CHECK(ceos::launch_transpose(tmp.p, rows, DR_, c.estP.p + r0, c.st(), n));
        }
        c.phase_mark(2, c.st());
        HIP_CHECK(hipStreamSynchronize(c.st()));
        c.last_build = ceos_phase_times{};
        c.last_build.copy_ms = c.phase_ms(0, 1);
        c.last_build.project_ms = c.phase_ms(1, 2);
        c.built = true;
    });
}

int ceos_search_est(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        if (!c.built || !c.estP.p) throw StateError("CEOs-Est index is not built (call build)");
        if (c.s > c.DR()) throw ArgError("Number of probed vectors must be smaller than n_proj * n_repeats");
        if (c.s <= 0 || c.c <= 0) throw ArgError("n_probed_vectors and n_cand must be positive");
        if (k < 0 || k > kMaxK) throw ArgError("n_neighbors out of range");
        if (nq <= 0 || k == 0) return;
        c.wait_reads();  // the scratch below is shared with device searches
        const int cc = (int)std::min<int64_t>(c.c, c.est_n);
        const int lds = ceos::est_search_lds_bytes(c.dims(), c.s, cc, k);
        if (lds > kMaxLdsBytes) throw ArgError("n_cand too large for the device path");
        const int64_t stride = c.est_n + cc;
        int grid = device_cus() * 2;
        // per-workgroup scratch holds n keys (the unbounded case); at most 4 GiB in all
        grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t(4) << 30) / (stride * 8)));
        c.scr.cand.ensure((size_t)grid * stride);
        if (!c.scr.work.p) c.scr.work.alloc(1);
        DBuf<float> dQ((size_t)nq * d);
        DBuf<int32_t> dI((size_t)nq * k);
        DBuf<float> dP((size_t)nq * k);
        HIP_CHECK(hipMemcpyAsync(dQ.p, Q, (size_t)nq * d * 4, hipMemcpyHostToDevice, c.st()));
        ceos::QueryBatch qb{dQ.p, nq, nullptr, nullptr, nullptr, k, c.s, cc, 0, nullptr, dI.p, dP.p};
        HIP_CHECK(ceos::launch_est_search(c.dims(), c.signs.p, c.estP.p, c.est_n, c.X.p, qb, grid, c.scr.cand.p, stride,
                                          c.scr.work.p, c.st()));
        HIP_CHECK(hipMemcpyAsync(ids, dI.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, c.st()));
        HIP_CHECK(hipMemcpyAsync(ips, dP.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, c.st()));
        HIP_CHECK(hipStreamSynchronize(c.st()));
    });
}

int ceos_get_buckets(const ceos_index* h, int32_t* cid, float* cval, int32_t* fid, float* fval) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        auto& c = const_cast<Core&>(h->core);
        std::lock_guard<std::mutex> g(c.mu);
        c.get_buckets(cid, cval, fid, fval, nullptr, nullptr);
    });
}

int ceos_get_sign_bits(const ceos_index* h, uint8_t* bits) {
    return guarded([&] {
        if (!h || !bits) throw ArgError("null handle");
        if (h->core.bits.empty()) throw StateError("index is not built");
        std::memcpy(bits, h->core.bits.data(), h->core.bits.size());
    });
}

int ceos_project(ceos_index* h, const float* rows, int64_t n, int d, float* out) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(rows, n, d, c.d);
        if (!c.signs.p) throw StateError("index is not built");
        if (n == 0) return;
        DBuf<float> dr((size_t)n * d), dout((size_t)n * c.DR());
        HIP_CHECK(hipMemcpyAsync(dr.p, rows, (size_t)n * d * 4, hipMemcpyHostToDevice, c.st()));
        HIP_CHECK(ceos::launch_project_rows(c.dims(), dr.p, 0, 1, n, c.signs.p, dout.p, c.DR(), false, c.st()));
        HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)n * c.DR() * 4, hipMemcpyDeviceToHost, c.st()));
        HIP_CHECK(hipStreamSynchronize(c.st()));
    });
}

int ceos_set_search_config(ceos_index* h, int wg_per_cu, int table_log2) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (wg_per_cu < 0 || wg_per_cu > 8) throw ArgError("wg_per_cu must be in [0, 8]");
        if (table_log2 != 0 && (table_log2 < 8 || table_log2 > 15)) throw ArgError("table_log2 must be 0 or in [8, 15]");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.wg_per_cu = wg_per_cu;
        h->core.table_log2 = table_log2;
    });
}

int ceos_set_build_budget(ceos_index* h, int64_t entries) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (entries <= 0) throw ArgError("the build budget must be positive");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.rescan_budget_entries = entries;
    });
}

int ceos_get_build_rescans(const ceos_index* h, int* under, int* over, int* passes) {
    return guarded([&] {
        if (!h || !under || !over || !passes) throw ArgError("null argument");
        const Core& c = h->core;
        *under = c.rescan.under;
        *over = c.rescan.over;
        *passes = c.rescan.ranges;
    });
}

int ceos_stream_set_build_budget(ceos_stream* h, int64_t entries) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (entries <= 0) throw ArgError("the build budget must be positive");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.rescan_budget_entries = entries;
    });
}

int ceos_set_rerank_bounds(ceos_index* h, int on) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.set_rerank_bounds(on != 0);
    });
}
int ceos_get_rerank_bounds(const ceos_index* h, int* on) {
    return guarded([&] {
        if (!h || !on) throw ArgError("null argument");
        *on = h->core.rerank_bounds && h->core.d % 4 == 0 ? 1 : 0;
    });
}
int ceos_stream_set_rerank_bounds(ceos_stream* h, int on) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.set_rerank_bounds(on != 0);
    });
}
int ceos_stream_get_rerank_bounds(const ceos_stream* h, int* on) {
    return guarded([&] {
        if (!h || !on) throw ArgError("null argument");
        *on = h->core.rerank_bounds && h->core.d % 4 == 0 ? 1 : 0;
    });
}

int ceos_get_search_config(const ceos_index* h, int* wg_per_cu, int* table_log2, int* lds_bytes) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (wg_per_cu) *wg_per_cu = h->core.last_wg;
        if (table_log2) *table_log2 = h->core.last_tlog2;
        if (lds_bytes) *lds_bytes = h->core.last_lds;
    });
}

int ceos_read_bandwidth(const float* x, int64_t n, float* out, int64_t n_out, int iters, void* stream,
                        double* gbs) {
    return guarded([&] {
        if (!x || !out || !gbs) throw ArgError("null argument");
        if (n <= 0 || n % 4 != 0 || (reinterpret_cast<uintptr_t>(x) & 15) != 0)
            throw ArgError("n must be a positive multiple of 4 and x 16-byte aligned");
        if (n_out < 256 || n_out % 256 != 0 || n_out / 256 > (int64_t)1 << 20)
            throw ArgError("n_out must be a multiple of 256 with at most 2^20 workgroups");
        if (iters <= 0) throw ArgError("iters must be positive");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        const int grid = (int)(n_out / 256);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_CHECK(hipEventCreate(&e0));
        hipError_t err = hipEventCreate(&e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = ceos::launch_stream_read(x, n, out, grid, st);  // warm
        if (err == hipSuccess) err = hipEventRecord(e0, st);
        for (int t = 0; t < iters && err == hipSuccess; ++t) err = ceos::launch_stream_read(x, n, out, grid, st);
        if (err == hipSuccess) err = hipEventRecord(e1, st);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        HIP_CHECK(err);
        *gbs = (double)n * 4.0 * iters / ((double)ms * 1e6);
    });
}

int ceos_set_kernel_timing(ceos_index* h, int on) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.ktiming = on != 0;
        h->core.kev_used = 0;
    });
}

int ceos_get_kernel_timing(ceos_index* h, double* front_ms, double* fused_ms, int* launches) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        double a = 0.0, b = 0.0;
        for (size_t i = 0; i < c.kev_used; ++i) {
            HIP_CHECK(hipEventSynchronize(c.kev[i][2]));
            float x = 0.0f, y = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&x, c.kev[i][0], c.kev[i][1]));
            HIP_CHECK(hipEventElapsedTime(&y, c.kev[i][1], c.kev[i][2]));
            a += x;
            b += y;
        }
        if (front_ms) *front_ms = a;
        if (fused_ms) *fused_ms = b;
        if (launches) *launches = (int)c.kev_used;
    });
}

void* ceos_get_stream(const ceos_index* h) {
    if (!h) return nullptr;
    try {
        return reinterpret_cast<void*>(const_cast<Core&>(h->core).st());
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// ------------------------------- streamCEOs -------------------------------------

int ceos_stream_create(int n_features, ceos_stream** out) {
    return guarded([&] {
        if (!out) throw ArgError("null output handle");
        if (n_features <= 0 || n_features > 4096) throw ArgError("n_features must be in [1, 4096]");
        auto* h = new ceos_stream();
        h->core.d = n_features;
        *out = h;
    });
}

void ceos_stream_destroy(ceos_stream* h) { delete h; }

int ceos_stream_set_index_param(ceos_stream* h, int n_proj, int n_repeats, int top_points, int n_threads,
                                int seed) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.set_index_param(n_proj, n_repeats, top_points, n_threads, seed);
        h->core.s = top_points;  // src/streamCEOs.h:58
    });
}

int ceos_stream_set_query_param(ceos_stream* h, int s, int mp, int c) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.s = s; h->core.mp = mp; h->core.c = c;
    });
}

int ceos_stream_get_query_param(const ceos_stream* h, int* s, int* mp, int* c) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (s) *s = h->core.s;
        if (mp) *mp = h->core.mp;
        if (c) *c = h->core.c;
    });
}

// centering (streamCEOs.h:44): the reference has the field but does not bind it
// (python_wrapper.cpp:52-71); it takes effect at the next build.
int ceos_stream_set_centering(ceos_stream* h, int centering) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.centering = centering != 0;
    });
}
int ceos_stream_get_centering(const ceos_stream* h, int* centering) {
    return guarded([&] {
        if (!h || !centering) throw ArgError("null argument");
        *centering = h->core.centering ? 1 : 0;
    });
}

int ceos_stream_set_threads(ceos_stream* h, int n_threads) {
    return guarded([&] { if (!h) throw ArgError("null handle"); h->core.threads = n_threads; });
}
int ceos_stream_get_threads(const ceos_stream* h, int* n_threads) {
    return guarded([&] { if (!h || !n_threads) throw ArgError("null argument"); *n_threads = h->core.threads; });
}
int ceos_stream_get_dims(const ceos_stream* h, int* fht, int* n_proj, int* n_repeats, int* top_m, int* seed) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        const Core& c = h->core;
        if (fht) *fht = c.F;
        if (n_proj) *n_proj = c.D;
        if (n_repeats) *n_repeats = c.R;
        if (top_m) *top_m = c.m;
        if (seed) *seed = (int)c.seed_used;
    });
}

int ceos_stream_clear(ceos_stream* h) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        h->core.clear();
    });
}

int ceos_stream_build(ceos_stream* h, const float* X, int64_t n, int d) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        if (!c.params_set) throw StateError("setIndexParam must be called before build");
        check_matrix(X, n, d, c.d);
        if (n <= 0) throw ArgError("empty dataset");
        if (n >= (int64_t(1) << 31)) throw ArgError("more than 2^31-1 points");
        c.wait_reads();
        const auto t0 = std::chrono::steady_clock::now();
        c.built = false;
        const int64_t capacity = std::max<int64_t>(n + n / 4, 1024);
        c.dcenter.release();
        c.center.clear();
        c.phase_mark(0, c.st());
        c.load_points(X, n, capacity, false, c.st());
        // centering: rows are stored minus the build rows' column mean, which stays
        // fixed for later updates (src/streamCEOs.cpp:22-34, 198-203)
        if (c.centering) c.center_rows(n, c.st());
        c.phase_mark(1, c.st());
        c.dead.alloc((size_t)(capacity + 31) / 32);
        HIP_CHECK(hipMemsetAsync(c.dead.p, 0, c.dead.n * 4, c.st()));
        c.make_signs();
        c.build_buckets(n, c.st());
        c.quantize_rows(0, n, n, c.st());
        HIP_CHECK(hipStreamSynchronize(c.st()));
        c.finish_build_times(t0);
        c.built = true;
    });
}

int ceos_stream_update(ceos_stream* h, const float* Xn, int64_t n_new, int d, int n_del) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        if (!c.built) throw StateError("stream index is not built");
        check_matrix(Xn, n_new, d, c.d);
        const int64_t n_cur = c.live;
        if (n_del < 0) throw ArgError("n_delPoints must be >= 0");
        if (n_del > n_cur) throw ArgError("Number of removed points must be smaller than the current number of points");
        if (c.m > n_cur + n_new - n_del) throw ArgError("There is not enough indexed top-points for coCEOs after update");
        if (n_cur + n_new >= (int64_t(1) << 31)) throw ArgError("more than 2^31-1 live points");
        c.wait_reads();
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t st = c.st();
        // grow the ring when the new rows do not fit
        if (n_cur + n_new > c.cap) {
            const int64_t ncap = std::max<int64_t>(2 * (n_cur + n_new), 1024);
            DBuf<float> nx((size_t)ncap * c.d);
            DBuf<uint32_t> nd((size_t)(ncap + 31) / 32);
            HIP_CHECK(hipMemsetAsync(nd.p, 0, nd.n * 4, st));
            const int64_t first = std::min<int64_t>(n_cur, c.cap - c.head);
            HIP_CHECK(hipMemcpyAsync(nx.p, c.X.p + c.head * c.d, (size_t)first * c.d * 4, hipMemcpyDeviceToDevice, st));
            if (n_cur > first)
                HIP_CHECK(hipMemcpyAsync(nx.p + first * c.d, c.X.p, (size_t)(n_cur - first) * c.d * 4,
                                         hipMemcpyDeviceToDevice, st));
            // carry random-delete marks over to the new slots
            if (c.dead.p) {
                std::vector<uint32_t> od(c.dead.n), ndh(nd.n, 0u);
                HIP_CHECK(hipMemcpyAsync(od.data(), c.dead.p, od.size() * 4, hipMemcpyDeviceToHost, st));
                HIP_CHECK(hipStreamSynchronize(st));
                for (int64_t i = 0; i < n_cur; ++i) {
                    int64_t sl = c.head + i;
                    if (sl >= c.cap) sl -= c.cap;
                    if ((od[sl >> 5] >> (sl & 31)) & 1u) ndh[i >> 5] |= 1u << (i & 31);
                }
                HIP_CHECK(hipMemcpyAsync(nd.p, ndh.data(), ndh.size() * 4, hipMemcpyHostToDevice, st));
            }
            HIP_CHECK(hipStreamSynchronize(st));
            c.X = std::move(nx);
            c.dead = std::move(nd);
            c.cap = ncap;
            c.head = 0;
        }
        // append rows at slots head+n_cur .. (ring), clearing their delete marks
        const uint32_t first_abs = c.base + (uint32_t)n_cur;
        int64_t slot0 = c.head + n_cur;
        if (slot0 >= c.cap) slot0 -= c.cap;
        const int64_t seg1 = std::min<int64_t>(n_new, c.cap - slot0);
        c.phase_mark(0, st);
        if (n_new > 0) {
            HIP_CHECK(hipMemcpyAsync(c.X.p + slot0 * c.d, Xn, (size_t)seg1 * c.d * 4, hipMemcpyHostToDevice, st));
            if (n_new > seg1)
                HIP_CHECK(hipMemcpyAsync(c.X.p, Xn + seg1 * c.d, (size_t)(n_new - seg1) * c.d * 4,
                                         hipMemcpyHostToDevice, st));
            // centering: new rows minus the build's centre (src/streamCEOs.cpp:198-203)
            if (c.centering && c.dcenter.p)
                HIP_CHECK(ceos::launch_subtract_center(c.X.p, slot0, n_new, c.cap, c.d, c.dcenter.p, st));
        }
        c.phase_mark(1, st);
        if (n_new > 0) {
            if (c.dead.p) HIP_CHECK(ceos::launch_clear_slot_bits(c.dead.p, slot0, n_new, c.cap, st));
            c.quantize_rows(slot0, n_new, n_cur + n_new, st);
        }
        // project the new rows: column-major [DR][n_new], through a row-major projection
        // (coalesced stores) and an LDS-tiled transpose; scratch kept on the index for
        // updates up to kUpdScratchKeep floats, allocated for this update above that
        const size_t proj_n = (size_t)c.DR() * std::max<int64_t>(n_new, 1);
        const bool keep = proj_n <= Core::kUpdScratchKeep;
        DBuf<float> np_own, tmp_own;
        if (keep) {
            c.upd_lists.ensure(proj_n);
        } else {
            np_own.alloc(proj_n);
        }
        float* np_p = keep ? c.upd_lists.p : np_own.p;
        if (n_new > 0 && n_new <= int64_t(65535) * 64) {
            if (keep) c.upd_rows.ensure(proj_n); else tmp_own.alloc(proj_n);
            float* tmp_p = keep ? c.upd_rows.p : tmp_own.p;
            HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p + slot0 * c.d, 0, 1, seg1, c.signs.p, tmp_p, c.DR(),
                                                false, st));
            if (n_new > seg1)
                HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p, 0, 1, n_new - seg1, c.signs.p,
                                                    tmp_p + seg1 * c.DR(), c.DR(), false, st));
            HIP_CHECK(ceos::launch_transpose(tmp_p, n_new, c.DR(), np_p, st, n_new));
            if (!keep) HIP_CHECK(hipStreamSynchronize(st));  // tmp_own is freed on return
        } else if (n_new > 0) {
            HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p + slot0 * c.d, 0, 1, seg1, c.signs.p, np_p, n_new,
                                                true, st));
            if (n_new > seg1)
                HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p, 0, 1, n_new - seg1, c.signs.p, np_p + seg1,
                                                    n_new, true, st));
        }
        const uint32_t new_base = c.base + (uint32_t)n_del;
        const uint32_t new_live = (uint32_t)(n_cur + n_new - n_del);
        c.phase_mark(2, st);
        HIP_CHECK(ceos::launch_stream_merge(c.DR(), c.m, c.bkt.p, c.bsize.p, np_p, n_new, first_abs, new_base,
                                            new_live, st));
        c.phase_mark(3, st);
        HIP_CHECK(hipStreamSynchronize(st));
        c.base = new_base;
        c.live = new_live;
        c.head = (c.head + n_del) % c.cap;
        c.finish_build_times(t0);
    });
}

int ceos_stream_delete_ids(ceos_stream* h, const int32_t* ids, int64_t count) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        if (!c.built) throw StateError("stream index is not built");
        if (count < 0 || (!ids && count > 0)) throw ArgError("invalid id list");
        if (count == 0) return;
        for (int64_t i = 0; i < count; ++i)
            if (ids[i] < 0 || (uint32_t)ids[i] >= c.live) throw ArgError("id " + std::to_string(ids[i]) + " is not live");
        c.wait_reads();
        // the ids go to mapped pinned memory and a kernel sets their bits: the bitmap
        // never crosses PCIe, and no copy-engine transfer is queued
        c.del_ids.ensure((size_t)count);
        std::memcpy(c.del_ids.p, ids, (size_t)count * sizeof(int32_t));
        HIP_CHECK(ceos::launch_mark_dead(c.del_ids.dp, count, c.head, c.cap, c.dead.p, c.st()));
        HIP_CHECK(ceos::launch_bucket_purge(2 * c.DR(), c.m, c.bkt.p, c.bsize.p, c.dead.p, c.rows(), c.st()));
        HIP_CHECK(hipStreamSynchronize(c.st()));
    });
}

int ceos_stream_estimate_search(ceos_stream* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        c.validate_query(c.s, c.mp, c.c, k, c.D, "number of projections");  // src/streamCEOs.cpp:386
        c.search_host(Q, nq, k, c.s, c.c, ids, ips);
    });
}

// hash_search(queries, n_neighbors) — src/streamCEOs.cpp:621-877
int ceos_stream_hash_search(ceos_stream* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(Q, nq, d, c.d);
        c.validate_query(c.s, c.mp, 1, k, c.D, "number of projections");  // src/streamCEOs.cpp:623-633
        c.search_host(Q, nq, k, c.s, 1, ids, ips, true);
    });
}

int ceos_stream_save(const ceos_stream* h, const char* path) {
    return guarded([&] {
        if (!h || !path) throw ArgError("null argument");
        auto& c = const_cast<Core&>(h->core);
        std::lock_guard<std::mutex> g(c.mu);
        c.save(path, 1);
    });
}

int ceos_stream_load(ceos_stream* h, const char* path) {
    return guarded([&] {
        if (!h || !path) throw ArgError("null argument");
        std::lock_guard<std::mutex> g(h->core.mu);
        (void)h->core.load(path, 1);
    });
}

int ceos_stream_size(const ceos_stream* h, int64_t* n) {
    return guarded([&] {
        if (!h || !n) throw ArgError("null handle");
        *n = h->core.live;
    });
}

int ceos_stream_get_buckets(const ceos_stream* h, int32_t* cid, float* cval, int32_t* fid, float* fval,
                            int32_t* csz, int32_t* fsz) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        auto& c = const_cast<Core&>(h->core);
        std::lock_guard<std::mutex> g(c.mu);
        c.get_buckets(cid, cval, fid, fval, csz, fsz);
    });
}

// ------------------------- phase times and the centre -----------------------------

static int phase_times(const Core& c, bool search, ceos_phase_times* out) {
    return guarded([&] {
        if (!out) throw ArgError("null argument");
        *out = search ? c.last_search : c.last_build;
    });
}
int ceos_get_build_times(const ceos_index* h, ceos_phase_times* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return phase_times(h->core, false, out);
}
int ceos_get_search_times(const ceos_index* h, ceos_phase_times* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return phase_times(h->core, true, out);
}
int ceos_stream_get_build_times(const ceos_stream* h, ceos_phase_times* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return phase_times(h->core, false, out);
}
int ceos_stream_get_search_times(const ceos_stream* h, ceos_phase_times* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return phase_times(h->core, true, out);
}

static int get_center(const Core& c, float* out) {
    return guarded([&] {
        if (!out) throw ArgError("null argument");
        if (c.center.empty()) throw StateError("no centre: centering was off at the last build");
        std::memcpy(out, c.center.data(), c.center.size() * 4);
    });
}
int ceos_get_center(const ceos_index* h, float* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return get_center(h->core, out);
}
int ceos_stream_get_center(const ceos_stream* h, float* out) {
    if (!h) { g_err = "null handle"; return CEOS_EINVAL; }
    return get_center(h->core, out);
}

// ------------------------------- search plans ------------------------------------

int ceos_plan_create(ceos_index* h, int nq_max, int k_max, int s_max, int c_max, int flags, ceos_plan** out) {
    if (!h) {
        g_err = "null handle";
        return CEOS_EINVAL;
    }
    return make_plan(h->core, false, nq_max, k_max, s_max, c_max, flags, out);
}

int ceos_stream_plan_create(ceos_stream* h, int nq_max, int k_max, int s_max, int c_max, int flags,
                            ceos_plan** out) {
    if (!h) {
        g_err = "null handle";
        return CEOS_EINVAL;
    }
    return make_plan(h->core, true, nq_max, k_max, s_max, c_max, flags, out);
}

void ceos_plan_destroy(ceos_plan* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(p->core->mu);
    delete p;
}

int ceos_plan_buffers(ceos_plan* p, float** dQ, int32_t** dk, int32_t** ds, int32_t** dc, int32_t** d_ids,
                      float** d_ips) {
    return guarded([&] {
        if (!p) throw ArgError("null plan");
        if (dQ) *dQ = p->Q.p;
        if (dk) *dk = p->k.p;
        if (ds) *ds = p->s.p;
        if (dc) *dc = p->c.p;
        if (d_ids) *d_ids = p->ids.p;
        if (d_ips) *d_ips = p->ips.p;
    });
}

int ceos_plan_run(ceos_plan* p, int nq, void* stream) {
    return guarded([&] {
        if (!p) throw ArgError("null plan");
        std::lock_guard<std::mutex> g(p->core->mu);
        p->run(nq, p->core->stream(stream));
    });
}

int ceos_plan_search(ceos_plan* p, const float* Q, int nq, int d, const int32_t* k, const int32_t* s,
                     const int32_t* c, int32_t* ids, float* ips) {
    return guarded([&] {
        if (!p) throw ArgError("null plan");
        std::lock_guard<std::mutex> g(p->core->mu);
        Core& x = *p->core;
        check_matrix(Q, nq, d, x.d);
        if (nq > p->nq_max) throw ArgError("batch of " + std::to_string(nq) + " exceeds the plan's nq_max");
        if (nq <= 0) return;
        if (!k || !s || !c || !ids || !ips) throw ArgError("null array");
        for (int q = 0; q < nq; ++q)
            if (k[q] < 0 || k[q] > p->k_max || s[q] <= 0 || s[q] > p->s_max || c[q] <= 0 || c[q] > p->c_max)
                throw ArgError("query " + std::to_string(q) + ": k, n_probed_vectors or n_cand outside the plan's bounds");
        hipStream_t st = x.st();
        HIP_CHECK(hipMemcpyAsync(p->Q.p, Q, (size_t)nq * d * 4, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(p->k.p, k, (size_t)nq * 4, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(p->s.p, s, (size_t)nq * 4, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(p->c.p, c, (size_t)nq * 4, hipMemcpyHostToDevice, st));
        p->run(nq, st);
        HIP_CHECK(hipMemcpyAsync(ids, p->ids.p, (size_t)nq * p->k_max * 4, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(ips, p->ips.p, (size_t)nq * p->k_max * 4, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        int32_t flag = 0;
        HIP_CHECK(hipMemcpy(&flag, p->scr.status.p, 4, hipMemcpyDeviceToHost));
        if (flag) {
            HIP_CHECK(hipMemset(p->scr.status.p, 0, 4));
            throw DevError("LDS score tile overflowed");
        }
    });
}

int ceos_plan_config(const ceos_plan* p, int* wg_per_cu, int* table_log2, double* tuned_ms) {
    return guarded([&] {
        if (!p) throw ArgError("null plan");
        if (wg_per_cu) *wg_per_cu = p->cfg.grid > 0 ? p->cfg.grid / std::max(1, device_cus()) : 0;
        if (table_log2) *table_log2 = p->tl;  // 0: the automatic tile (most buckets that fit)
        if (tuned_ms) *tuned_ms = p->tuned ? p->tuned_ms : 0.0;
    });
}

int ceos_plan_tuning(const ceos_plan* p, int cap, int32_t* shapes, double* ms, int* count) {
    return guarded([&] {
        if (!p || !count || cap < 0 || (cap > 0 && (!shapes || !ms))) throw ArgError("null argument");
        const int nrec = (int)p->tune_log.size();
        for (int i = 0; i < std::min(cap, nrec); ++i) {
            shapes[3 * i] = p->tune_log[i].wg;
            shapes[3 * i + 1] = p->tune_log[i].tl;
            shapes[3 * i + 2] = p->tune_log[i].pw;
            ms[i] = p->tune_log[i].ms;
        }
        *count = nrec;
    });
}

int ceos_plan_proj_waves(const ceos_plan* p, int* waves_per_wg) {
    return guarded([&] {
        if (!p || !waves_per_wg) throw ArgError("null argument");
        *waves_per_wg = p->pw;
    });
}

}  // extern "C"
