// pybind11 module `CEOs`: the reference's Python surface (python/python_wrapper.cpp:13-75)
// over the C ABI of libceos_amd (include/ceos_amd.h).
//
// Same class, method, argument and field names and defaults. Differences, all
// stricter: invalid parameters raise ValueError instead of exit(1)
// (src/CEOs.cpp:452-461); device failures raise RuntimeError; the GIL is released
// while the device works.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ceos_amd.h"

namespace py = pybind11;

namespace {

using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using IArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

void check(int rc) {
    if (rc == CEOS_OK) return;
    const std::string msg = ceos_last_error();
    if (rc == CEOS_EINVAL) throw py::value_error(msg);
    if (rc == CEOS_ENOMEM) throw std::runtime_error("device memory: " + msg);
    throw std::runtime_error(msg);
}

void require_2d(const FArr& a, const char* what) {
    if (a.ndim() != 2) throw py::value_error(std::string(what) + " must be a 2-D array (rows x n_features)");
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The search phase lines printed when verbose (src/CEOs.cpp:676-680): the device runs
// score accumulation, candidate selection and the rerank as one fused kernel, so they
// share one line.
void print_search_phases(const ceos_phase_times& t, const char* proj_label) {
    std::printf("%s: %g ms\n", proj_label, t.project_ms);
    std::printf("Estimating, extracting candidates and computing distance time (one fused kernel): %g ms\n",
                t.select_ms);
    std::printf("Querying time: %g ms\n", t.total_ms);
}

py::tuple result(int nq, int k, py::array_t<int32_t>& ids, py::array_t<float>& ips) {
    (void)nq; (void)k;
    return py::make_tuple(ids, ips);
}

// A captured batched-search plan (ceos_plan_*); keeps its index alive.
class PyPlan {
   public:
    explicit PyPlan(ceos_plan* p, int d, int k_max) : p_(p), d_(d), k_max_(k_max) {}
    ~PyPlan() { ceos_plan_destroy(p_); }
    PyPlan(const PyPlan&) = delete;
    PyPlan& operator=(const PyPlan&) = delete;

    py::tuple search(FArr Q, IArr k, IArr s, IArr c) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k.size() != nq || s.size() != nq || c.size() != nq)
            throw py::value_error("k, n_probed_vectors and n_cand arrays must have one entry per query");
        py::array_t<int32_t> ids({nq, k_max_});
        py::array_t<float> ips({nq, k_max_});
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_plan_search(p_, Q.data(), nq, (int)Q.shape(1), k.data(), s.data(), c.data(),
                                  ids.mutable_data(), ips.mutable_data());
        }
        check(rc);
        return py::make_tuple(ids, ips);
    }
    void run(int nq, uintptr_t stream) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_plan_run(p_, nq, reinterpret_cast<void*>(stream));
        }
        check(rc);
    }
    py::dict buffers() {
        float *q, *ip;
        int32_t *k, *s, *c, *id;
        check(ceos_plan_buffers(p_, &q, &k, &s, &c, &id, &ip));
        py::dict out;
        out["queries"] = reinterpret_cast<uintptr_t>(q);
        out["k"] = reinterpret_cast<uintptr_t>(k);
        out["n_probed_vectors"] = reinterpret_cast<uintptr_t>(s);
        out["n_cand"] = reinterpret_cast<uintptr_t>(c);
        out["ids"] = reinterpret_cast<uintptr_t>(id);
        out["ips"] = reinterpret_cast<uintptr_t>(ip);
        return out;
    }
    py::tuple config() const {
        int wg = 0, tl = 0;
        double ms = 0.0;
        check(ceos_plan_config(p_, &wg, &tl, &ms));
        return py::make_tuple(wg, tl, ms);
    }
    int proj_waves() const {
        int pw = 0;
        check(ceos_plan_proj_waves(p_, &pw));
        return pw;
    }
    // [(wg_per_cu, table_log2, proj_waves, median ms), ...] of the autotuner
    py::list tuning() const {
        int cnt = 0;
        check(ceos_plan_tuning(p_, 0, nullptr, nullptr, &cnt));
        std::vector<int32_t> sh(3 * (size_t)cnt);
        std::vector<double> ms((size_t)cnt);
        if (cnt > 0) check(ceos_plan_tuning(p_, cnt, sh.data(), ms.data(), &cnt));
        py::list out;
        for (int i = 0; i < cnt; ++i) out.append(py::make_tuple(sh[3 * i], sh[3 * i + 1], sh[3 * i + 2], ms[i]));
        return out;
    }

   private:
    ceos_plan* p_;
    int d_, k_max_;
};

class PyCEOs {
   public:
    PyCEOs(int n_points, int n_features) : n_(n_points), d_(n_features) { check(ceos_create(n_points, n_features, &h_)); }
    ~PyCEOs() { ceos_destroy(h_); }

    void setIndexParam(int n_proj, int n_repeats, int top_m, int n_threads, int random_seed) {
        check(ceos_set_index_param(h_, n_proj, n_repeats, top_m, n_threads, random_seed));
    }

    int get_s() const { int s, mp, c; check(ceos_get_query_param(h_, &s, &mp, &c)); return s; }
    int get_mp() const { int s, mp, c; check(ceos_get_query_param(h_, &s, &mp, &c)); return mp; }
    int get_c() const { int s, mp, c; check(ceos_get_query_param(h_, &s, &mp, &c)); return c; }
    void set_s(int v) { check(ceos_set_query_param(h_, v, get_mp(), get_c())); }
    void set_mp(int v) { check(ceos_set_query_param(h_, get_s(), v, get_c())); }
    void set_c(int v) { check(ceos_set_query_param(h_, get_s(), get_mp(), v)); }
    bool get_centering() const { int v; check(ceos_get_centering(h_, &v)); return v != 0; }
    void set_centering(bool v) { check(ceos_set_centering(h_, v ? 1 : 0)); }
    void set_threads(int t) { check(ceos_set_threads(h_, t)); }
    void clear() { check(ceos_clear(h_)); }

    void print_params(const char* title, bool top_m = true) const {
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        std::printf("%s\nn_points: %d\nn_features: %d\nn_repeats: %d\nn_proj: %d\n", title, n_, d_, R, D);
        if (top_m) std::printf("top_m: %d\n", m);
        std::printf("fhtDim: %d\n", F);
    }
    int threads() const { int t = 0; check(ceos_get_threads(h_, &t)); return t; }
    py::array_t<float> center() const {
        py::array_t<float> out(d_);
        check(ceos_get_center(h_, out.mutable_data()));
        return out;
    }
    ceos_phase_times build_times() const { ceos_phase_times t{}; check(ceos_get_build_times(h_, &t)); return t; }
    ceos_phase_times search_times() const { ceos_phase_times t{}; check(ceos_get_search_times(h_, &t)); return t; }

    void build_coCEOs_Est(FArr X) {
        require_2d(X, "dataset");
        print_params("Building coCEOs-Estimate index...");
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_build_coceos(h_, X.data(), X.shape(0), (int)X.shape(1));
        }
        check(rc);
        const ceos_phase_times t = build_times();  // src/CEOs.cpp:326-328, 434-440
        std::printf("Copying data time (in seconds): %g\n", t.copy_ms / 1000.0);
        std::printf("Size of coCEOs-Est index in GB: %g\n", t.index_gb);
        std::printf("Projecting time: %g ms\n", t.project_ms);
        std::printf("Extracting top-m points time: %g ms\n", t.select_ms);
        std::printf("Constructing time (in seconds): %.3f\n", ms_since(t0) / 1000.0);
        std::fflush(stdout);
    }

    void build_CEOs_Hash(FArr X) {
        require_2d(X, "dataset");
        print_params("Building CEOs-Hash index...");
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_build_hash(h_, X.data(), X.shape(0), (int)X.shape(1));
        }
        check(rc);
        const ceos_phase_times t = build_times();  // src/CEOs.cpp:731, 834-839
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        std::printf("Copying and centering data time in seconds: %g\n", t.copy_ms / 1000.0);
        std::printf("Size of CEOs-Hash index in GB: %g\n",
                    (2.0 * m * D * R + (double)X.shape(0) * d_) * 4.0 / (double)(1 << 30));
        std::printf("Projecting time: %g ms\n", t.project_ms);
        std::printf("Extracting top-points time: %g ms\n", t.select_ms);
        std::printf("Constructing time (in seconds): %.3f\n", ms_since(t0) / 1000.0);
        std::fflush(stdout);
    }

    py::tuple search_CEOs_Hash(FArr Q, int k, bool verbose) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k < 0) throw py::value_error("n_neighbors must be >= 0");
        py::array_t<int32_t> ids({nq, k});
        py::array_t<float> ips({nq, k});
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_search_hash(h_, Q.data(), nq, (int)Q.shape(1), k, ids.mutable_data(), ips.mutable_data());
        }
        check(rc);
        if (verbose) {  // src/CEOs.cpp:866-870, 1047-1052
            std::printf("n_queries: %d\nn_probedVectors: %d\nn_probedPoints: %d\nn_threads: %d\n", nq, get_s(), get_mp(),
                        threads());
            print_search_phases(search_times(), "Projecting and extracting top-vectors time");
            std::fflush(stdout);
        }
        return result(nq, k, ids, ips);
    }

    py::tuple search_coCEOs_Est(FArr Q, int k, bool verbose) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k < 0) throw py::value_error("n_neighbors must be >= 0");
        py::array_t<int32_t> ids({nq, k});
        py::array_t<float> ips({nq, k});
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_search_coceos(h_, Q.data(), nq, (int)Q.shape(1), k, ids.mutable_data(), ips.mutable_data());
        }
        check(rc);
        if (verbose) {  // src/CEOs.cpp:466-470, 676-680
            std::printf("n_queries: %d\nn_probedVectors: %d\nn_probedPoints: %d\nn_cand: %d\nn_threads: %d\n", nq,
                        get_s(), get_mp(), get_c(), threads());
            print_search_phases(search_times(), "Projecting and extracting top-vectors time");
            std::fflush(stdout);
        }
        return result(nq, k, ids, ips);
    }

    // Batched form: per-query k, n_probed_vectors and n_cand. Returns (ids, ips, offsets)
    // with query q's results at offsets[q] .. offsets[q+1]-1.
    void save(const std::string& path) const {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_save(h_, path.c_str());
        }
        check(rc);
    }
    void load(const std::string& path) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_load(h_, path.c_str());
        }
        check(rc);
    }

    PyPlan* make_plan(int nq_max, int k_max, int s_max, int c_max, bool autotune) {
        ceos_plan* p = nullptr;
        check(ceos_plan_create(h_, nq_max, k_max, s_max, c_max, autotune ? CEOS_PLAN_AUTOTUNE : 0, &p));
        return new PyPlan(p, d_, k_max);
    }

    py::tuple search_batched(FArr Q, IArr k, IArr s, IArr c) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k.size() != nq || s.size() != nq || c.size() != nq)
            throw py::value_error("k, n_probed_vectors and n_cand arrays must have one entry per query");
        py::array_t<int64_t> off(nq + 1);
        int64_t tot = 0;
        auto o = off.mutable_unchecked<1>();
        for (int q = 0; q < nq; ++q) { o(q) = tot; tot += std::max<int32_t>(0, k.data()[q]); }
        o(nq) = tot;
        py::array_t<int32_t> ids(tot);
        py::array_t<float> ips(tot);
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_search_batched(h_, Q.data(), nq, (int)Q.shape(1), k.data(), s.data(), c.data(),
                                     ids.mutable_data(), ips.mutable_data());
        }
        check(rc);
        return py::make_tuple(ids, ips, off);
    }

    void build(FArr X) {
        require_2d(X, "dataset");
        print_params("Building CEOs index...", false);
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_build_est(h_, X.data(), X.shape(0), (int)X.shape(1));
        }
        check(rc);
        const ceos_phase_times t = build_times();  // src/CEOs.cpp:47, 91-97
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        std::printf("Copying data time (in seconds): %g\n", t.copy_ms / 1000.0);
        std::printf("Size of CEOs index in GB: %g\n", (double)X.shape(0) * D * R * 4.0 / (double)(1 << 30));
        std::printf("Size of data set in GB: %g\n", (double)X.shape(0) * d_ * 4.0 / (double)(1 << 30));
        std::printf("Index construction time (in seconds): %.3f\n", ms_since(t0) / 1000.0);
        std::fflush(stdout);
    }

    py::tuple search(FArr Q, int k, bool verbose) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k < 0) throw py::value_error("n_neighbors must be >= 0");
        py::array_t<int32_t> ids({nq, k});
        py::array_t<float> ips({nq, k});
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_search_est(h_, Q.data(), nq, (int)Q.shape(1), k, ids.mutable_data(), ips.mutable_data());
        }
        check(rc);
        if (verbose) {  // src/CEOs.cpp:113-117, 280-284 (one fused kernel: the total only)
            std::printf("n_queries: %d\nn_probed_vectors: %d\nn_probed_points: %d\nn_cand: %d\nn_threads: %d\n", nq,
                        get_s(), get_mp(), get_c(), threads());
            std::printf("Querying time: %g ms\n", ms_since(t0));
            std::fflush(stdout);
        }
        return result(nq, k, ids, ips);
    }

    // Parity helpers (not in the reference surface).
    py::tuple get_buckets() const {
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        const int DR = D * R;
        py::array_t<int32_t> cid({DR, m}), fid({DR, m});
        py::array_t<float> cv({DR, m}), fv({DR, m});
        check(ceos_get_buckets(h_, cid.mutable_data(), cv.mutable_data(), fid.mutable_data(), fv.mutable_data()));
        return py::make_tuple(cid, cv, fid, fv);
    }
    py::array_t<uint8_t> get_sign_bits() const {
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        py::array_t<uint8_t> out((size_t)F * 3 * R);
        check(ceos_get_sign_bits(h_, out.mutable_data()));
        return out;
    }
    py::array_t<float> project(FArr rows) {
        require_2d(rows, "rows");
        int F, D, R, m, seed;
        check(ceos_get_dims(h_, &F, &D, &R, &m, &seed));
        py::array_t<float> out({(int64_t)rows.shape(0), (int64_t)D * R});
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_project(h_, rows.data(), rows.shape(0), (int)rows.shape(1), out.mutable_data());
        }
        check(rc);
        return out;
    }
    int fht_dim() const { int F, D, R, m, seed; check(ceos_get_dims(h_, &F, &D, &R, &m, &seed)); return F; }
    void set_search_config(int wg_per_cu, int table_log2) { check(ceos_set_search_config(h_, wg_per_cu, table_log2)); }
    bool get_rerank_bounds() const { int v; check(ceos_get_rerank_bounds(h_, &v)); return v != 0; }
    void set_rerank_bounds(bool v) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_set_rerank_bounds(h_, v ? 1 : 0);
        }
        check(rc);
    }
    py::tuple get_search_config() const {
        int wg, t, lds;
        check(ceos_get_search_config(h_, &wg, &t, &lds));
        return py::make_tuple(wg, t, lds);
    }
    uintptr_t handle() const { return reinterpret_cast<uintptr_t>(h_); }
    ceos_index* raw() const { return h_; }
    int dim() const { return d_; }

   private:
    ceos_index* h_ = nullptr;
    int n_, d_;
};

class PyStream {
   public:
    explicit PyStream(int n_features) : d_(n_features) { check(ceos_stream_create(n_features, &h_)); }
    ~PyStream() { ceos_stream_destroy(h_); }

    ceos_stream* raw() const { return h_; }
    bool get_centering() const { int v; check(ceos_stream_get_centering(h_, &v)); return v != 0; }
    void set_centering(bool v) { check(ceos_stream_set_centering(h_, v ? 1 : 0)); }
    bool get_rerank_bounds() const { int v; check(ceos_stream_get_rerank_bounds(h_, &v)); return v != 0; }
    void set_rerank_bounds(bool v) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_set_rerank_bounds(h_, v ? 1 : 0);
        }
        check(rc);
    }
    int dim() const { return d_; }
    int threads() const { int t = 0; check(ceos_stream_get_threads(h_, &t)); return t; }
    ceos_phase_times search_times() const {
        ceos_phase_times t{};
        check(ceos_stream_get_search_times(h_, &t));
        return t;
    }
    py::array_t<float> center() const {
        py::array_t<float> out(d_);
        check(ceos_stream_get_center(h_, out.mutable_data()));
        return out;
    }

    void save(const std::string& path) const {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_save(h_, path.c_str());
        }
        check(rc);
    }
    void load(const std::string& path) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_load(h_, path.c_str());
        }
        check(rc);
    }

    PyPlan* make_plan(int nq_max, int k_max, int s_max, int c_max, bool autotune) {
        ceos_plan* p = nullptr;
        check(ceos_stream_plan_create(h_, nq_max, k_max, s_max, c_max, autotune ? CEOS_PLAN_AUTOTUNE : 0, &p));
        return new PyPlan(p, d_, k_max);
    }

    void setIndexParam(int n_proj, int n_repeats, int top_points, int n_threads, int random_seed) {
        check(ceos_stream_set_index_param(h_, n_proj, n_repeats, top_points, n_threads, random_seed));
    }
    int get_s() const { int s, mp, c; check(ceos_stream_get_query_param(h_, &s, &mp, &c)); return s; }
    int get_mp() const { int s, mp, c; check(ceos_stream_get_query_param(h_, &s, &mp, &c)); return mp; }
    int get_c() const { int s, mp, c; check(ceos_stream_get_query_param(h_, &s, &mp, &c)); return c; }
    void set_s(int v) { check(ceos_stream_set_query_param(h_, v, get_mp(), get_c())); }
    void set_mp(int v) { check(ceos_stream_set_query_param(h_, get_s(), v, get_c())); }
    void set_c(int v) { check(ceos_stream_set_query_param(h_, get_s(), get_mp(), v)); }
    void set_threads(int t) { check(ceos_stream_set_threads(h_, t)); }
    void clear() { check(ceos_stream_clear(h_)); }

    void build(FArr X) {
        require_2d(X, "dataset");
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_build(h_, X.data(), X.shape(0), (int)X.shape(1));
        }
        check(rc);
        int F, D, R, m, seed;
        check(ceos_stream_get_dims(h_, &F, &D, &R, &m, &seed));
        ceos_phase_times t{};
        check(ceos_stream_get_build_times(h_, &t));  // src/streamCEOs.cpp:11-20, 37, 154-160
        std::printf("n_features: %d\nn_proj: %d\ntop_m: %d\nfhtDim: %d\ncentering: %d\nn_points: %lld\n", d_, D, m, F,
                    get_centering() ? 1 : 0, (long long)X.shape(0));
        std::printf("Copying data time (in ms): %g ms\n", t.copy_ms);
        std::printf("Size of coCEOs-Est index in GB: %g\n", t.index_gb);
        std::printf("Projecting time (in ms): %g ms\n", t.project_ms);
        std::printf("Extracting top-points time (in ms): %g ms\n", t.select_ms);
        std::printf("Constructing time (in seconds): %.3f seconds\n", ms_since(t0) / 1000.0);
        std::fflush(stdout);
    }
    void update(FArr newX, int n_del) {
        require_2d(newX, "new_dataset");
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_update(h_, newX.data(), newX.shape(0), (int)newX.shape(1), n_del);
        }
        check(rc);
        ceos_phase_times t{};
        check(ceos_stream_get_build_times(h_, &t));  // src/streamCEOs.cpp:212-215, 362-368
        std::printf("n_points (after update): %lld\n", (long long)size());
        std::printf("Updating data time (in ms): %g ms\n", t.copy_ms);
        std::printf("Size of coCEOs-Est index in GB: %g\n", t.index_gb);
        std::printf("Projecting time (in ms): %g ms\n", t.project_ms);
        std::printf("Extracting top-points time (in ms): %g ms\n", t.select_ms);
        std::printf("Updating time (in ms): %g ms\n", t.total_ms);
        std::fflush(stdout);
    }
    void delete_ids(IArr ids) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_delete_ids(h_, ids.data(), ids.size());
        }
        check(rc);
    }
    py::tuple estimate_search(FArr Q, int k, bool verbose) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k < 0) throw py::value_error("n_neighbors must be >= 0");
        py::array_t<int32_t> ids({nq, k});
        py::array_t<float> ips({nq, k});
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_estimate_search(h_, Q.data(), nq, (int)Q.shape(1), k, ids.mutable_data(),
                                             ips.mutable_data());
        }
        check(rc);
        (void)t0;
        if (verbose) {  // src/streamCEOs.cpp:396-399, 594-598
            std::printf("n_queries: %d\nn_probedVectors: %d\nn_cand: %d\nn_threads: %d\n", nq, get_s(), get_c(), threads());
            print_search_phases(search_times(), "Projecting time");
            std::fflush(stdout);
        }
        return py::make_tuple(ids, ips);
    }

    py::tuple hash_search(FArr Q, int k, bool verbose) {
        require_2d(Q, "queries");
        const int nq = (int)Q.shape(0);
        if (k < 0) throw py::value_error("n_neighbors must be >= 0");
        py::array_t<int32_t> ids({nq, k});
        py::array_t<float> ips({nq, k});
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_stream_hash_search(h_, Q.data(), nq, (int)Q.shape(1), k, ids.mutable_data(),
                                             ips.mutable_data());
        }
        check(rc);
        (void)t0;
        if (verbose) {  // src/streamCEOs.cpp:638-642, 858-863
            std::printf("n_queries: %d\nn_probedVectors: %d\nn_probedPoints: %d\nn_cand: %d\nn_threads: %d\n", nq, get_s(),
                        get_mp(), get_c(), threads());
            print_search_phases(search_times(), "Projecting time");
            std::fflush(stdout);
        }
        return py::make_tuple(ids, ips);
    }
    int64_t size() const { int64_t n; check(ceos_stream_size(h_, &n)); return n; }
    py::tuple get_buckets(int DR, int m) const {
        py::array_t<int32_t> cid({DR, m}), fid({DR, m}), cs(DR), fs(DR);
        py::array_t<float> cv({DR, m}), fv({DR, m});
        check(ceos_stream_get_buckets(h_, cid.mutable_data(), cv.mutable_data(), fid.mutable_data(),
                                      fv.mutable_data(), cs.mutable_data(), fs.mutable_data()));
        return py::make_tuple(cid, cv, fid, fv, cs, fs);
    }

   private:
    ceos_stream* h_ = nullptr;
    int d_;
};

}  // namespace

// Continuous-batching server over a CEOs or streamCEOs index (ceos_server_*).
class PyServer {
   public:
    PyServer(ceos_index* hi, ceos_stream* hs, int d, int max_batch, int k_max, int s_max, int c_max, int inflight,
             int64_t hbm_budget, int64_t max_pending)
        : d_(d), k_max_(k_max) {
        ceos_server_config cfg{max_batch, k_max, s_max, c_max, inflight, hbm_budget, max_pending};
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = hi ? ceos_server_create(hi, &cfg, &s_) : ceos_stream_server_create(hs, &cfg, &s_);
        }
        check(rc);
    }
    ~PyServer() { close(); }
    void close() {
        if (!s_) return;
        py::gil_scoped_release nogil;
        ceos_server_destroy(s_);
        s_ = nullptr;
    }
    ceos_server* live() const {
        if (!s_) throw std::runtime_error("server is closed");
        return s_;
    }

    int64_t submit(FArr Q, IArr k, IArr s, IArr c, py::object arrival_s) {
        require_2d(Q, "queries");
        const int n = (int)Q.shape(0);
        if (k.size() != n || s.size() != n || c.size() != n)
            throw py::value_error("k, n_probed_vectors and n_cand arrays must have one entry per query");
        py::array_t<double, py::array::c_style | py::array::forcecast> arr;
        const double* ap = nullptr;
        if (!arrival_s.is_none()) {
            arr = py::array_t<double, py::array::c_style | py::array::forcecast>(arrival_s);
            if (arr.size() != n) throw py::value_error("arrival_s must have one entry per query");
            ap = arr.data();
        }
        int64_t t0 = 0;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_server_submit(live(), Q.data(), n, (int)Q.shape(1), k.data(), s.data(), c.data(), ap, &t0);
        }
        check(rc);
        return t0;
    }
    py::object result(int64_t ticket, bool wait) {
        std::vector<int32_t> ids(k_max_);
        std::vector<float> ips(k_max_);
        int k = 0;
        double lat = 0.0;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_server_result(live(), ticket, wait ? 1 : 0, ids.data(), ips.data(), &k, &lat);
        }
        if (rc == CEOS_EAGAIN) return py::none();
        check(rc);
        py::array_t<int32_t> I(k);
        py::array_t<float> P(k);
        std::copy(ids.begin(), ids.begin() + k, I.mutable_data());
        std::copy(ips.begin(), ips.begin() + k, P.mutable_data());
        return py::make_tuple(I, P, lat);
    }
    void update(FArr X, int n_del) {
        require_2d(X, "new_dataset");
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_server_update(live(), X.data(), X.shape(0), (int)X.shape(1), n_del);
        }
        check(rc);
    }
    void remove(IArr ids) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_server_delete(live(), ids.data(), ids.size());
        }
        check(rc);
    }
    void drain() {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = ceos_server_drain(live());
        }
        check(rc);
    }
    py::dict stats() {
        ceos_server_stats st{};
        check(ceos_server_get_stats(live(), &st));
        py::dict o;
        o["submitted"] = st.submitted;
        o["completed"] = st.completed;
        o["batches"] = st.batches;
        o["mean_batch"] = st.mean_batch;
        o["p50_ms"] = st.p50_ms;
        o["p99_ms"] = st.p99_ms;
        o["max_ms"] = st.max_ms;
        o["elapsed_s"] = st.elapsed_s;
        o["qps"] = st.qps;
        o["inflight_slots"] = st.inflight_slots;
        o["device_bytes"] = st.device_bytes;
        o["inserts"] = st.inserts;
        o["deletes"] = st.deletes;
        o["insert_mean_ms"] = st.insert_mean_ms;
        o["insert_max_ms"] = st.insert_max_ms;
        o["delete_mean_ms"] = st.delete_mean_ms;
        o["delete_max_ms"] = st.delete_max_ms;
        o["wait_p50_ms"] = st.wait_p50_ms;
        o["wait_p99_ms"] = st.wait_p99_ms;
        o["wait_max_ms"] = st.wait_max_ms;
        o["service_p50_ms"] = st.service_p50_ms;
        o["service_p99_ms"] = st.service_p99_ms;
        o["service_max_ms"] = st.service_max_ms;
        return o;
    }
    // Per-batch trace: columns launch_s, enqueue_ms, done_ms, oldest_wait_ms, queries, slot.
    py::array_t<double> batches() {
        int64_t n = 0;
        check(ceos_server_get_batches(live(), nullptr, 0, &n));
        std::vector<ceos_server_batch> b((size_t)n);
        int64_t got = 0;
        check(ceos_server_get_batches(live(), b.data(), n, &got));
        n = std::min(n, got);
        py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)11});
        auto m = out.mutable_unchecked<2>();
        for (int64_t i = 0; i < n; ++i) {
            m(i, 0) = b[i].launch_s;
            m(i, 1) = b[i].enqueue_ms;
            m(i, 2) = b[i].done_ms;
            m(i, 3) = b[i].oldest_wait_ms;
            m(i, 4) = b[i].queries;
            m(i, 5) = b[i].slot;
            for (int j = 0; j < 5; ++j) m(i, 6 + j) = b[i].steps_ms[j];
        }
        return out;
    }
    void reset_stats() { check(ceos_server_reset_stats(live())); }

   private:
    ceos_server* s_ = nullptr;
    int d_, k_max_;
};

PYBIND11_MODULE(CEOs, m) {
    m.doc() = "MI355X-native coCEOs (HIP, gfx950) with the reference CEOs module's Python API";

    py::class_<PyCEOs>(m, "CEOs")
        .def(py::init<int, int>(), py::arg("n_points"), py::arg("n_features"),
             "Constructor for CEOs.\nn_points: number of points.\nn_features: number of dimensions.")
        .def("setIndexParam", &PyCEOs::setIndexParam, py::arg("n_proj"), py::arg("n_repeats") = 1,
             py::arg("top_m") = 10, py::arg("n_threads") = -1, py::arg("random_seed") = -1)
        .def_property("n_probed_vectors", &PyCEOs::get_s, &PyCEOs::set_s)
        .def_property("n_probed_points", &PyCEOs::get_mp, &PyCEOs::set_mp)
        .def_property("n_cand", &PyCEOs::get_c, &PyCEOs::set_c)
        .def_property("centering", &PyCEOs::get_centering, &PyCEOs::set_centering)
        .def_property_readonly("center", &PyCEOs::center)
        .def_property("rerank_bounds", &PyCEOs::get_rerank_bounds, &PyCEOs::set_rerank_bounds)
        .def("set_threads", &PyCEOs::set_threads, py::arg("n_threads"))
        .def("clear", &PyCEOs::clear)
        .def("build", &PyCEOs::build, py::arg("dataset"))
        .def("search", &PyCEOs::search, py::arg("queries"), py::arg("n_neighbors"), py::arg("verbose") = false)
        .def("build_coCEOs_Est", &PyCEOs::build_coCEOs_Est, py::arg("dataset"))
        .def("search_coCEOs_Est", &PyCEOs::search_coCEOs_Est, py::arg("queries"), py::arg("n_neighbors"),
             py::arg("verbose") = false)
        .def("build_CEOs_Hash", &PyCEOs::build_CEOs_Hash, py::arg("dataset"))
        .def("search_CEOs_Hash", &PyCEOs::search_CEOs_Hash, py::arg("queries"), py::arg("n_neighbors"),
             py::arg("verbose") = false)
        .def("save", &PyCEOs::save, py::arg("path"))
        .def("load", &PyCEOs::load, py::arg("path"))
        .def("make_plan", &PyCEOs::make_plan, py::arg("nq_max"), py::arg("k_max"), py::arg("s_max"),
             py::arg("c_max"), py::arg("autotune") = true, py::keep_alive<0, 1>())
        .def("search_batched", &PyCEOs::search_batched, py::arg("queries"), py::arg("n_neighbors"),
             py::arg("n_probed_vectors"), py::arg("n_cand"))
        .def("get_buckets", &PyCEOs::get_buckets)
        .def("get_sign_bits", &PyCEOs::get_sign_bits)
        .def("project", &PyCEOs::project, py::arg("rows"))
        .def_property_readonly("fhtDim", &PyCEOs::fht_dim)
        .def("set_search_config", &PyCEOs::set_search_config, py::arg("wg_per_cu") = 0, py::arg("table_log2") = 0)
        .def("get_search_config", &PyCEOs::get_search_config)
        .def_property_readonly("_handle", &PyCEOs::handle);

    py::class_<PyPlan>(m, "SearchPlan")
        .def("search", &PyPlan::search, py::arg("queries"), py::arg("n_neighbors"), py::arg("n_probed_vectors"),
             py::arg("n_cand"))
        .def("run", &PyPlan::run, py::arg("nq"), py::arg("stream") = 0)
        .def("buffers", &PyPlan::buffers)
        .def("config", &PyPlan::config)
        .def("proj_waves", &PyPlan::proj_waves)
        .def("tuning", &PyPlan::tuning);

    py::class_<PyServer>(m, "Server")
        .def(py::init([](PyCEOs& ix, int max_batch, int k_max, int s_max, int c_max, int inflight, int64_t hbm_budget,
                         int64_t max_pending) {
                 return new PyServer(ix.raw(), nullptr, ix.dim(), max_batch, k_max, s_max, c_max, inflight,
                                     hbm_budget, max_pending);
             }),
             py::arg("index"), py::arg("max_batch"), py::arg("k_max"), py::arg("s_max"), py::arg("c_max"),
             py::arg("inflight") = 2, py::arg("hbm_budget") = 0, py::arg("max_pending") = 0, py::keep_alive<1, 2>())
        .def(py::init([](PyStream& ix, int max_batch, int k_max, int s_max, int c_max, int inflight, int64_t hbm_budget,
                         int64_t max_pending) {
                 return new PyServer(nullptr, ix.raw(), ix.dim(), max_batch, k_max, s_max, c_max, inflight,
                                     hbm_budget, max_pending);
             }),
             py::arg("index"), py::arg("max_batch"), py::arg("k_max"), py::arg("s_max"), py::arg("c_max"),
             py::arg("inflight") = 2, py::arg("hbm_budget") = 0, py::arg("max_pending") = 0, py::keep_alive<1, 2>())
        .def("submit", &PyServer::submit, py::arg("queries"), py::arg("n_neighbors"), py::arg("n_probed_vectors"),
             py::arg("n_cand"), py::arg("arrival_s") = py::none())
        .def("result", &PyServer::result, py::arg("ticket"), py::arg("wait") = true)
        .def("update", &PyServer::update, py::arg("new_dataset"), py::arg("n_delPoints") = 0)
        .def("delete_ids", &PyServer::remove, py::arg("ids"))
        .def("drain", &PyServer::drain)
        .def("stats", &PyServer::stats)
        .def("batches", &PyServer::batches)
        .def("reset_stats", &PyServer::reset_stats)
        .def("close", &PyServer::close);

    py::class_<PyStream>(m, "streamCEOs")
        .def(py::init<int>(), py::arg("n_features"))
        .def("setIndexParam", &PyStream::setIndexParam, py::arg("n_proj"), py::arg("n_repeats") = 1,
             py::arg("top_points") = 1, py::arg("n_threads") = -1, py::arg("random_seed") = -1)
        .def_property("n_probed_vectors", &PyStream::get_s, &PyStream::set_s)
        .def_property("n_probed_points", &PyStream::get_mp, &PyStream::set_mp)
        .def_property("n_cand", &PyStream::get_c, &PyStream::set_c)
        .def_property("centering", &PyStream::get_centering, &PyStream::set_centering)
        .def_property_readonly("center", &PyStream::center)
        .def_property("rerank_bounds", &PyStream::get_rerank_bounds, &PyStream::set_rerank_bounds)
        .def("set_threads", &PyStream::set_threads, py::arg("n_threads"))
        .def("clear", &PyStream::clear)
        .def("build", &PyStream::build, py::arg("dataset"))
        .def("save", &PyStream::save, py::arg("path"))
        .def("load", &PyStream::load, py::arg("path"))
        .def("make_plan", &PyStream::make_plan, py::arg("nq_max"), py::arg("k_max"), py::arg("s_max"),
             py::arg("c_max"), py::arg("autotune") = true, py::keep_alive<0, 1>())
        .def("update", &PyStream::update, py::arg("new_dataset"), py::arg("n_delPoints") = 0)
        .def("estimate_search", &PyStream::estimate_search, py::arg("queries"), py::arg("n_neighbors"),
             py::arg("verbose") = false)
        .def("hash_search", &PyStream::hash_search, py::arg("queries"), py::arg("n_neighbors"),
             py::arg("verbose") = false)
        .def("delete_ids", &PyStream::delete_ids, py::arg("ids"))
        .def("size", &PyStream::size)
        .def("get_buckets", &PyStream::get_buckets, py::arg("n_lists"), py::arg("top_m"));
}
