// Host runtime of libceos_amd: device-resident coCEOs / CEOs-Est / streamCEOs indexes
// behind the C ABI declared in include/ceos_amd.h.
//
// Each index owns one HIP stream, its points (HBM copy of X), the sign diagonals,
// the bucket lists and the query scratch. Calls on one index are serialised by a
// mutex; different indexes run independently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ceos_amd.h"
#include "ceos_kernels.h"

using ceos::Entry;
using ceos::IndexDims;

namespace {

thread_local std::string g_err;

struct ArgError : std::runtime_error { using std::runtime_error::runtime_error; };
struct DevError : std::runtime_error { using std::runtime_error::runtime_error; };
struct MemError : std::runtime_error { using std::runtime_error::runtime_error; };
struct StateError : std::runtime_error { using std::runtime_error::runtime_error; };

#define HIP_CHECK(expr)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw DevError(std::string(#expr) + " failed: " + hipGetErrorString(e_));           \
    } while (0)

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_err.clear();
        return CEOS_OK;
    } catch (const ArgError& e) {
        g_err = e.what();
        return CEOS_EINVAL;
    } catch (const MemError& e) {
        g_err = e.what();
        return CEOS_ENOMEM;
    } catch (const StateError& e) {
        g_err = e.what();
        return CEOS_ESTATE;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CEOS_EDEVICE;
    }
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        if (hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw MemError("device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// src/CEOs.h:76-79
int fht_dim(int n_proj, int d) {
    if (n_proj < d) return 1 << (int)std::ceil(std::log2((double)d));
    return 1 << (int)std::ceil(std::log2((double)n_proj));
}

// Sign bits of src/Utilities.cpp:21-41: std::default_random_engine (minstd_rand0,
// x <- 16807 x mod (2^31-1)) drawn through libstdc++'s uniform_int_distribution(0,1),
// which rejects x-1 >= 2147483644 and returns (x-1) / 1073741822.
std::vector<uint8_t> sign_bits(int nbits, unsigned seed) {
    uint64_t x = seed % 2147483647u;
    if (x == 0) x = 1;
    std::vector<uint8_t> bits(nbits);
    for (int i = 0; i < nbits; ++i) {
        uint64_t r;
        do {
            x = (x * 16807u) % 2147483647u;
            r = x - 1;
        } while (r >= 2147483644u);
        bits[i] = (uint8_t)(r / 1073741822u);
    }
    return bits;
}

int device_cus() {
    static int cus = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        return v;
    }();
    if (cus <= 0) throw DevError("no HIP device available");
    return cus;
}

constexpr int kMaxLdsBytes = 160 * 1024 - 2048;  // minus the kernels' static LDS
constexpr int kMaxTopM = 8192;  // bucket depth; searches read at most kMaxProbedPoints of it
constexpr int kMaxProbes = 4096;
constexpr int kMaxK = 4096;

// Device state shared by CEOs and streamCEOs.
// Device scratch of the coCEOs search for one launch configuration.
struct ScratchSet {
    DBuf<uint64_t> cand;
    DBuf<float> qproj;
    DBuf<uint2> pairs;
    DBuf<uint4> stash;
    DBuf<uint32_t> tkeys;
    DBuf<int32_t> probes;
    DBuf<int32_t> work, status;

    ceos::QueryScratch ensure(const ceos::QueryConfig& cfg, int nq, int DR) {
        cand.ensure((size_t)cfg.grid * cfg.cand_stride);
        pairs.ensure((size_t)cfg.grid * cfg.cand_stride);
        stash.ensure((size_t)cfg.grid * ceos::kSearchStashList);
        tkeys.ensure((size_t)cfg.grid << cfg.table_log2);
        qproj.ensure((size_t)std::max(nq, 1) * DR);
        probes.ensure((size_t)std::max(nq, 1) * cfg.s_max);
        if (!work.p) work.alloc(1);
        if (!status.p) {
            status.alloc(1);
            HIP_CHECK(hipMemset(status.p, 0, 4));
        }
        return ceos::QueryScratch{qproj.p, probes.p, cand.p, pairs.p, stash.p, tkeys.p, work.p, status.p};
    }
    void release() {
        cand.release(); qproj.release(); pairs.release(); stash.release(); tkeys.release(); probes.release();
    }
};

struct Core {
    int d = 0;
    int D = 256, R = 1, m = 100, F = 0, L = 0;
    int seed = -1;
    unsigned seed_used = 0;
    int threads = 1;
    int s = 10, mp = 10, c = 10;  // n_probed_vectors, n_probed_points, n_cand
    bool centering = false;
    bool params_set = false;
    bool built = false;
    hipStream_t st_ = nullptr;

    std::vector<uint8_t> bits;
    DBuf<uint32_t> signs;
    DBuf<Entry> bkt;
    DBuf<int32_t> bsize;
    // stored points: ring of `cap` rows, logical id 0 at slot `head`, absolute id `base`
    DBuf<float> X;
    int64_t cap = 0, head = 0;
    uint32_t base = 0, live = 0;
    DBuf<uint32_t> dead;  // stream random-delete bitmap over ring slots

    // query scratch
    ScratchSet scr;
    int wg_per_cu = 0, table_log2 = 0;
    int last_lds = 0, last_tlog2 = 0, last_wg = 0;
    // optional per-launch kernel timing (HIP events on the launch stream)
    bool ktiming = false;
    std::vector<std::array<hipEvent_t, 3>> kev;
    size_t kev_used = 0;

    // CEOs-Est projections [DR][n]
    DBuf<float> estP;
    int64_t est_n = 0;

    std::mutex mu;

    Core() = default;
    ~Core() {
        for (auto& e : kev)
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        if (st_) (void)hipStreamDestroy(st_);
    }
    static constexpr size_t kMaxTimedLaunches = 4096;
    // Attaches the next event triple to a search launch when timing is on.
    void arm_timing(ceos::QueryConfig& cfg) {
        cfg.ev[0] = cfg.ev[1] = cfg.ev[2] = nullptr;
        if (!ktiming || kev_used >= kMaxTimedLaunches) return;
        if (kev_used == kev.size()) {
            std::array<hipEvent_t, 3> e{};
            for (auto& x : e) HIP_CHECK(hipEventCreate(&x));
            kev.push_back(e);
        }
        for (int i = 0; i < 3; ++i) cfg.ev[i] = kev[kev_used][i];
        ++kev_used;
    }
    // The stream is created on first device use, so handles can be made and
    // configured on a host without a GPU.
    hipStream_t st() {
        if (!st_) HIP_CHECK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        return st_;
    }

    int DR() const { return D * R; }
    IndexDims dims() const { return IndexDims{d, L, D, R, m}; }
    hipStream_t stream(void* s_) { return s_ ? reinterpret_cast<hipStream_t>(s_) : st(); }

    void set_index_param(int n_proj, int n_repeats, int top_m, int n_threads, int seed_) {
        if (n_proj <= 0 || n_repeats <= 0 || top_m <= 0)
            throw ArgError("n_proj, n_repeats and top_m must be positive");
        const int f = fht_dim(n_proj, d);
        if (f > 4096) throw ArgError("fhtDim " + std::to_string(f) + " exceeds 4096 (n_proj and n_features <= 4096)");
        if (top_m > kMaxTopM) throw ArgError("top_m above " + std::to_string(kMaxTopM) + " is not supported");
        D = n_proj; R = n_repeats; m = top_m; seed = seed_;
        threads = n_threads;
        F = f;
        L = 0;
        while ((1 << L) < F) ++L;
        params_set = true;
        built = false;
    }

    void clear() {
        bkt.release(); bsize.release(); X.release(); signs.release(); dead.release();
        scr.release(); estP.release();
        bits.clear();
        cap = head = 0; base = live = 0; est_n = 0;
        built = false;
    }

    void make_signs() {
        seed_used = seed >= 0 ? (unsigned)seed
                              : (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
        bits = sign_bits(F * 3 * R, seed_used);
        const int W = (F + 31) / 32;
        std::vector<uint32_t> words((size_t)W * 3 * R, 0u);
        for (int r = 0; r < R; ++r)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < F; ++j)
                    if (bits[(size_t)r * 3 * F + (size_t)i * F + j])
                        words[(size_t)(r * 3 + i) * W + j / 32] |= 1u << (j % 32);
        signs.alloc(words.size());
        HIP_CHECK(hipMemcpyAsync(signs.p, words.data(), words.size() * 4, hipMemcpyHostToDevice, st()));
        HIP_CHECK(hipStreamSynchronize(st()));
    }

    // Copies n rows (host or device pointer) into a fresh ring starting at slot 0.
    void load_points(const float* src, int64_t n, int64_t capacity, bool src_device, hipStream_t s_) {
        X.alloc((size_t)capacity * d);
        cap = capacity; head = 0; base = 0; live = (uint32_t)n;
        HIP_CHECK(hipMemcpyAsync(X.p, src, (size_t)n * d * 4,
                                 src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s_));
    }

    // coCEOs buckets over rows 0..n-1 of X (ring slots 0..n-1): see ceos_build.hip.
    void build_buckets(int64_t n, hipStream_t s_) {
        const int DR_ = DR();
        const int nl = 2 * DR_;
        const IndexDims dm = dims();
        if (m > n) throw ArgError("top_m must not exceed the number of points");

        int64_t S_target = std::max<int64_t>(65536, (64 * n + m - 1) / m);
        S_target = std::min<int64_t>(S_target, (int64_t(4) << 30) / (4LL * DR_));
        S_target = std::max<int64_t>(S_target, 1);
        const int64_t stride = std::max<int64_t>(1, n / S_target);
        const int64_t S = (n + stride - 1) / stride;
        const bool exact = stride == 1;
        const int rank = exact ? m : (int)std::min<int64_t>(S, (2 * (int64_t)m * S + n - 1) / n + 8);

        DBuf<float> thr(nl);
        {
            DBuf<float> sample((size_t)DR_ * S);
            HIP_CHECK(ceos::launch_project_rows(dm, X.p, 0, stride, S, signs.p, sample.p, S, true, s_));
            HIP_CHECK(ceos::launch_sample_thresholds(sample.p, S, DR_, rank, thr.p, s_));
            HIP_CHECK(hipStreamSynchronize(s_));
        }
        std::vector<int64_t> hbase(nl);
        std::vector<int32_t> hcap(nl, (int32_t)std::min<int64_t>(n, 4LL * m + 256));
        std::vector<int32_t> hcount(nl);
        std::vector<float> hthr;
        DBuf<int64_t> dbase(nl);
        DBuf<int32_t> dcap(nl), dcount(nl);
        DBuf<Entry> ent;
        for (int iter = 0;; ++iter) {
            int64_t total = 0;
            for (int l = 0; l < nl; ++l) { hbase[l] = total; total += hcap[l]; }
            ent.alloc((size_t)total);
            HIP_CHECK(hipMemcpyAsync(dbase.p, hbase.data(), nl * 8, hipMemcpyHostToDevice, s_));
            HIP_CHECK(hipMemcpyAsync(dcap.p, hcap.data(), nl * 4, hipMemcpyHostToDevice, s_));
            HIP_CHECK(ceos::launch_fill_i32(dcount.p, 0, nl, s_));
            HIP_CHECK(ceos::launch_scan_append(dm, X.p, n, signs.p, thr.p, nullptr, dbase.p, dcap.p, dcount.p,
                                               ent.p, s_));
            HIP_CHECK(hipMemcpyAsync(hcount.data(), dcount.p, nl * 4, hipMemcpyDeviceToHost, s_));
            HIP_CHECK(hipStreamSynchronize(s_));
            bool redo = false;
            for (int l = 0; l < nl; ++l) {
                if (hcount[l] < m) {
                    if (hthr.empty()) {
                        hthr.resize(nl);
                        HIP_CHECK(hipMemcpy(hthr.data(), thr.p, nl * 4, hipMemcpyDeviceToHost));
                    }
                    hthr[l] = -INFINITY;  // accept every point for this list
                    hcap[l] = (int32_t)n;
                    redo = true;
                } else if (hcount[l] > hcap[l]) {
                    hcap[l] = hcount[l];
                    redo = true;
                }
            }
            if (!redo) break;
            if (iter >= 2) throw DevError("bucket construction did not converge");
            if (!hthr.empty()) HIP_CHECK(hipMemcpy(thr.p, hthr.data(), nl * 4, hipMemcpyHostToDevice));
        }
        bkt.alloc((size_t)nl * m);
        bsize.alloc(nl);
        HIP_CHECK(ceos::launch_list_select(nl, m, dbase.p, dcap.p, dcount.p, ent.p, bkt.p, bsize.p, s_));
        HIP_CHECK(hipStreamSynchronize(s_));
    }

    // Picks the LDS score-tile size and grid for a search with the given bounds:
    // the largest tile that fits the per-workgroup LDS share (wg_per_cu workgroups
    // of 512 threads per CU; default 1, i.e. a 32K-slot tile of 4-byte ids).
    ceos::QueryConfig search_config(int s_max, int c_max, int k_max) {
        ceos::QueryConfig cfg = search_config_with(s_max, c_max, k_max, wg_per_cu, table_log2);
        arm_timing(cfg);
        return cfg;
    }
    // The same for an explicit launch shape (0 = automatic), without kernel timing.
    ceos::QueryConfig search_config_with(int s_max, int c_max, int k_max, int wg_req, int tl_req) {
        const int wg = wg_req > 0 ? wg_req : 1;
        const int budget = std::min(kMaxLdsBytes, (160 * 1024) / wg - 2048);
        int t = 0;
        if (tl_req > 0) {
            t = tl_req;
        } else {
            for (int cand_t = 15; cand_t >= 8; --cand_t)
                if (ceos::search_lds_bytes(d, cand_t, s_max, c_max, k_max) <= budget) { t = cand_t; break; }
            if (t == 0) t = 8;
        }
        const int lds = ceos::search_lds_bytes(d, t, s_max, c_max, k_max);
        if (lds > kMaxLdsBytes)
            throw ArgError("search parameters need " + std::to_string(lds) +
                           " bytes of LDS per workgroup (n_cand or n_probed_vectors too large)");
        if (std::min(mp, m) > ceos::kMaxProbedPoints)
            throw ArgError("n_probed_points above " + std::to_string(ceos::kMaxProbedPoints) + " is not supported");
        if ((int64_t)s_max * std::min(mp, m) > (int64_t(1) << 26))
            throw ArgError("n_probed_vectors * n_probed_points above 2^26 is not supported");
        ceos::QueryConfig cfg{};
        cfg.table_log2 = t;
        cfg.lds_bytes = lds;
        const int fit = std::max(1, (160 * 1024) / (lds + 2048));
        cfg.grid = device_cus() * std::min(wg, fit);
        cfg.s_max = s_max; cfg.c_max = c_max; cfg.k_max = k_max;
        cfg.cand_stride = (int64_t)s_max * std::max(1, std::min(mp, m));
        last_lds = lds; last_tlog2 = t; last_wg = std::min(wg, fit);
        return cfg;
    }

    ceos::QueryScratch scratch(const ceos::QueryConfig& cfg, int nq) { return scr.ensure(cfg, nq, DR()); }

    ceos::RowRing rows() const { return ceos::RowRing{X.p, cap, head, base, live}; }

    void check_status() {
        int32_t flag = 0;
        HIP_CHECK(hipMemcpy(&flag, scr.status.p, 4, hipMemcpyDeviceToHost));
        if (flag) {
            HIP_CHECK(hipMemset(scr.status.p, 0, 4));
            throw DevError("LDS score tile overflowed; rerun with a larger table (ceos_set_search_config)");
        }
    }

    void validate_query(int s_, int mp_, int c_, int k, int s_limit, const char* s_limit_name) const {
        if (!built) throw StateError("index is not built");
        if (mp_ > m) throw ArgError("Number of probed points must be smaller than number of indexed top-m points");
        if (s_ > s_limit)
            throw ArgError(std::string("Number of probed vectors must be smaller than ") + s_limit_name);
        if (s_ <= 0 || mp_ <= 0 || c_ <= 0) throw ArgError("n_probed_vectors, n_probed_points and n_cand must be positive");
        if (k < 0 || k > kMaxK) throw ArgError("n_neighbors must be in [0, " + std::to_string(kMaxK) + "]");
        if (s_ > kMaxProbes) throw ArgError("n_probed_vectors above " + std::to_string(kMaxProbes) + " is not supported");
    }

    // Uniform-parameter coCEOs search on device pointers.
    void search_device(const float* dQ, int nq, int k, int s_, int c_, int32_t* ids, float* ips, hipStream_t s__) {
        if (nq <= 0 || k == 0) return;
        const ceos::QueryConfig cfg = search_config(s_, c_, k);
        const ceos::QueryScratch scr = scratch(cfg, nq);
        ceos::QueryBatch qb{dQ, nq, nullptr, nullptr, nullptr, k, s_, c_, mp, nullptr, ids, ips};
        HIP_CHECK(ceos::launch_coceos_search(dims(), signs.p, bkt.p, bsize.p, rows(), qb, cfg, scr, s__));
    }

    void search_host(const float* Q, int nq, int k, int s_, int c_, int32_t* ids, float* ips) {
        if (nq <= 0 || k == 0) return;
        DBuf<float> dQ((size_t)nq * d);
        DBuf<int32_t> dI((size_t)nq * k);
        DBuf<float> dP((size_t)nq * k);
        HIP_CHECK(hipMemcpyAsync(dQ.p, Q, (size_t)nq * d * 4, hipMemcpyHostToDevice, st()));
        search_device(dQ.p, nq, k, s_, c_, dI.p, dP.p, st());
        HIP_CHECK(hipMemcpyAsync(ids, dI.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st()));
        HIP_CHECK(hipMemcpyAsync(ips, dP.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st()));
        HIP_CHECK(hipStreamSynchronize(st()));
        check_status();
    }

    void get_buckets(int32_t* cid, float* cval, int32_t* fid, float* fval, int32_t* csz, int32_t* fsz) {
        if (!built) throw StateError("index is not built");
        const int DR_ = DR();
        std::vector<Entry> h((size_t)2 * DR_ * m);
        std::vector<int32_t> sz(2 * DR_);
        HIP_CHECK(hipMemcpy(h.data(), bkt.p, h.size() * sizeof(Entry), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(sz.data(), bsize.p, sz.size() * 4, hipMemcpyDeviceToHost));
        for (int col = 0; col < DR_; ++col) {
            for (int side = 0; side < 2; ++side) {
                const int l = 2 * col + side;
                int32_t* oid = side ? fid : cid;
                float* ov = side ? fval : cval;
                for (int t = 0; t < m; ++t) {
                    const Entry e = h[(size_t)l * m + t];
                    const bool valid = t < sz[l];
                    oid[(size_t)col * m + t] = valid ? (int32_t)(e.x - base) : 0;
                    uint32_t bits_ = valid ? e.y : 0u;
                    std::memcpy(&ov[(size_t)col * m + t], &bits_, 4);
                }
                if (side == 0 && csz) csz[col] = sz[l];
                if (side == 1 && fsz) fsz[col] = sz[l];
            }
        }
    }
};

void check_matrix(const void* p, int64_t n, int d, int expect_d) {
    if (!p && n > 0) throw ArgError("null data pointer");
    if (n < 0) throw ArgError("negative row count");
    if (d != expect_d)
        throw ArgError("feature dimension " + std::to_string(d) + " does not match the index (" +
                       std::to_string(expect_d) + ")");
}

}  // namespace

struct ceos_index {
    Core core;
    int n_points = 0;
};

struct ceos_stream {
    Core core;
};

// A batched search captured in a hipGraph over plan-owned buffers (see ceos_plan_create).
struct ceos_plan {
    Core* core = nullptr;
    bool stream_index = false;  // streamCEOs: n_probed_vectors is bounded by n_proj
    int nq_max = 0, k_max = 0, s_max = 0, c_max = 0, flags = 0;
    DBuf<float> Q, ips;
    DBuf<int32_t> k, s, c, ids, nq_dev;
    ScratchSet scr;
    ceos::QueryConfig cfg{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    // everything the captured launches hold by value
    struct Snap {
        const void *signs, *bkt, *bsize, *X;
        int64_t cap, head;
        uint32_t base, live;
        int mp, m, D, R, L, d;
        bool operator==(const Snap& o) const {
            return signs == o.signs && bkt == o.bkt && bsize == o.bsize && X == o.X && cap == o.cap &&
                   head == o.head && base == o.base && live == o.live && mp == o.mp && m == o.m && D == o.D &&
                   R == o.R && L == o.L && d == o.d;
        }
    } snap{};
    bool tuned = false;
    double tuned_ms = 0.0;
    int wg = 0, tl = 0;

    ~ceos_plan() {
        drop();
        if (cap) (void)hipStreamDestroy(cap);
    }
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    Snap current() const {
        const Core& x = *core;
        return Snap{x.signs.p, x.bkt.p, x.bsize.p, x.X.p, x.cap, x.head, x.base, x.live, x.mp, x.m, x.D, x.R, x.L, x.d};
    }
    void capture(int wg_, int tl_) {
        drop();
        Core& x = *core;
        cfg = x.search_config_with(s_max, c_max, k_max, wg_, tl_);
        const ceos::QueryScratch sc = scr.ensure(cfg, nq_max, x.DR());
        // k0 = k_max: rows at q * k_max, zero padded to k_max
        ceos::QueryBatch qb{Q.p, nq_max, k.p, s.p, c.p, k_max, 0, 0, x.mp, nullptr, ids.p, ips.p, nq_dev.p};
        if (!cap) HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        const hipError_t e1 =
            ceos::launch_coceos_search(x.dims(), x.signs.p, x.bkt.p, x.bsize.p, x.rows(), qb, cfg, sc, cap);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(cap, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            throw DevError(std::string("search plan capture failed: ") +
                           hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        }
        graph = g;
        HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        snap = current();
        wg = wg_;
        tl = tl_;
    }
    void launch(int nq, hipStream_t st) {
        HIP_CHECK(ceos::launch_fill_i32(nq_dev.p, nq, 1, st));  // a kernel, outside the graph
        HIP_CHECK(hipGraphLaunch(exec, st));
    }
    // Times each candidate launch shape once on this batch and keeps the fastest. Every
    // candidate computes the same results, so the batch's outputs are valid afterwards.
    void tune(int nq, hipStream_t st) {
        Core& x = *core;
        const int t1 = x.search_config_with(s_max, c_max, k_max, 1, 0).table_log2;
        const int t2 = x.search_config_with(s_max, c_max, k_max, 2, 0).table_log2;
        std::vector<std::pair<int, int>> cands = {{1, 0}, {2, 0}};
        if (t1 - 1 >= 8) cands.push_back({1, t1 - 1});
        if (t2 - 1 >= 8) cands.push_back({2, t2 - 1});
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        double best = 1e30;
        std::pair<int, int> pick = cands[0];
        try {
            for (const auto& cd : cands) {
                capture(cd.first, cd.second);
                launch(nq, st);  // warm
                HIP_CHECK(hipEventRecord(e0, st));
                launch(nq, st);
                HIP_CHECK(hipEventRecord(e1, st));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    pick = cd;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        capture(pick.first, pick.second);
        tuned = true;
        tuned_ms = best;
    }
    void validate() const {
        const Core& x = *core;
        x.validate_query(s_max, x.mp, c_max, k_max, stream_index ? x.D : x.DR(),
                         stream_index ? "number of projections" : "n_proj * n_repeats");
    }
    void run(int nq, hipStream_t st) {
        if (nq < 0 || nq > nq_max) throw ArgError("batch of " + std::to_string(nq) + " exceeds the plan's nq_max");
        validate();
        if (nq == 0) return;
        if (!exec || !(snap == current())) capture(wg, tl);
        if ((flags & CEOS_PLAN_AUTOTUNE) && !tuned)
            tune(nq, st);
        else
            launch(nq, st);
    }
};

namespace {
int make_plan(Core& x, bool stream_index, int nq_max, int k_max, int s_max, int c_max, int flags, ceos_plan** out) {
    return guarded([&] {
        if (!out) throw ArgError("null output handle");
        if (nq_max <= 0 || k_max <= 0 || s_max <= 0 || c_max <= 0)
            throw ArgError("nq_max, k_max, s_max and c_max must be positive");
        std::lock_guard<std::mutex> g(x.mu);
        auto* p = new ceos_plan();
        try {
            p->core = &x;
            p->stream_index = stream_index;
            p->nq_max = nq_max; p->k_max = k_max; p->s_max = s_max; p->c_max = c_max; p->flags = flags;
            if (x.built) p->validate();
            p->Q.alloc((size_t)nq_max * x.d);
            p->k.alloc(nq_max); p->s.alloc(nq_max); p->c.alloc(nq_max);
            p->ids.alloc((size_t)nq_max * k_max);
            p->ips.alloc((size_t)nq_max * k_max);
            p->nq_dev.alloc(1);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}
}  // namespace

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

static void build_coceos_impl(ceos_index* h, const float* X, int64_t n, int d, bool dev, hipStream_t st) {
    Core& c = h->core;
    if (!c.params_set) throw StateError("setIndexParam must be called before build");
    check_matrix(X, n, d, c.d);
    if (n <= 0) throw ArgError("empty dataset");
    if (n >= (int64_t(1) << 31)) throw ArgError("more than 2^31-1 points");
    c.built = false;
    h->n_points = (int)n;
    c.estP.release();
    c.load_points(X, n, n, dev, st);
    c.make_signs();
    c.build_buckets(n, st);
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
        const ceos::QueryScratch scr = c.scratch(cfg, nq);
        ceos::QueryBatch qb{dQ, nq, d_k, d_s, d_c, 0, 0, 0, c.mp, d_off, d_ids, d_ips};
        HIP_CHECK(ceos::launch_coceos_search(c.dims(), c.signs.p, c.bkt.p, c.bsize.p, c.rows(), qb, cfg, scr,
                                             c.stream(stream)));
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
        const ceos::QueryScratch scr = c.scratch(cfg, nq);
        ceos::QueryBatch qb{dQ.p, nq, dk.p, ds.p, dc.p, 0, 0, 0, c.mp, doff.p, dI.p, dP.p};
        HIP_CHECK(ceos::launch_coceos_search(c.dims(), c.signs.p, c.bkt.p, c.bsize.p, c.rows(), qb, cfg, scr, c.st()));
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
        c.built = false;
        h->n_points = (int)n;
        c.bkt.release();
        c.bsize.release();
        c.load_points(X, n, n, false, c.st());
        c.make_signs();
        c.estP.alloc((size_t)c.DR() * n);
        c.est_n = n;
        HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p, 0, 1, n, c.signs.p, c.estP.p, n, true, c.st()));
        HIP_CHECK(hipStreamSynchronize(c.st()));
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
        const int cc = (int)std::min<int64_t>(c.c, c.est_n);
        const int lds = ceos::est_search_lds_bytes(c.dims(), c.s, cc, k);
        if (lds > kMaxLdsBytes) throw ArgError("n_cand too large for the device path");
        const int64_t stride = c.est_n + cc;
        int grid = device_cus() * 2;
        grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t(2) << 30) / (stride * 8)));
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

int ceos_get_search_config(const ceos_index* h, int* wg_per_cu, int* table_log2, int* lds_bytes) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        if (wg_per_cu) *wg_per_cu = h->core.last_wg;
        if (table_log2) *table_log2 = h->core.last_tlog2;
        if (lds_bytes) *lds_bytes = h->core.last_lds;
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

int ceos_stream_set_threads(ceos_stream* h, int n_threads) {
    return guarded([&] { if (!h) throw ArgError("null handle"); h->core.threads = n_threads; });
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
        c.built = false;
        const int64_t capacity = std::max<int64_t>(n + n / 4, 1024);
        c.load_points(X, n, capacity, false, c.st());
        c.dead.alloc((size_t)(capacity + 31) / 32);
        HIP_CHECK(hipMemsetAsync(c.dead.p, 0, c.dead.n * 4, c.st()));
        c.make_signs();
        c.build_buckets(n, c.st());
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
        if (n_new > 0) {
            HIP_CHECK(hipMemcpyAsync(c.X.p + slot0 * c.d, Xn, (size_t)seg1 * c.d * 4, hipMemcpyHostToDevice, st));
            if (n_new > seg1)
                HIP_CHECK(hipMemcpyAsync(c.X.p, Xn + seg1 * c.d, (size_t)(n_new - seg1) * c.d * 4,
                                         hipMemcpyHostToDevice, st));
            std::vector<uint32_t> dh(c.dead.n);
            HIP_CHECK(hipMemcpyAsync(dh.data(), c.dead.p, dh.size() * 4, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            for (int64_t j = 0; j < n_new; ++j) {
                int64_t sl = slot0 + j;
                if (sl >= c.cap) sl -= c.cap;
                dh[sl >> 5] &= ~(1u << (sl & 31));
            }
            HIP_CHECK(hipMemcpyAsync(c.dead.p, dh.data(), dh.size() * 4, hipMemcpyHostToDevice, st));
        }
        // project the new rows: column-major [DR][n_new]
        DBuf<float> np((size_t)c.DR() * std::max<int64_t>(n_new, 1));
        if (n_new > 0) {
            HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p + slot0 * c.d, 0, 1, seg1, c.signs.p, np.p, n_new,
                                                true, st));
            if (n_new > seg1)
                HIP_CHECK(ceos::launch_project_rows(c.dims(), c.X.p, 0, 1, n_new - seg1, c.signs.p, np.p + seg1,
                                                    n_new, true, st));
        }
        const uint32_t new_base = c.base + (uint32_t)n_del;
        const uint32_t new_live = (uint32_t)(n_cur + n_new - n_del);
        HIP_CHECK(ceos::launch_stream_merge(c.DR(), c.m, c.bkt.p, c.bsize.p, np.p, n_new, first_abs, new_base,
                                            new_live, st));
        HIP_CHECK(hipStreamSynchronize(st));
        c.base = new_base;
        c.live = new_live;
        c.head = (c.head + n_del) % c.cap;
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
        std::vector<uint32_t> dh(c.dead.n);
        HIP_CHECK(hipMemcpy(dh.data(), c.dead.p, dh.size() * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < count; ++i) {
            if (ids[i] < 0 || (uint32_t)ids[i] >= c.live) throw ArgError("id " + std::to_string(ids[i]) + " is not live");
            int64_t sl = c.head + ids[i];
            if (sl >= c.cap) sl -= c.cap;
            dh[sl >> 5] |= 1u << (sl & 31);
        }
        HIP_CHECK(hipMemcpyAsync(c.dead.p, dh.data(), dh.size() * 4, hipMemcpyHostToDevice, c.st()));
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
        if (table_log2) *table_log2 = p->cfg.table_log2;
        if (tuned_ms) *tuned_ms = p->tuned ? p->tuned_ms : 0.0;
    });
}

}  // extern "C"
