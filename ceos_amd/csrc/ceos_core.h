// Internal host runtime of libceos_amd (not installed): device buffers, the index
// core shared by CEOs and streamCEOs handles, search scratch and captured search plans.
// Included by the C-ABI entry points (ceos_runtime.cpp) and the serving engine
// (ceos_serve.cpp).
//
// Each index owns one HIP stream, its points (HBM copy of X), the sign diagonals,
// the bucket lists and the query scratch. Calls on one index are serialised by a
// mutex; different indexes run independently.
#pragma once

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

namespace ceos_rt {

// Last error text of the calling thread (ceos_last_error).
inline thread_local std::string g_err;

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

// Pinned host buffer (device copies into it run without a staging round trip).
template <typename T>
struct HBuf {
    T* p = nullptr;
    size_t n = 0;
    HBuf() = default;
    HBuf(const HBuf&) = delete;
    HBuf& operator=(const HBuf&) = delete;
    ~HBuf() { release(); }
    void ensure(size_t count) {
        if (count <= n) return;
        release();
        if (hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw MemError("pinned host allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// Pinned host buffer mapped into the device's address space and coherent: a kernel
// reads what the host wrote before its launch (through `dp`), with no copy-engine
// transfer.
template <typename T>
struct MappedBuf {
    T* p = nullptr;
    T* dp = nullptr;
    size_t n = 0;
    MappedBuf() = default;
    MappedBuf(const MappedBuf&) = delete;
    MappedBuf& operator=(const MappedBuf&) = delete;
    ~MappedBuf() { release(); }
    void ensure(size_t count) {
        if (count <= n) return;
        release();
        size_t cap = 1024;
        while (cap < count) cap *= 2;
        if (hipHostMalloc(reinterpret_cast<void**>(&p), cap * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw MemError("pinned host allocation of " + std::to_string(cap * sizeof(T)) + " bytes failed");
        }
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess || !d) {
            (void)hipGetLastError();
            (void)hipHostFree(p);
            p = nullptr;
            throw MemError("pinned host memory is not mapped for the device");
        }
        dp = static_cast<T*>(d);
        n = cap;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = dp = nullptr;
        n = 0;
    }
};

// src/CEOs.h:76-79
inline int fht_dim(int n_proj, int d) {
    if (n_proj < d) return 1 << (int)std::ceil(std::log2((double)d));
    return 1 << (int)std::ceil(std::log2((double)n_proj));
}

// Sign bits of src/Utilities.cpp:21-41: std::default_random_engine (minstd_rand0,
// x <- 16807 x mod (2^31-1)) drawn through libstdc++'s uniform_int_distribution(0,1),
// which rejects x-1 >= 2147483644 and returns (x-1) / 1073741822.
inline std::vector<uint8_t> sign_bits(int nbits, unsigned seed) {
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

inline int device_cus() {
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
    DBuf<float> est;

    bool fits(const ceos::QueryConfig& cfg, int nq, int DR) const {
        return cand.n >= (size_t)cfg.grid * cfg.cand_stride && pairs.n >= (size_t)cfg.grid * cfg.cand_stride &&
               stash.n >= (size_t)cfg.grid * ceos::kSearchStashList &&
               tkeys.n >= (size_t)cfg.grid * ceos::search_tile_slots_nbk(cfg.tile_buckets) &&
               qproj.n >= (size_t)std::max(nq, 1) * DR && probes.n >= (size_t)std::max(nq, 1) * cfg.s_max && work.p &&
               status.p && est.p;
    }
    ceos::QueryScratch ensure(const ceos::QueryConfig& cfg, int nq, int DR) {
        cand.ensure((size_t)cfg.grid * cfg.cand_stride);
        pairs.ensure((size_t)cfg.grid * cfg.cand_stride);
        stash.ensure((size_t)cfg.grid * ceos::kSearchStashList);
        tkeys.ensure((size_t)cfg.grid * ceos::search_tile_slots_nbk(cfg.tile_buckets));
        qproj.ensure((size_t)std::max(nq, 1) * DR);
        probes.ensure((size_t)std::max(nq, 1) * cfg.s_max);
        if (!work.p) work.alloc(1);
        if (!status.p) {
            status.alloc(1);
            HIP_CHECK(hipMemset(status.p, 0, 4));
        }
        if (!est.p) {
            est.alloc(1);
            const float half = 0.5f;
            HIP_CHECK(hipMemcpy(est.p, &half, 4, hipMemcpyHostToDevice));
        }
        return ceos::QueryScratch{qproj.p, probes.p, cand.p, pairs.p, stash.p, tkeys.p, work.p, status.p, est.p};
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
    std::vector<float> center;  // streamCEOs with centering: column mean of the build rows
    DBuf<uint32_t> signs;
    DBuf<Entry> bkt;
    DBuf<int32_t> bsize;
    // stored points: ring of `cap` rows, logical id 0 at slot `head`, absolute id `base`
    DBuf<float> X;
    int64_t cap = 0, head = 0;
    uint32_t base = 0, live = 0;
    DBuf<uint32_t> dead;  // stream random-delete bitmap over ring slots
    MappedBuf<int32_t> del_ids;  // ids of a random delete, read by the marking kernel
    // stream update scratch (the new rows' projections, row-major then per list), kept
    // between updates up to kUpdScratchKeep floats each: a 10,000-row update at
    // n_proj*n_repeats 4096 needs 164 MB of each, and allocating and freeing them took
    // 0.4 of the update's 0.58 ms projection phase
    static constexpr size_t kUpdScratchKeep = size_t(64) << 20;
    DBuf<float> upd_rows, upd_lists;
    // rerank bounds: int8 row copies [cap][quant_stride(d)] holding each row's codes and
    // (s, r1, r2, 0), kept in step with X for coCEOs and stream indexes (see
    // k_quantize_rows); off: not kept, every candidate gets the exact rerank
    DBuf<int8_t> Xq;
    bool rerank_bounds = true;

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
    // build workspace, kept between builds (grow-only; released by clear()): sample
    // projections row-major and column-major, per-list thresholds, candidate segments
    DBuf<float> bw_rows, bw_sample, bw_thr;
    DBuf<int64_t> bw_base;
    DBuf<int32_t> bw_cap, bw_count;
    HBuf<int32_t> bw_hcount;
    DBuf<uint8_t> bw_active;
    // list rescans of a build (rare): entries per rescan range (at least n)
    int64_t rescan_budget_entries = int64_t(1) << 27;
    struct RescanStats {
        int under = 0, over = 0, ranges = 0;
    } rescan;
    DBuf<Entry> bw_ent;
    int64_t est_n = 0;

    std::mutex mu;

    // Cross-stream ordering. Device searches may run on caller streams and return before
    // they finish; they share this index's query scratch, and updates rewrite the lists
    // and rows every search reads. Each launch that reads the index records an event on
    // its stream. Work that rewrites or frees the index (builds, updates, deletes, loads,
    // clear, destruction) first waits on the host for every such event, and a launch
    // that uses the index's own scratch first makes its stream wait for the previous
    // scratch user when that ran on another stream.
    struct StreamEvent {
        hipStream_t s;
        hipEvent_t e;
        bool pending;
    };
    std::vector<StreamEvent> rd_ev;
    hipEvent_t scr_ev = nullptr;
    hipStream_t scr_stream = nullptr;
    bool scr_pending = false;

    void note_read(hipStream_t s_) {
        StreamEvent* se = nullptr;
        for (auto& x : rd_ev)
            if (x.s == s_) se = &x;
        if (!se) {
            if (rd_ev.size() >= 64) wait_reads();  // many caller streams: retire the old events
            if (rd_ev.size() >= 64) {
                for (auto& x : rd_ev) (void)hipEventDestroy(x.e);
                rd_ev.clear();
            }
            hipEvent_t e = nullptr;
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            rd_ev.push_back({s_, e, false});
            se = &rd_ev.back();
        }
        HIP_CHECK(hipEventRecord(se->e, s_));
        se->pending = true;
    }
    // Host wait for every launch that reads the index (before it is rewritten or freed).
    void wait_reads() {
        for (auto& x : rd_ev)
            if (x.pending) {
                HIP_CHECK(hipEventSynchronize(x.e));
                x.pending = false;
            }
        scr_pending = false;
    }
    // Orders a launch that uses the index's scratch after the previous one.
    void before_scratch(hipStream_t s_) {
        if (scr_pending && scr_stream != s_) HIP_CHECK(hipStreamWaitEvent(s_, scr_ev, 0));
    }
    void after_scratch(hipStream_t s_) {
        if (!scr_ev) HIP_CHECK(hipEventCreateWithFlags(&scr_ev, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(scr_ev, s_));
        scr_stream = s_;
        scr_pending = true;
        note_read(s_);
    }

    // Phase times of the last build / update (copy, projection and candidate scan, list
    // selection) and of the last host-entry search (query projection + probe selection,
    // fused search kernel), from HIP events on the stream the work ran on.
    hipEvent_t ph_ev[5] = {};
    ceos_phase_times last_build{}, last_search{};
    void phase_mark(int i, hipStream_t s_) {
        if (!ph_ev[i]) HIP_CHECK(hipEventCreate(&ph_ev[i]));
        HIP_CHECK(hipEventRecord(ph_ev[i], s_));
    }
    double phase_ms(int a, int b) {
        float ms = 0.0f;
        HIP_CHECK(hipEventSynchronize(ph_ev[b]));
        HIP_CHECK(hipEventElapsedTime(&ms, ph_ev[a], ph_ev[b]));
        return ms;
    }
    double index_gb() const {  // the reference's figure: 4 m x DR list floats plus the stored rows
        return (4.0 * m * DR() + (double)live * d) * 4.0 / (double)(1 << 30);
    }

    Core() = default;
    ~Core() {
        for (hipEvent_t e : ph_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : sev)
            if (e) (void)hipEventDestroy(e);
        for (auto& x : rd_ev) {
            if (x.pending) (void)hipEventSynchronize(x.e);
            (void)hipEventDestroy(x.e);
        }
        if (scr_ev) (void)hipEventDestroy(scr_ev);
        for (auto& e : kev)
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        if (st_) (void)hipStreamDestroy(st_);
        if (aux_) {
            (void)hipStreamSynchronize(aux_);
            (void)hipStreamDestroy(aux_);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
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
    // A second stream for build work that can overlap the bucket scan (default priority).
    hipStream_t aux_ = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool aux_pending = false;
    hipStream_t aux() {
        if (!aux_) HIP_CHECK(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
        if (!fork_ev) HIP_CHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        if (!join_ev) HIP_CHECK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
        return aux_;
    }
    // Makes s_ wait for the work put on the second stream (no host wait).
    void join_aux(hipStream_t s_) {
        if (!aux_pending) return;
        HIP_CHECK(hipStreamWaitEvent(s_, join_ev, 0));
        aux_pending = false;
    }
    // Host wait for the second stream (error paths: its work may read caller memory).
    void drain_aux() {
        if (aux_) (void)hipStreamSynchronize(aux_);
        aux_pending = false;
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
        wait_reads();
        dcenter.release();
        center.clear();
        bkt.release(); bsize.release(); X.release(); signs.release(); dead.release();
        upd_rows.release(); upd_lists.release();
        Xq.release();
        scr.release(); estP.release();
        bw_rows.release(); bw_sample.release(); bw_thr.release(); bw_base.release(); bw_cap.release();
        bw_count.release(); bw_ent.release(); bw_hcount.release(); bw_active.release();
        bits.clear();
        cap = head = 0; base = live = 0; est_n = 0;
        built = false;
    }

    void make_signs() {
        make_signs_from(seed >= 0 ? (unsigned)seed
                                  : (unsigned)std::chrono::system_clock::now().time_since_epoch().count());
    }
    void make_signs_from(unsigned seed_value) {
        seed_used = seed_value;
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
    // A ring of the same capacity is reused (a rebuild of the same shape allocates nothing);
    // the int8 copies are kept for quantize_rows to overwrite or resize.
    void load_points(const float* src, int64_t n, int64_t capacity, bool src_device, hipStream_t s_) {
        if (X.n != (size_t)capacity * d) X.alloc((size_t)capacity * d);
        cap = capacity; head = 0; base = 0; live = (uint32_t)n;
        HIP_CHECK(hipMemcpyAsync(X.p, src, (size_t)n * d * 4,
                                 src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s_));
    }

    // A ring for n rows that the first bucket scan fills (build_buckets with src).
    void reserve_points(int64_t n, int64_t capacity) {
        if (X.n != (size_t)capacity * d) X.alloc((size_t)capacity * d);
        cap = capacity; head = 0; base = 0; live = (uint32_t)n;
    }

    // Centering (src/streamCEOs.cpp:22-34): the column mean of rows 0..n-1 of X (ring slots
    // 0..n-1) is computed on the device, subtracted from those rows, kept on the device
    // for later inserts and copied to `center` on the host (index files, get_center).
    DBuf<float> dcenter;
    void center_rows(int64_t n, hipStream_t s_) {
        DBuf<double> partial((size_t)ceos::column_mean_blocks(n) * d);
        dcenter.alloc(d);
        HIP_CHECK(ceos::launch_column_mean(X.p, n, d, partial.p, dcenter.p, s_));
        HIP_CHECK(ceos::launch_subtract_center(X.p, 0, n, cap, d, dcenter.p, s_));
        center.resize(d);
        HIP_CHECK(hipMemcpyAsync(center.data(), dcenter.p, (size_t)d * 4, hipMemcpyDeviceToHost, s_));
        HIP_CHECK(hipStreamSynchronize(s_));  // partial is freed on return
    }

    // coCEOs buckets over rows 0..n-1 of X (ring slots 0..n-1): see ceos_build.hip.
    // src (optional, a device copy of those rows, 16-byte aligned): the scans read it
    // instead, and the first one also stores the rows into X (after reserve_points), so
    // the points are read from HBM once for both.
    void build_buckets(int64_t n, hipStream_t s_, const float* src = nullptr) {
        const int DR_ = DR();
        const int nl = 2 * DR_;
        const IndexDims dm = dims();
        if (m > n) throw ArgError("top_m must not exceed the number of points");
        const float* rows = src ? src : X.p;
        float* copy_to = src ? X.p : nullptr;

        // sample size: about 16 n/m points (at least 32768), so a list's top m hold about
        // lambda = m S/n >= 16 sample points. The threshold is the sample value of rank
        // k = 2.5 lambda + 8 (about 48); the points passing it number about n/S * Gamma(k).
        // A list under m (P(Poisson(lambda) >= k), about 1e-10 at lambda = 16) or over its
        // space n/S (k + 8 sqrt(k) + 16) + 256 (an 8-sigma Gamma tail) is checked for and
        // rescanned.
        int64_t S_target = std::max<int64_t>(32768, (16 * n + m - 1) / m);
        S_target = std::min<int64_t>(S_target, std::min<int64_t>((int64_t(2) << 30) / (4LL * DR_), 65535LL * 64));
        S_target = std::max<int64_t>(S_target, 1);
        const int64_t stride = std::max<int64_t>(1, n / S_target);
        const int64_t S = (n + stride - 1) / stride;
        const bool exact = stride == 1;
        const int rank = exact ? m : (int)std::min<int64_t>(S, (5 * (int64_t)m * S + 2 * n - 1) / (2 * n) + 8);

        DBuf<float>& thr = bw_thr;
        thr.ensure(nl);
        {
            // project the sample row-major (coalesced), then transpose to one column per list
            bw_rows.ensure((size_t)DR_ * S);
            bw_sample.ensure((size_t)DR_ * S);
            HIP_CHECK(ceos::launch_project_rows(dm, rows, 0, stride, S, signs.p, bw_rows.p, DR_, false, s_));
            HIP_CHECK(ceos::launch_transpose(bw_rows.p, S, DR_, bw_sample.p, s_));
            HIP_CHECK(ceos::launch_sample_thresholds(bw_sample.p, S, DR_, rank, thr.p, s_));
        }
        // One scan appends every point that passes its list's threshold; the lists are then
        // selected without waiting for the counts (the common case ends here, with one
        // synchronisation). A list left under m entries (threshold too high for it) is
        // rescanned with no threshold and capacity n; one over its capacity, with capacity
        // = its count. Rescans cover contiguous ranges of lists whose capacities fit the
        // entry budget, the lists outside a range inactive, so the memory stays bounded
        // however many lists the strided sample misjudged.
        const int32_t cap0 = (int32_t)std::min<int64_t>(
            n, (int64_t)std::ceil((double)n / (double)S * (rank + 8.0 * std::sqrt((double)rank) + 16.0)) + 256);
        std::vector<int64_t> hbase(nl);
        std::vector<int32_t> hcap(nl, cap0);
        DBuf<int64_t>& dbase = bw_base;
        DBuf<int32_t>& dcap = bw_cap;
        DBuf<int32_t>& dcount = bw_count;
        DBuf<Entry>& ent = bw_ent;
        HBuf<int32_t>& hcount = bw_hcount;
        dbase.ensure(nl);
        dcap.ensure(nl);
        dcount.ensure(nl);
        hcount.ensure(nl);
        bkt.ensure((size_t)nl * m);
        bsize.ensure(nl);
        auto scan_select = [&](int lo, int hi, const uint8_t* active) {
            int64_t total = 0;
            for (int l = lo; l < hi; ++l) { hbase[l] = total; total += hcap[l]; }
            ent.ensure((size_t)total);
            HIP_CHECK(hipMemcpyAsync(dbase.p + lo, hbase.data() + lo, (size_t)(hi - lo) * 8, hipMemcpyHostToDevice, s_));
            HIP_CHECK(hipMemcpyAsync(dcap.p + lo, hcap.data() + lo, (size_t)(hi - lo) * 4, hipMemcpyHostToDevice, s_));
            HIP_CHECK(ceos::launch_fill_i32(dcount.p, 0, nl, s_));
            HIP_CHECK(ceos::launch_scan_append(dm, rows, n, signs.p, thr.p, active, dbase.p, dcap.p, dcount.p,
                                               ent.p, s_, copy_to));
            copy_to = nullptr;  // rescans leave X as it is
            HIP_CHECK(hipMemcpyAsync(hcount.p + lo, dcount.p + lo, (size_t)(hi - lo) * 4, hipMemcpyDeviceToHost, s_));
            phase_mark(2, s_);
            HIP_CHECK(ceos::launch_list_select(hi - lo, m, dbase.p + lo, dcap.p + lo, dcount.p + lo, ent.p,
                                               bkt.p + (size_t)lo * m, bsize.p + lo, s_));
            phase_mark(3, s_);
            HIP_CHECK(hipStreamSynchronize(s_));
        };
        scan_select(0, nl, nullptr);

        std::vector<uint8_t> bad(nl, 0);
        std::vector<float> hthr;
        bool any_bad = false;
        rescan = RescanStats{};
        for (int l = 0; l < nl; ++l) {
            if (hcount.p[l] < m) {
                ++rescan.under;
                if (hthr.empty()) {
                    hthr.resize(nl);
                    HIP_CHECK(hipMemcpy(hthr.data(), thr.p, nl * 4, hipMemcpyDeviceToHost));
                }
                hthr[l] = -INFINITY;  // accept every point for this list
                hcap[l] = (int32_t)n;
                bad[l] = any_bad = 1;
            } else if (hcount.p[l] > hcap[l]) {
                ++rescan.over;
                hcap[l] = hcount.p[l];  // the same threshold admits exactly this many again
                bad[l] = any_bad = 1;
            }
        }
        if (!any_bad) return;
        if (!hthr.empty()) HIP_CHECK(hipMemcpy(thr.p, hthr.data(), nl * 4, hipMemcpyHostToDevice));
        const int64_t budget = std::max<int64_t>(rescan_budget_entries, n);
        std::vector<uint8_t> hact(nl);
        bw_active.ensure(nl);
        for (int lo = 0; lo < nl;) {
            if (!bad[lo]) { ++lo; continue; }
            // [lo, hi): from a misjudged list, as far as the budget allows, ending on one
            int hi = lo + 1, end = lo + 1;
            int64_t total = hcap[lo];
            while (hi < nl && total + hcap[hi] <= budget) {
                total += hcap[hi];
                ++hi;
                if (bad[hi - 1]) end = hi;
            }
            hi = end;
            std::fill(hact.begin(), hact.end(), 0);
            std::fill(hact.begin() + lo, hact.begin() + hi, 1);
            HIP_CHECK(hipMemcpyAsync(bw_active.p, hact.data(), nl, hipMemcpyHostToDevice, s_));
            scan_select(lo, hi, bw_active.p);
            ++rescan.ranges;
            for (int l = lo; l < hi; ++l)
                if (hcount.p[l] < m || hcount.p[l] > hcap[l]) throw DevError("bucket construction did not converge");
            lo = hi;
        }
    }
    // Phase record of a build that marked 0 / 1 around its copy and ran build_buckets.
    void finish_build_times(std::chrono::steady_clock::time_point t0) {
        last_build.copy_ms = phase_ms(0, 1);
        last_build.project_ms = phase_ms(1, 2);
        last_build.select_ms = phase_ms(2, 3);
        last_build.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        last_build.index_gb = index_gb();
    }

    // Picks the LDS score-tile size and grid for a search with the given bounds:
    // the largest tile that fits the per-workgroup LDS share (wg_per_cu workgroups
    // of 512 threads per CU; default 1): the most 4-slot buckets, a multiple of 64.
    ceos::QueryConfig search_config(int s_max, int c_max, int k_max) {
        ceos::QueryConfig cfg = search_config_with(s_max, c_max, k_max, wg_per_cu, table_log2);
        arm_timing(cfg);
        return cfg;
    }
    // The same for an explicit launch shape (0 = automatic), without kernel timing.
    ceos::QueryConfig search_config_with(int s_max, int c_max, int k_max, int wg_req, int tl_req) {
        // a query never has more distinct candidates than its probed entries
        c_max = (int)std::min<int64_t>(c_max, (int64_t)s_max * std::max(1, std::min(mp, m)));
        const int wg = wg_req > 0 ? wg_req : 1;
        const int budget = std::min(kMaxLdsBytes, (160 * 1024) / wg - 2048);
        // explicit: 2^tl_req slots; automatic: the most buckets (a multiple of 64, at
        // least 64, at most 2^13) that fit the budget
        int nbk = 0;
        if (tl_req > 0) {
            nbk = 1 << (tl_req - 2);
        } else {
            for (int cand = 8192; cand >= 64; cand -= 64)
                if (ceos::search_lds_bytes(d, cand, s_max, c_max, k_max) <= budget) { nbk = cand; break; }
            if (nbk == 0) nbk = 64;
        }
        int t = 0;  // floor(log2(4 nbk))
        while ((2 << t) <= 4 * nbk) ++t;
        const int lds = ceos::search_lds_bytes(d, nbk, s_max, c_max, k_max);
        if (lds > kMaxLdsBytes)
            throw ArgError("search parameters need " + std::to_string(lds) +
                           " bytes of LDS per workgroup (n_cand or n_probed_vectors too large)");
        if (std::min(mp, m) > ceos::kMaxProbedPoints)
            throw ArgError("n_probed_points above " + std::to_string(ceos::kMaxProbedPoints) + " is not supported");
        if ((int64_t)s_max * std::min(mp, m) > (int64_t(1) << 26))
            throw ArgError("n_probed_vectors * n_probed_points above 2^26 is not supported");
        ceos::QueryConfig cfg{};
        cfg.table_log2 = t;
        cfg.tile_buckets = nbk;
        cfg.lds_bytes = lds;
        const int fit = std::max(1, (160 * 1024) / (lds + 2048));
        cfg.grid = device_cus() * std::min(wg, fit);
        cfg.s_max = s_max; cfg.c_max = c_max; cfg.k_max = k_max;
        cfg.cand_stride = (int64_t)s_max * std::max(1, std::min(mp, m));
        last_lds = lds; last_tlog2 = t; last_wg = std::min(wg, fit);
        return cfg;
    }

    // The index's scratch sized for a launch; a launch still in flight keeps its buffers
    // until it has finished.
    ceos::QueryScratch scratch(const ceos::QueryConfig& cfg, int nq) {
        if (scr_pending && !scr.fits(cfg, nq, DR())) HIP_CHECK(hipEventSynchronize(scr_ev));
        return scr.ensure(cfg, nq, DR());
    }

    ceos::RowRing rows() const { return ceos::RowRing{X.p, cap, head, base, live, Xq.p}; }

    // Brings the rerank bounds up to date after rows were written at ring slots
    // [slot0, slot0 + count) mod cap. When the copies are missing or sized for another
    // ring, they are (re)made for the `total` rows from slot `head`. Needs d % 4 == 0.
    void quantize_rows(int64_t slot0, int64_t count, int64_t total, hipStream_t s_) {
        if (!rerank_bounds || d % 4 != 0 || cap <= 0) {
            Xq.release();
            return;
        }
        if (Xq.n != (size_t)cap * ceos::quant_stride(d)) {
            Xq.alloc((size_t)cap * ceos::quant_stride(d));
            slot0 = head;
            count = total;
        }
        HIP_CHECK(ceos::launch_quantize_rows(X.p, slot0, count, cap, d, Xq.p, s_));
    }
    // The rerank bounds of rows 0..n-1 made from src (a device copy of those rows, read
    // only) on the second stream, after the work queued on s_ so far: the memory-bound
    // pass overlaps the compute-bound bucket scan. join_aux(s_) orders later work after it.
    void quantize_rows_side(const float* src, int64_t n, hipStream_t s_) {
        if (!rerank_bounds || d % 4 != 0 || cap <= 0) {
            Xq.release();
            return;
        }
        if (Xq.n != (size_t)cap * ceos::quant_stride(d)) {
            Xq.alloc((size_t)cap * ceos::quant_stride(d));
        }
        hipStream_t a = aux();
        HIP_CHECK(hipEventRecord(fork_ev, s_));
        HIP_CHECK(hipStreamWaitEvent(a, fork_ev, 0));
        HIP_CHECK(ceos::launch_quantize_rows(src, 0, n, n, d, Xq.p, a));
        HIP_CHECK(hipEventRecord(join_ev, a));
        aux_pending = true;
    }
    void set_rerank_bounds(bool on) {
        wait_reads();
        rerank_bounds = on;
        if (!on) {
            Xq.release();
        } else if (built && bkt.p && X.p) {
            quantize_rows(head, live, live, st());
            HIP_CHECK(hipStreamSynchronize(st()));
        }
    }

    void check_status() {
        int32_t flag = 0;
        HIP_CHECK(hipMemcpy(&flag, scr.status.p, 4, hipMemcpyDeviceToHost));
        if (flag) {
            HIP_CHECK(hipMemset(scr.status.p, 0, 4));
            throw DevError("LDS score tile overflowed; rerun with a larger table (ceos_set_search_config)");
        }
    }

    // The coCEOs / CEOs-Hash lists exist (a CEOs-Est build keeps dense projections instead).
    void require_lists() const {
        if (!built) throw StateError("index is not built");
        if (!bkt.p || !bsize.p || !signs.p || !X.p)
            throw StateError("coCEOs / CEOs-Hash index is not built (build_coCEOs_Est or build_CEOs_Hash)");
    }

    void validate_query(int s_, int mp_, int c_, int k, int s_limit, const char* s_limit_name) const {
        require_lists();
        if (mp_ > m) throw ArgError("Number of probed points must be smaller than number of indexed top-m points");
        if (s_ > s_limit)
            throw ArgError(std::string("Number of probed vectors must be smaller than ") + s_limit_name);
        if (s_ <= 0 || mp_ <= 0 || c_ <= 0) throw ArgError("n_probed_vectors, n_probed_points and n_cand must be positive");
        if (k < 0 || k > kMaxK) throw ArgError("n_neighbors must be in [0, " + std::to_string(kMaxK) + "]");
        if (s_ > kMaxProbes) throw ArgError("n_probed_vectors above " + std::to_string(kMaxProbes) + " is not supported");
    }

    // Uniform-parameter coCEOs search on device pointers.
    // Uniform-parameter search on device pointers; hash = CEOs-Hash (every distinct
    // candidate reranked, c unused).
    hipEvent_t sev[3] = {};  // host-entry searches: before the front kernels, before and after the fused kernel
    void search_device(const float* dQ, int nq, int k, int s_, int c_, int32_t* ids, float* ips, hipStream_t s__,
                       bool hash = false, bool phase_events = false) {
        if (nq <= 0 || k == 0) return;
        if (hash) c_ = 1;
        ceos::QueryConfig cfg = search_config(s_, c_, k);
        if (phase_events && !cfg.ev[0]) {
            for (auto& e : sev)
                if (!e) HIP_CHECK(hipEventCreate(&e));
            for (int i = 0; i < 3; ++i) cfg.ev[i] = sev[i];
        }
        before_scratch(s__);
        const ceos::QueryScratch scr = scratch(cfg, nq);
        ceos::QueryBatch qb{dQ, nq, nullptr, nullptr, nullptr, k, s_, c_, mp, nullptr, ids, ips, nullptr, hash ? 1 : 0};
        HIP_CHECK(ceos::launch_coceos_search(dims(), signs.p, bkt.p, bsize.p, rows(), qb, cfg, scr, s__));
        after_scratch(s__);
    }

    void search_host(const float* Q, int nq, int k, int s_, int c_, int32_t* ids, float* ips, bool hash = false) {
        if (nq <= 0 || k == 0) return;
        const auto t0 = std::chrono::steady_clock::now();
        DBuf<float> dQ((size_t)nq * d);
        DBuf<int32_t> dI((size_t)nq * k);
        DBuf<float> dP((size_t)nq * k);
        HIP_CHECK(hipMemcpyAsync(dQ.p, Q, (size_t)nq * d * 4, hipMemcpyHostToDevice, st()));
        search_device(dQ.p, nq, k, s_, c_, dI.p, dP.p, st(), hash, !ktiming);
        HIP_CHECK(hipMemcpyAsync(ids, dI.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st()));
        HIP_CHECK(hipMemcpyAsync(ips, dP.p, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st()));
        HIP_CHECK(hipStreamSynchronize(st()));
        check_status();
        last_search = ceos_phase_times{};
        if (!ktiming && sev[2]) {
            float a = 0.0f, b = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&a, sev[0], sev[1]));
            HIP_CHECK(hipEventElapsedTime(&b, sev[1], sev[2]));
            last_search.project_ms = a;
            last_search.select_ms = b;
        }
        last_search.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        last_search.index_gb = index_gb();
    }

    // ---- index files (no reference counterpart): parameters, points in logical order,
    // buckets with logical ids, the stream's random-delete flags. Little-endian binary:
    //   "CEOSIDX1" | i32 kind, d, D, R, m, F, L, s, mp, c, centering | u32 seed |
    //   i64 live | X [live][d] f32 | buckets [2 DR][m] (u32 id, f32 value) | sizes [2 DR] i32 |
    //   delete flags [(live + 31) / 32] u32 | stream with centering: center [d] f32
    static constexpr char kMagic[8] = {'C', 'E', 'O', 'S', 'I', 'D', 'X', '1'};
    void save(const char* path, int kind) {
        if (!built) throw StateError("index is not built");
        FILE* f = std::fopen(path, "wb");
        if (!f) throw ArgError(std::string("cannot open ") + path + " for writing");
        auto put = [&](const void* p, size_t bytes) {
            if (bytes && std::fwrite(p, 1, bytes, f) != bytes) {
                std::fclose(f);
                throw DevError(std::string("write to ") + path + " failed");
            }
        };
        const int32_t hdr[11] = {kind, d, D, R, m, F, L, s, mp, c, centering ? 1 : 0};
        const int64_t n = live;
        put(kMagic, 8);
        put(hdr, sizeof(hdr));
        put(&seed_used, 4);
        put(&n, 8);
        // rows in logical order, through a bounded host buffer
        const int64_t chunk = std::max<int64_t>(1, (int64_t(64) << 20) / (4LL * d));
        std::vector<float> buf;
        for (int64_t i = 0; i < n;) {
            int64_t slot = (head + i) % cap;
            const int64_t len = std::min({chunk, n - i, cap - slot});
            buf.resize((size_t)len * d);
            HIP_CHECK(hipMemcpy(buf.data(), X.p + slot * d, (size_t)len * d * 4, hipMemcpyDeviceToHost));
            put(buf.data(), buf.size() * 4);
            i += len;
        }
        const int nl = 2 * DR();
        std::vector<Entry> h((size_t)nl * m);
        std::vector<int32_t> sz(nl);
        HIP_CHECK(hipMemcpy(h.data(), bkt.p, h.size() * sizeof(Entry), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(sz.data(), bsize.p, sz.size() * 4, hipMemcpyDeviceToHost));
        for (auto& e : h) e.x -= base;  // logical ids
        put(h.data(), h.size() * sizeof(Entry));
        put(sz.data(), sz.size() * 4);
        std::vector<uint32_t> flags((size_t)(n + 31) / 32, 0u);
        if (dead.p) {
            std::vector<uint32_t> ring((size_t)(cap + 31) / 32);
            HIP_CHECK(hipMemcpy(ring.data(), dead.p, ring.size() * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) {
                const int64_t slot = (head + i) % cap;
                if ((ring[slot >> 5] >> (slot & 31)) & 1u) flags[i >> 5] |= 1u << (i & 31);
            }
        }
        put(flags.data(), flags.size() * 4);
        if (kind == 1 && centering) put(center.data(), (size_t)d * 4);  // the stream's fixed center
        if (std::fclose(f) != 0) throw DevError(std::string("closing ") + path + " failed");
    }
    // Returns the number of points.
    int64_t load(const char* path, int kind) {
        FILE* f = std::fopen(path, "rb");
        if (!f) throw ArgError(std::string("cannot open ") + path);
        struct Closer {
            FILE* f;
            ~Closer() { std::fclose(f); }
        } closer{f};
        auto get = [&](void* p, size_t bytes) {
            if (bytes && std::fread(p, 1, bytes, f) != bytes)
                throw ArgError(std::string(path) + ": truncated index file");
        };
        auto invalid = [&](const char* what) { throw ArgError(std::string(path) + ": invalid " + what + " in the index file"); };
        char magic[8];
        int32_t hdr[11];
        uint32_t seed_v;
        int64_t n;
        get(magic, 8);
        if (std::memcmp(magic, kMagic, 8) != 0) throw ArgError(std::string(path) + " is not an index file");
        get(hdr, sizeof(hdr));
        get(&seed_v, 4);
        get(&n, 8);
        if (hdr[0] != kind)
            throw ArgError(std::string(path) + (kind ? " holds a CEOs index, not a streamCEOs one"
                                                     : " holds a streamCEOs index, not a CEOs one"));
        if (hdr[1] != d)
            throw ArgError("the index file has " + std::to_string(hdr[1]) + " features, the handle " +
                           std::to_string(d));
        // Everything in the file is checked before anything is allocated or copied to the
        // device: a foreign or damaged file must not size allocations or make the search
        // index rows out of bounds. The rows are skipped first and read last.
        const int fD = hdr[2], fR = hdr[3], fm = hdr[4], fF = hdr[5], fL = hdr[6];
        if (n < 0 || n >= (int64_t(1) << 31)) invalid("point count");
        if (fD <= 0 || fR <= 0 || fm <= 0) invalid("n_proj / n_repeats / top_m");
        if (fm > kMaxTopM || fm > n) invalid("top_m");
        if ((int64_t)fD * fR > (int64_t(1) << 20)) invalid("n_proj * n_repeats");
        if (fht_dim(fD, d) > 4096 || fF != fht_dim(fD, d) || fL < 0 || fL > 12 || (1 << fL) != fF) invalid("fhtDim");
        if (hdr[7] <= 0 || hdr[8] <= 0 || hdr[9] <= 0) invalid("query parameters");
        const int64_t rows_at = (int64_t)ftello(f);
        if (rows_at < 0 || fseeko(f, (off_t)(rows_at + n * (int64_t)d * 4), SEEK_SET) != 0)
            throw ArgError(std::string(path) + ": truncated index file");
        const int nl = 2 * fD * fR;
        std::vector<Entry> h((size_t)nl * fm);
        std::vector<int32_t> sz(nl);
        std::vector<uint32_t> flags((size_t)(n + 31) / 32);
        get(h.data(), h.size() * sizeof(Entry));
        get(sz.data(), sz.size() * 4);
        get(flags.data(), flags.size() * 4);
        std::vector<float> ctr;
        if (kind == 1 && hdr[10] != 0) {
            ctr.resize(d);
            get(ctr.data(), (size_t)d * 4);
        }
        // every list holds at most top_m valid entries, each a stored point
        for (int l = 0; l < nl; ++l) {
            if (sz[l] < 0 || sz[l] > fm) invalid("bucket size");
            for (int t = 0; t < sz[l]; ++t)
                if ((int64_t)h[(size_t)l * fm + t].x >= n) invalid("bucket id");
        }
        if (fseeko(f, (off_t)rows_at, SEEK_SET) != 0) throw ArgError(std::string(path) + ": cannot read the rows");

        wait_reads();
        built = false;
        bkt.release();
        bsize.release();
        dead.release();
        estP.release();
        Xq.release();
        D = fD; R = fR; m = fm; F = fF; L = fL;
        s = hdr[7]; mp = hdr[8]; c = hdr[9]; centering = hdr[10] != 0;
        seed = (int)(seed_v & 0x7fffffffu) == (int)seed_v ? (int)seed_v : -1;
        params_set = true;
        center = std::move(ctr);
        dcenter.release();
        if (!center.empty()) {  // a stream's centre, subtracted from later inserts
            dcenter.alloc(d);
            HIP_CHECK(hipMemcpy(dcenter.p, center.data(), (size_t)d * 4, hipMemcpyHostToDevice));
        }
        X.alloc((size_t)std::max<int64_t>(n, 1) * d);
        cap = std::max<int64_t>(n, 1); head = 0; base = 0; live = (uint32_t)n;
        const int64_t chunk = std::max<int64_t>(1, (int64_t(64) << 20) / (4LL * d));
        std::vector<float> buf;
        for (int64_t i = 0; i < n; i += chunk) {
            const int64_t len = std::min(chunk, n - i);
            buf.resize((size_t)len * d);
            get(buf.data(), buf.size() * 4);
            HIP_CHECK(hipMemcpy(X.p + i * d, buf.data(), buf.size() * 4, hipMemcpyHostToDevice));
        }
        bkt.alloc(h.size());
        bsize.alloc(sz.size());
        HIP_CHECK(hipMemcpy(bkt.p, h.data(), h.size() * sizeof(Entry), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(bsize.p, sz.data(), sz.size() * 4, hipMemcpyHostToDevice));
        if (kind == 1) {  // streams always carry the flags (ring slots = logical rows here)
            flags.resize((size_t)(cap + 31) / 32, 0u);
            dead.alloc(flags.size());
            HIP_CHECK(hipMemcpy(dead.p, flags.data(), flags.size() * 4, hipMemcpyHostToDevice));
        }
        make_signs_from(seed_v);
        quantize_rows(0, n, n, st());
        HIP_CHECK(hipStreamSynchronize(st()));
        built = true;
        return n;
    }

    void get_buckets(int32_t* cid, float* cval, int32_t* fid, float* fval, int32_t* csz, int32_t* fsz) {
        require_lists();
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

inline void check_matrix(const void* p, int64_t n, int d, int expect_d) {
    if (!p && n > 0) throw ArgError("null data pointer");
    if (n < 0) throw ArgError("negative row count");
    if (d != expect_d)
        throw ArgError("feature dimension " + std::to_string(d) + " does not match the index (" +
                       std::to_string(expect_d) + ")");
}

}  // namespace ceos_rt

using namespace ceos_rt;

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
        const void *signs, *bkt, *bsize, *X, *Xq;
        int64_t cap, head;
        uint32_t base, live;
        int mp, m, D, R, L, d;
        bool operator==(const Snap& o) const {
            return signs == o.signs && bkt == o.bkt && bsize == o.bsize && X == o.X && Xq == o.Xq && cap == o.cap &&
                   head == o.head && base == o.base && live == o.live && mp == o.mp && m == o.m && D == o.D &&
                   R == o.R && L == o.L && d == o.d;
        }
    } snap{};
    bool tuned = false;
    double tuned_ms = 0.0;
    static constexpr int kTuneRuns = 3;
    struct TuneRecord {
        int wg, tl, pw;
        double ms;
    };
    std::vector<TuneRecord> tune_log;  // every shape the tuner timed, in order
    int wg = 0, tl = 0, pw = 4;  // launch shape: workgroups per CU, tile log2, projection waves per workgroup

    // the last run: later runs on another stream wait for it; capture and destruction
    // wait for it on the host (the plan's scratch and graph are in use until it ends)
    hipEvent_t run_ev = nullptr;
    hipStream_t run_stream = nullptr;
    bool run_pending = false;

    ~ceos_plan() {
        if (run_ev) {
            if (run_pending) (void)hipEventSynchronize(run_ev);
            (void)hipEventDestroy(run_ev);
        }
        drop();
        if (cap) (void)hipStreamDestroy(cap);
    }
    void wait_runs() {
        if (run_pending) HIP_CHECK(hipEventSynchronize(run_ev));
        run_pending = false;
    }
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    Snap current() const {
        const Core& x = *core;
        return Snap{x.signs.p, x.bkt.p, x.bsize.p, x.X.p, x.Xq.p, x.cap, x.head, x.base, x.live, x.mp, x.m, x.D, x.R,
                    x.L, x.d};
    }
    void capture(int wg_, int tl_, int pw_) {
        wait_runs();
        drop();
        Core& x = *core;
        cfg = x.search_config_with(s_max, c_max, k_max, wg_, tl_);
        cfg.proj_waves = pw_;
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
        pw = pw_;
    }
    void launch(int nq, hipStream_t st) {
        if (run_pending && run_stream != st) HIP_CHECK(hipStreamWaitEvent(st, run_ev, 0));
        HIP_CHECK(ceos::launch_fill_i32(nq_dev.p, nq, 1, st));  // a kernel, outside the graph
        HIP_CHECK(hipGraphLaunch(exec, st));
        if (!run_ev) HIP_CHECK(hipEventCreateWithFlags(&run_ev, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(run_ev, st));
        run_stream = st;
        run_pending = true;
        core->note_read(st);
    }
    // Times each candidate launch shape once on this batch and keeps the fastest. First
    // the fused kernel's shape (workgroups per CU, score-tile size), then, with that, the
    // query projection's waves per workgroup. Every candidate computes the same results,
    // so the batch's outputs are valid afterwards.
    void tune(int nq, hipStream_t st) {
        Core& x = *core;
        // automatic tiles (the most buckets that fit) and the power of two below each
        const ceos::QueryConfig c1 = x.search_config_with(s_max, c_max, k_max, 1, 0);
        const ceos::QueryConfig c2 = x.search_config_with(s_max, c_max, k_max, 2, 0);
        struct Shape { int wg, tl, pw; };
        std::vector<Shape> cands = {{1, 0, 4}, {2, 0, 4}};
        const int p1 = (1 << c1.table_log2) == 4 * c1.tile_buckets ? c1.table_log2 - 1 : c1.table_log2;
        const int p2 = (1 << c2.table_log2) == 4 * c2.tile_buckets ? c2.table_log2 - 1 : c2.table_log2;
        if (p1 >= 8) cands.push_back({1, p1, 4});
        if (p2 >= 8) cands.push_back({2, p2, 4});
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        double best = 1e30;
        Shape pick = cands[0];
        tune_log.clear();
        // each shape: one warm-up launch, then the median of kTuneRuns timed launches
        auto time_shape = [&](const Shape& cd) {
            capture(cd.wg, cd.tl, cd.pw);
            launch(nq, st);  // warm
            float t[kTuneRuns];
            for (int i = 0; i < kTuneRuns; ++i) {
                HIP_CHECK(hipEventRecord(e0, st));
                launch(nq, st);
                HIP_CHECK(hipEventRecord(e1, st));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipEventElapsedTime(&t[i], e0, e1));
            }
            std::sort(t, t + kTuneRuns);
            const double ms = t[kTuneRuns / 2];
            tune_log.push_back({cd.wg, cd.tl, cd.pw, ms});
            if (ms < best) {
                best = ms;
                pick = cd;
            }
        };
        try {
            for (const auto& cd : cands) time_shape(cd);
            const Shape base = pick;
            for (int pw_ : {2, 8}) time_shape({base.wg, base.tl, pw_});
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        capture(pick.wg, pick.tl, pick.pw);
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
        if (!exec || !(snap == current())) capture(wg, tl, pw);
        if ((flags & CEOS_PLAN_AUTOTUNE) && !tuned)
            tune(nq, st);
        else
            launch(nq, st);
    }
};

namespace ceos_rt {
inline int make_plan(Core& x, bool stream_index, int nq_max, int k_max, int s_max, int c_max, int flags, ceos_plan** out) {
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
}  // namespace ceos_rt

