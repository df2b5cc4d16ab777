// Serving engine of libceos_amd: continuous batching of coCEOs queries over a resident
// index (include/ceos_amd.h, "serving"). No reference counterpart: the reference only
// searches a whole query matrix per call (src/CEOs.cpp:450-694).
//
// One worker thread owns the device side. Requests (one query each, with its own k,
// n_probed_vectors and n_cand) come in submissions: a submission holds its rows,
// parameters and arrival offsets, prepared outside the server lock and queued in one
// step, and the arrival clock starts when it is queued. The worker admits due requests
// from each submission's cursor, earliest arrival first. Whenever one of the
// server's `inflight` batch slots is free, the worker takes the requests that have
// arrived into that slot (up to max_batch; a backlog is shared among the free slots):
// pinned host staging -> device buffers of a captured search plan -> plan replay ->
// results back, all on the slot's own stream, ending in an event. Requests that arrive
// while batches run are admitted into the next free slot, so batches overlap on the
// device (the next persistent grid fills the CUs as the previous one drains) and with
// the host copies. The number of slots
// is bounded by the HBM budget: each slot holds one plan's buffers and scratch.
// streamCEOs updates and deletes are applied between batches: the worker waits for
// the batches in flight, applies them in submission order, and the next batch sees
// the new index.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <memory>
#include <queue>
#include <thread>
#include <unordered_map>

#include "ceos_core.h"

namespace {

using Clock = std::chrono::steady_clock;

// One submission: n requests with tickets first_ticket .. first_ticket + n - 1.
struct Feed {
    int64_t first_ticket = 0;
    int n = 0;
    std::vector<float> Q;              // n x d
    std::vector<int32_t> k, s, c;
    std::vector<double> off;           // arrival offsets (seconds after t0)
    std::vector<int32_t> order;        // request indices by ascending offset (admission order)
    int next = 0;                      // admitted so far (position in order)
    Clock::time_point t0;
    enum : uint8_t { kQueued = 0, kDone = 1, kFailed = 2, kTaken = 3 };
    std::vector<uint8_t> state;
    std::vector<std::vector<int32_t>> ids;
    std::vector<std::vector<float>> ips;
    std::vector<double> latency_ms;
    int untaken = 0;                   // results not yet returned by ceos_server_result

    Clock::time_point arrival(int i) const {
        return t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(off[i]));
    }
    bool pending() const { return next < n; }
    Clock::time_point next_arrival() const { return arrival(order[next]); }
};

// Append-only record of fixed-size chunks that are kept across clear(): appending never
// moves what is recorded (a vector doubling under the server lock copied megabytes into
// freshly faulted pages and stalled the worker for milliseconds at 2^19, 2^20, ...
// requests), and a cleared record refills its chunks without allocating.
template <typename T>
struct Record {
    static constexpr size_t kChunk = size_t(1) << 15;
    std::vector<std::unique_ptr<T[]>> chunks;
    size_t n = 0;
    void push_back(const T& v) {
        if (n == chunks.size() * kChunk) chunks.emplace_back(new T[kChunk]);
        chunks[n / kChunk][n % kChunk] = v;
        ++n;
    }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    void clear() { n = 0; }
    const T& operator[](size_t i) const { return chunks[i / kChunk][i % kChunk]; }
    const T& back() const { return (*this)[n - 1]; }
    // The first m entries into out.
    void copy_to(T* out, size_t m) const {
        for (size_t i = 0; i < m; i += kChunk) std::copy_n(chunks[i / kChunk].get(), std::min(kChunk, m - i), out + i);
    }
    std::vector<T> to_vector() const {
        std::vector<T> v(n);
        copy_to(v.data(), n);
        return v;
    }
};

struct Update {
    bool is_delete = false;
    std::vector<float> X;  // new rows (insert)
    int64_t n_new = 0;
    int n_del = 0;
    std::vector<int32_t> ids;  // random delete
};

// Pinned host memory (hipHostMalloc), mapped into the device's address space: `dp` is the
// address kernels use. Coherent, so a kernel reads what the host wrote before its launch
// and the host reads what a kernel wrote once the kernel's completion event is seen.
template <typename T>
struct HBuf {
    T* p = nullptr;
    T* dp = nullptr;
    void alloc(size_t n) {
        if (hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(1, n) * sizeof(T),
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw MemError("pinned host allocation failed");
        }
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess || !d) {
            (void)hipGetLastError();
            (void)hipHostFree(p);
            p = nullptr;
            throw MemError("pinned host memory is not mapped for the device");
        }
        dp = static_cast<T*>(d);
    }
    ~HBuf() {
        if (p) (void)hipHostFree(p);
    }
};

struct Slot {
    ceos_plan* plan = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    HBuf<float> hQ, hP;
    HBuf<int32_t> hk, hs, hc, hI;
    std::vector<int64_t> tickets;
    bool busy = false;
    Clock::time_point t_launch{};
    ceos_server_batch cur{};  // trace record of the batch in flight

    // Enqueues a staged batch of nb requests on the slot's stream: the inputs move from the
    // mapped staging into the plan's buffers by a kernel, the plan runs, the results move
    // back by a kernel, and the completion event is recorded. Copy-engine transfers
    // (hipMemcpyAsync) blocked the worker for several milliseconds at times. ts[1] and
    // ts[2] (optional) get the times after the inputs' launch and after the plan's.
    void enqueue(int nb, int d, int k_max, Clock::time_point* ts = nullptr) {
        float* dQ;
        int32_t *dk, *ds, *dc, *dI;
        float* dP;
        if (ceos_plan_buffers(plan, &dQ, &dk, &ds, &dc, &dI, &dP) != CEOS_OK) throw DevError(g_err);
        ceos::CopySegs in{};
        in.n = 4;
        in.src[0] = reinterpret_cast<const uint32_t*>(hQ.dp);
        in.dst[0] = reinterpret_cast<uint32_t*>(dQ);
        in.words[0] = (int64_t)nb * d;
        in.src[1] = reinterpret_cast<const uint32_t*>(hk.dp);
        in.dst[1] = reinterpret_cast<uint32_t*>(dk);
        in.src[2] = reinterpret_cast<const uint32_t*>(hs.dp);
        in.dst[2] = reinterpret_cast<uint32_t*>(ds);
        in.src[3] = reinterpret_cast<const uint32_t*>(hc.dp);
        in.dst[3] = reinterpret_cast<uint32_t*>(dc);
        in.words[1] = in.words[2] = in.words[3] = nb;
        HIP_CHECK(ceos::launch_copy_segments(in, st));
        if (ts) ts[1] = Clock::now();
        if (ceos_plan_run(plan, nb, st) != CEOS_OK) throw DevError(g_err);
        if (ts) ts[2] = Clock::now();
        ceos::CopySegs out{};
        out.n = 2;
        out.src[0] = reinterpret_cast<const uint32_t*>(dI);
        out.dst[0] = reinterpret_cast<uint32_t*>(hI.dp);
        out.src[1] = reinterpret_cast<const uint32_t*>(dP);
        out.dst[1] = reinterpret_cast<uint32_t*>(hP.dp);
        out.words[0] = out.words[1] = (int64_t)nb * k_max;
        HIP_CHECK(ceos::launch_copy_segments(out, st));
        HIP_CHECK(hipEventRecord(ev, st));
    }
};

}  // namespace

struct ceos_server {
    Core* core = nullptr;
    ceos_index* index = nullptr;
    ceos_stream* stream = nullptr;
    ceos_server_config cfg{};
    std::vector<std::unique_ptr<Slot>> slots;
    int64_t device_bytes = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Feed>> feeds;  // by first ticket
    // (next arrival, submission) for every submission with requests not yet admitted
    std::priority_queue<std::pair<Clock::time_point, Feed*>, std::vector<std::pair<Clock::time_point, Feed*>>,
                        std::greater<>>
        due;
    int64_t queued = 0;                       // requests submitted and not yet admitted
    std::deque<Update> updates;
    int64_t next_ticket = 1;
    bool stop = false;
    std::string worker_error;  // fatal: the worker stopped
    std::string update_error;  // the last failed update (reported by ceos_server_drain)
    std::thread worker;

    // statistics
    int64_t submitted = 0, completed = 0, batches = 0, batched_queries = 0;
    Record<double> lat_ms;
    Record<double> wait_ms;     // per request: arrival -> its batch launched
    Record<double> service_ms;  // per batch: launch -> results in host memory
    Record<ceos_server_batch> trace;  // per batch (ceos_server_get_batches), at most kMaxTrace
    static constexpr size_t kMaxTrace = 1 << 20;
    int64_t n_upd[2] = {0, 0};                           // inserts, deletes applied
    double upd_ms_sum[2] = {0.0, 0.0}, upd_ms_max[2] = {0.0, 0.0};
    Clock::time_point first_arrival{}, last_done{};
    bool any_arrival = false;

    ~ceos_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (auto& sl : slots) {
            if (sl->plan) ceos_plan_destroy(sl->plan);
            if (sl->ev) (void)hipEventDestroy(sl->ev);
            if (sl->st) (void)hipStreamDestroy(sl->st);
        }
    }

    int k_max() const { return cfg.k_max; }

    // The submission holding a ticket and the request's index in it, or null (mu held).
    Feed* find(int64_t ticket, int* idx) {
        auto it = std::upper_bound(feeds.begin(), feeds.end(), ticket,
                                   [](int64_t t, const std::unique_ptr<Feed>& f) { return t < f->first_ticket; });
        if (it == feeds.begin()) return nullptr;
        Feed* f = (--it)->get();
        if (ticket >= f->first_ticket + f->n) return nullptr;
        *idx = (int)(ticket - f->first_ticket);
        return f;
    }
    // The submission whose next request arrives first, or null when none is queued (mu held).
    Feed* earliest() { return due.empty() ? nullptr : due.top().second; }
    // Admits the next request of the earliest submission: its index (mu held).
    int admit() {
        Feed* f = due.top().second;
        due.pop();
        const int i = f->order[f->next++];
        if (f->pending()) due.push({f->next_arrival(), f});
        --queued;
        return i;
    }
    // Removes a submission whose results were all collected and returns it, to be freed
    // by the caller after releasing the lock (its rows can be gigabytes: freeing them
    // under the lock would stall the worker).
    std::unique_ptr<Feed> release_if_taken(Feed* f) {
        if (f->untaken > 0 || f->pending()) return nullptr;
        int idx = 0;
        if (find(f->first_ticket, &idx) != f) return nullptr;
        auto it = std::upper_bound(feeds.begin(), feeds.end(), f->first_ticket,
                                   [](int64_t t, const std::unique_ptr<Feed>& g) { return t < g->first_ticket; });
        --it;
        std::unique_ptr<Feed> out = std::move(*it);
        feeds.erase(it);
        return out;
    }

    // Moves a finished slot's results to its requests (mu held).
    void reap(Slot& sl) {
        const auto now = Clock::now();
        for (size_t i = 0; i < sl.tickets.size(); ++i) {
            int idx = 0;
            Feed* f = find(sl.tickets[i], &idx);
            // a request failed (fail_all) or already collected while its batch ran keeps
            // that state: it is neither given results nor counted again
            if (!f || f->state[idx] != Feed::kQueued) continue;
            const int kq = f->k[idx];
            f->ids[idx].assign(sl.hI.p + i * cfg.k_max, sl.hI.p + i * cfg.k_max + kq);
            f->ips[idx].assign(sl.hP.p + i * cfg.k_max, sl.hP.p + i * cfg.k_max + kq);
            f->latency_ms[idx] = std::chrono::duration<double, std::milli>(now - f->arrival(idx)).count();
            f->state[idx] = Feed::kDone;
            lat_ms.push_back(f->latency_ms[idx]);
            ++completed;
        }
        last_done = now;
        service_ms.push_back(std::chrono::duration<double, std::milli>(now - sl.t_launch).count());
        sl.cur.done_ms = service_ms.back();
        if (trace.size() < kMaxTrace) trace.push_back(sl.cur);
        sl.tickets.clear();
        sl.busy = false;
        cv_done.notify_all();
    }

    void fail_all(const std::string& msg) {  // mu held; the server stops serving
        worker_error = msg;
        for (auto& f : feeds) {
            for (int i = 0; i < f->n; ++i)
                if (f->state[i] == Feed::kQueued) f->state[i] = Feed::kFailed;
            f->next = f->n;
        }
        queued = 0;
        while (!due.empty()) due.pop();
        cv_done.notify_all();
    }

    // Launches one batch of arrived requests into a free slot (mu held, released
    // around the device calls).
    struct Taken {
        Feed* f;
        int i;
        float cost;
    };
    std::vector<Taken> taken_;  // the batch being launched (worker thread only)

    // Requests that have arrived by `now` and are not yet admitted, counted up to `limit`
    // (mu held).
    int64_t due_count(Clock::time_point now, int64_t limit) const {
        int64_t n = 0;
        for (const auto& f : feeds) {
            if (!f->pending()) continue;
            const auto first = f->order.begin() + f->next;
            const auto last = std::partition_point(first, f->order.end(), [&](int32_t i) { return f->arrival(i) <= now; });
            n += last - first;
            if (n >= limit) break;
        }
        return n;
    }

    // Queries for the next batch. A backlog is shared among the free slots (at least
    // kSplitMin each): rows are copied to the staging before a batch is launched, so one
    // slot taking the whole backlog after all slots finished together would leave the
    // device idle for that whole copy, and the next backlog would be larger still.
    static constexpr int kSplitMin = 256;
    int batch_cap(Clock::time_point now) const {
        int free_slots = 0;
        for (const auto& sl : slots) free_slots += !sl->busy;
        if (free_slots <= 1) return cfg.max_batch;
        const int64_t due = due_count(now, (int64_t)free_slots * cfg.max_batch);
        if (due <= kSplitMin) return cfg.max_batch;
        const int64_t share = (due + free_slots - 1) / free_slots;
        return (int)std::min<int64_t>(cfg.max_batch, std::max<int64_t>(kSplitMin, share));
    }

    void launch(Slot& sl, std::unique_lock<std::mutex>& lk) {
        const auto now = Clock::now();
        const int d = core->d;
        const float mp = (float)std::min(core->mp, core->m);
        const int cap = batch_cap(now);
        taken_.clear();
        while ((int)taken_.size() < cap) {
            Feed* f = earliest();
            if (!f || f->next_arrival() > now) break;
            const int i = f->order[f->next];
            admit();
            wait_ms.push_back(std::chrono::duration<double, std::milli>(now - f->arrival(i)).count());
            // bytes the search reads for this request: its probed bucket entries, its
            // candidates' int8 rows (or FP32 rows), its k exact rows
            const float cost = 8.0f * (float)f->s[i] * mp + (float)(d + 16) * (float)f->c[i] + 4.0f * d * (float)f->k[i];
            taken_.push_back({f, i, cost});
        }
        if (taken_.empty()) return;
        // Largest first: the persistent grid hands the batch's queries out in order, so the
        // heavy ones start first and light ones fill the workgroups that finish early; a
        // batch lasts as long as its slowest workgroup.
        std::stable_sort(taken_.begin(), taken_.end(), [](const Taken& a, const Taken& b) { return a.cost > b.cost; });
        const int nb = (int)taken_.size();
        for (int b = 0; b < nb; ++b) {
            const Feed* f = taken_[b].f;
            const int i = taken_[b].i;
            sl.hk.p[b] = f->k[i];
            sl.hs.p[b] = f->s[i];
            sl.hc.p[b] = f->c[i];
            sl.tickets.push_back(f->first_ticket + i);
        }
        sl.busy = true;
        sl.t_launch = now;
        sl.cur = ceos_server_batch{};
        sl.cur.launch_s = std::chrono::duration<double>(now - first_arrival).count();
        sl.cur.oldest_wait_ms = 0.0;
        for (const Taken& t : taken_)
            sl.cur.oldest_wait_ms =
                std::max(sl.cur.oldest_wait_ms, std::chrono::duration<double, std::milli>(now - t.f->arrival(t.i)).count());
        sl.cur.queries = nb;
        for (size_t j = 0; j < slots.size(); ++j)
            if (slots[j].get() == &sl) sl.cur.slot = (int32_t)j;
        ++batches;
        batched_queries += nb;
        lk.unlock();
        // rows to the pinned staging outside the lock; a submission keeps its rows until
        // all its results are collected (it is freed then, outside the lock), and no result
        // of this batch can be collected before it completes, so the submissions stay alive
        for (int b = 0; b < nb; ++b)
            std::memcpy(sl.hQ.p + (size_t)b * d, taken_[b].f->Q.data() + (size_t)taken_[b].i * d, (size_t)d * 4);
        Clock::time_point ts[5];
        ts[0] = Clock::now();
        try {
            sl.enqueue(nb, d, cfg.k_max, ts);
            ts[3] = Clock::now();
        } catch (const std::exception& e) {
            lk.lock();
            sl.busy = false;
            fail_all(e.what());
            return;
        }
        lk.lock();
        ts[4] = Clock::now();
        sl.cur.enqueue_ms = std::chrono::duration<double, std::milli>(ts[4] - sl.t_launch).count();
        for (int i = 0; i < 5; ++i)
            sl.cur.steps_ms[i] = std::chrono::duration<double, std::milli>(ts[i] - (i ? ts[i - 1] : sl.t_launch)).count();
    }

    void apply_update(Update& u, std::unique_lock<std::mutex>& lk) {
        lk.unlock();
        const auto t0 = Clock::now();
        int rc;
        if (u.is_delete)
            rc = ceos_stream_delete_ids(stream, u.ids.data(), (int64_t)u.ids.size());
        else
            rc = ceos_stream_update(stream, u.X.data(), u.n_new, core->d, u.n_del);
        const std::string err = rc == CEOS_OK ? std::string() : g_err;
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        lk.lock();
        const int kind = u.is_delete ? 1 : 0;
        ++n_upd[kind];
        upd_ms_sum[kind] += ms;
        upd_ms_max[kind] = std::max(upd_ms_max[kind], ms);
        if (rc != CEOS_OK) update_error = "stream update failed: " + err;  // searches go on
        cv_done.notify_all();
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // reap finished batches
            for (auto& sl : slots) {
                if (!sl->busy) continue;
                const hipError_t q = hipEventQuery(sl->ev);
                if (q == hipSuccess) {
                    reap(*sl);
                } else if (q == hipErrorNotReady) {
                    (void)hipGetLastError();  // not an error; keep it out of later launch checks
                } else {
                    fail_all(std::string("batch failed: ") + hipGetErrorString(q));
                    return;
                }
            }
            bool any_busy = false;
            for (auto& sl : slots) any_busy |= sl->busy;
            if (stop && queued == 0 && updates.empty() && !any_busy) break;
            // updates: between batches
            if (!updates.empty()) {
                if (any_busy) {
                    for (auto& sl : slots)
                        if (sl->busy) {
                            hipEvent_t ev = sl->ev;
                            lk.unlock();
                            (void)hipEventSynchronize(ev);
                            lk.lock();
                            reap(*sl);
                        }
                    continue;
                }
                Update u = std::move(updates.front());
                updates.pop_front();
                apply_update(u, lk);
                continue;
            }
            // launch into a free slot when a request has arrived
            const auto now = Clock::now();
            Feed* nf = earliest();
            const bool ready = nf && nf->next_arrival() <= now;
            Slot* free_slot = nullptr;
            for (auto& sl : slots)
                if (!sl->busy) {
                    free_slot = sl.get();
                    break;
                }
            if (ready && free_slot) {
                launch(*free_slot, lk);
                continue;
            }
            // wait: for a batch to finish (poll), an arrival, or a submission
            if (any_busy) {
                cv_work.wait_for(lk, std::chrono::microseconds(20));
            } else if (nf) {
                cv_work.wait_until(lk, nf->next_arrival());
            } else {
                cv_work.wait(lk, [&] { return stop || queued > 0 || !updates.empty(); });
            }
        }
    }
};

namespace {

int64_t plan_bytes(const Core& x, const ceos_server_config& c, const ceos::QueryConfig& qc) {
    int64_t b = (int64_t)c.max_batch * (x.d * 4 + 3 * 4 + (int64_t)c.k_max * 8);
    b += (int64_t)c.max_batch * (x.DR() * 4 + (int64_t)qc.s_max * 4);
    b += (int64_t)qc.grid * (qc.cand_stride * 16 + (int64_t)ceos::kSearchStashList * 16 +
                             4ll * ceos::search_tile_slots_nbk(qc.tile_buckets));
    return b;
}

int make_server(Core& x, ceos_index* hi, ceos_stream* hs, const ceos_server_config* cfg, ceos_server** out) {
    return guarded([&] {
        if (!out || !cfg) throw ArgError("null argument");
        if (cfg->max_batch <= 0 || cfg->k_max <= 0 || cfg->s_max <= 0 || cfg->c_max <= 0)
            throw ArgError("max_batch, k_max, s_max and c_max must be positive");
        {
            std::lock_guard<std::mutex> g(x.mu);
            x.require_lists();
        }
        auto srv = std::make_unique<ceos_server>();
        srv->core = &x;
        srv->index = hi;
        srv->stream = hs;
        srv->cfg = *cfg;
        int want = std::max(1, cfg->inflight);
        ceos::QueryConfig qc;
        {
            std::lock_guard<std::mutex> g(x.mu);
            qc = x.search_config_with(cfg->s_max, cfg->c_max, cfg->k_max, 0, 0);
        }
        const int64_t per = plan_bytes(x, *cfg, qc);
        if (cfg->hbm_budget > 0) {
            const int64_t fit = cfg->hbm_budget / per;
            if (fit < 1)
                throw MemError("hbm_budget of " + std::to_string(cfg->hbm_budget) + " bytes is below one batch slot (" +
                               std::to_string(per) + " bytes)");
            want = (int)std::min<int64_t>(want, fit);
        }
        for (int i = 0; i < want; ++i) {
            auto sl = std::make_unique<Slot>();
            const int rc = hi ? ceos_plan_create(hi, cfg->max_batch, cfg->k_max, cfg->s_max, cfg->c_max, 0, &sl->plan)
                              : ceos_stream_plan_create(hs, cfg->max_batch, cfg->k_max, cfg->s_max, cfg->c_max, 0,
                                                        &sl->plan);
            if (rc != CEOS_OK) throw DevError(g_err);
            HIP_CHECK(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
            HIP_CHECK(hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming));
            {
                // capture the slot's graph now and replay it once with an empty batch, so the
                // first requests do not pay the capture and the graph upload
                std::lock_guard<std::mutex> g(x.mu);
                sl->plan->capture(0, 0, 4);
                sl->plan->launch(0, sl->st);
                HIP_CHECK(hipStreamSynchronize(sl->st));
            }
            sl->hQ.alloc((size_t)cfg->max_batch * x.d);
            sl->hk.alloc(cfg->max_batch);
            sl->hs.alloc(cfg->max_batch);
            sl->hc.alloc(cfg->max_batch);
            sl->hI.alloc((size_t)cfg->max_batch * cfg->k_max);
            sl->hP.alloc((size_t)cfg->max_batch * cfg->k_max);
            {
                // one real batch through the path the worker takes (staging copies, the
                // graph on a non-empty batch, result copies), so that the first requests do
                // not pay first-use costs: a slow first batch lets requests pile up into
                // larger, slower batches at the start
                const int nb = std::min(cfg->max_batch, 64);
                for (int i = 0; i < nb; ++i) {
                    for (int j = 0; j < x.d; ++j)
                        sl->hQ.p[(size_t)i * x.d + j] = (float)(((i * 131 + j * 71) % 17) - 8) / 8.0f;
                    sl->hk.p[i] = std::min(cfg->k_max, 10);
                    sl->hs.p[i] = std::min(cfg->s_max, 16);
                    sl->hc.p[i] = std::min(cfg->c_max, 500);
                }
                sl->enqueue(nb, x.d, cfg->k_max);
                HIP_CHECK(hipStreamSynchronize(sl->st));
            }
            srv->slots.push_back(std::move(sl));
        }
        srv->device_bytes = per * want;
        ceos_server* p = srv.release();
        p->worker = std::thread([p] {
            try {
                p->run();
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> g(p->mu);
                p->fail_all(e.what());
            }
        });
        *out = p;
    });
}

void check_req(const ceos_server* s, int k, int sv, int c) {
    if (k < 0 || k > s->cfg.k_max || sv <= 0 || sv > s->cfg.s_max || c <= 0 || c > s->cfg.c_max)
        throw ArgError("k, n_probed_vectors or n_cand outside the server's bounds");
}

}  // namespace

extern "C" {

int ceos_server_create(ceos_index* h, const ceos_server_config* cfg, ceos_server** out) {
    if (!h) {
        g_err = "null handle";
        return CEOS_EINVAL;
    }
    return make_server(h->core, h, nullptr, cfg, out);
}

int ceos_stream_server_create(ceos_stream* h, const ceos_server_config* cfg, ceos_server** out) {
    if (!h) {
        g_err = "null handle";
        return CEOS_EINVAL;
    }
    return make_server(h->core, nullptr, h, cfg, out);
}

void ceos_server_destroy(ceos_server* s) { delete s; }

int ceos_server_submit(ceos_server* s, const float* Q, int n, int d, const int32_t* k, const int32_t* sv,
                       const int32_t* c, const double* arrival_s, int64_t* first_ticket) {
    return guarded([&] {
        if (!s || !Q || !k || !sv || !c || !first_ticket) throw ArgError("null argument");
        if (d != s->core->d) throw ArgError("feature dimension does not match the index");
        if (n <= 0) throw ArgError("n must be positive");
        for (int i = 0; i < n; ++i) check_req(s, k[i], sv[i], c[i]);
        // the submission is prepared outside the lock; its clock starts when it is queued
        auto f = std::make_unique<Feed>();
        f->n = n;
        f->Q.assign(Q, Q + (size_t)n * d);
        f->k.assign(k, k + n);
        f->s.assign(sv, sv + n);
        f->c.assign(c, c + n);
        f->off.assign(n, 0.0);
        if (arrival_s)
            for (int i = 0; i < n; ++i) f->off[i] = std::max(0.0, arrival_s[i]);
        f->order.resize(n);
        for (int i = 0; i < n; ++i) f->order[i] = i;
        std::stable_sort(f->order.begin(), f->order.end(), [&](int32_t a, int32_t b) { return f->off[a] < f->off[b]; });
        f->state.assign(n, Feed::kQueued);
        f->ids.resize(n);
        f->ips.resize(n);
        f->latency_ms.assign(n, 0.0);
        f->untaken = n;
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (!s->worker_error.empty()) throw DevError("server stopped: " + s->worker_error);
            if (s->cfg.max_pending > 0 && s->queued + n > s->cfg.max_pending)
                throw MemError("server queue is full");
            f->first_ticket = s->next_ticket;
            s->next_ticket += n;
            *first_ticket = f->first_ticket;
            f->t0 = Clock::now();
            const auto first = f->arrival(f->order[0]);
            if (!s->any_arrival || first < s->first_arrival) s->first_arrival = first;
            s->any_arrival = true;
            s->queued += n;
            s->submitted += n;
            s->due.push({first, f.get()});
            s->feeds.push_back(std::move(f));
        }
        s->cv_work.notify_one();
    });
}

int ceos_server_result(ceos_server* s, int64_t ticket, int wait, int32_t* ids, float* ips, int* k_out,
                       double* latency_ms) {
    if (!s) {
        g_err = "null server";
        return CEOS_EINVAL;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    int idx = 0;
    Feed* f = s->find(ticket, &idx);
    if (!f || f->state[idx] == Feed::kTaken) {
        g_err = "unknown ticket " + std::to_string(ticket);
        return CEOS_EINVAL;
    }
    if (f->state[idx] == Feed::kQueued) {
        if (!wait) {
            g_err = "not finished";
            return CEOS_EAGAIN;
        }
        s->cv_done.wait(lk, [&] { return f->state[idx] != Feed::kQueued; });
    }
    const bool failed = f->state[idx] == Feed::kFailed;
    if (!failed) {
        if (ids) std::copy(f->ids[idx].begin(), f->ids[idx].end(), ids);
        if (ips) std::copy(f->ips[idx].begin(), f->ips[idx].end(), ips);
        if (k_out) *k_out = f->k[idx];
        if (latency_ms) *latency_ms = f->latency_ms[idx];
    }
    f->state[idx] = Feed::kTaken;
    f->ids[idx] = std::vector<int32_t>();
    f->ips[idx] = std::vector<float>();
    --f->untaken;
    std::unique_ptr<Feed> gone = s->release_if_taken(f);
    const std::string err = failed ? s->worker_error : std::string();
    lk.unlock();
    gone.reset();  // a finished submission is freed outside the lock
    if (failed) {
        g_err = err;
        return CEOS_EDEVICE;
    }
    g_err.clear();
    return CEOS_OK;
}

int ceos_server_update(ceos_server* s, const float* Xnew, int64_t n_new, int d, int n_del) {
    return guarded([&] {
        if (!s) throw ArgError("null server");
        if (!s->stream) throw ArgError("updates need a streamCEOs server");
        if (n_new < 0 || n_del < 0 || (n_new > 0 && !Xnew)) throw ArgError("invalid update");
        if (d != s->core->d) throw ArgError("feature dimension does not match the index");
        Update u;
        u.X.assign(Xnew, Xnew + (size_t)n_new * d);
        u.n_new = n_new;
        u.n_del = n_del;
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->updates.push_back(std::move(u));
        }
        s->cv_work.notify_one();
    });
}

int ceos_server_delete(ceos_server* s, const int32_t* ids, int64_t count) {
    return guarded([&] {
        if (!s) throw ArgError("null server");
        if (!s->stream) throw ArgError("deletes need a streamCEOs server");
        if (count < 0 || (count > 0 && !ids)) throw ArgError("invalid delete");
        Update u;
        u.is_delete = true;
        u.ids.assign(ids, ids + count);
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->updates.push_back(std::move(u));
        }
        s->cv_work.notify_one();
    });
}

int ceos_server_drain(ceos_server* s) {
    return guarded([&] {
        if (!s) throw ArgError("null server");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] {
            if (!s->worker_error.empty()) return true;
            if (s->queued > 0 || !s->updates.empty()) return false;
            for (auto& sl : s->slots)
                if (sl->busy) return false;
            return true;
        });
        if (!s->worker_error.empty()) throw DevError(s->worker_error);
        if (!s->update_error.empty()) {
            const std::string e = s->update_error;
            s->update_error.clear();
            throw ArgError(e);
        }
    });
}

int ceos_server_get_stats(ceos_server* s, ceos_server_stats* out) {
    return guarded([&] {
        if (!s || !out) throw ArgError("null argument");
        ceos_server_stats st{};
        std::vector<double> v, w, sv;
        {  // copies under the lock; the sorting after it, so the worker is not held up
            std::lock_guard<std::mutex> g(s->mu);
            st.submitted = s->submitted;
            st.completed = s->completed;
            st.batches = s->batches;
            st.mean_batch = s->batches ? (double)s->batched_queries / (double)s->batches : 0.0;
            v = s->lat_ms.to_vector();
            w = s->wait_ms.to_vector();
            sv = s->service_ms.to_vector();
            if (!v.empty()) {
                st.elapsed_s = std::chrono::duration<double>(s->last_done - s->first_arrival).count();
                st.qps = st.elapsed_s > 0 ? (double)s->completed / st.elapsed_s : 0.0;
            }
            st.inflight_slots = (int)s->slots.size();
            st.device_bytes = s->device_bytes;
            st.inserts = s->n_upd[0];
            st.deletes = s->n_upd[1];
            st.insert_mean_ms = s->n_upd[0] ? s->upd_ms_sum[0] / (double)s->n_upd[0] : 0.0;
            st.insert_max_ms = s->upd_ms_max[0];
            st.delete_mean_ms = s->n_upd[1] ? s->upd_ms_sum[1] / (double)s->n_upd[1] : 0.0;
            st.delete_max_ms = s->upd_ms_max[1];
        }
        auto pcts = [](std::vector<double>& x, double& p50, double& p99, double& mx) {
            if (x.empty()) return;
            std::sort(x.begin(), x.end());
            auto at = [&](double p) { return x[std::min(x.size() - 1, (size_t)(p * (double)(x.size() - 1) + 0.5))]; };
            p50 = at(0.50);
            p99 = at(0.99);
            mx = x.back();
        };
        pcts(v, st.p50_ms, st.p99_ms, st.max_ms);
        pcts(w, st.wait_p50_ms, st.wait_p99_ms, st.wait_max_ms);
        pcts(sv, st.service_p50_ms, st.service_p99_ms, st.service_max_ms);
        *out = st;
    });
}

int ceos_server_get_batches(ceos_server* s, ceos_server_batch* out, int64_t max, int64_t* count) {
    return guarded([&] {
        if (!s || !count || (max > 0 && !out)) throw ArgError("null argument");
        std::lock_guard<std::mutex> g(s->mu);
        *count = (int64_t)s->trace.size();
        const int64_t n = std::min<int64_t>(std::max<int64_t>(max, 0), (int64_t)s->trace.size());
        if (n > 0) s->trace.copy_to(out, (size_t)n);
    });
}

int ceos_server_reset_stats(ceos_server* s) {
    return guarded([&] {
        if (!s) throw ArgError("null server");
        std::lock_guard<std::mutex> g(s->mu);
        s->submitted = s->completed = s->batches = s->batched_queries = 0;
        s->lat_ms.clear();
        s->wait_ms.clear();
        s->service_ms.clear();
        s->trace.clear();
        s->any_arrival = false;
        s->n_upd[0] = s->n_upd[1] = 0;
        s->upd_ms_sum[0] = s->upd_ms_sum[1] = s->upd_ms_max[0] = s->upd_ms_max[1] = 0.0;
    });
}

}  // extern "C"
