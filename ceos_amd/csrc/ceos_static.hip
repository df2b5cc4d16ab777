// coCEOs device search that allocates nothing (ceos_search_coceos_device_noalloc).
//
// The search's scratch (query projections, probe lists, per-workgroup candidate keys,
// bucket-position slots, stash lists and tile-id copies, the work counter, the status
// flag and the distinct-id estimate) lives in `scratch`, a file-scope device array of
// kStaticScratchBytes. Each GPU that loads the library holds its own copy until the
// process ends; its device address is looked up once per device with
// hipGetSymbolAddress. Every call on one GPU shares that GPU's array: two calls that
// overlap on one GPU race (calls ordered on one stream do not).
//
// To fit the array, the persistent grid is narrowed (any grid size gives the same
// results: each query is searched by one workgroup) and the queries are searched in
// chunks, one launch sequence per chunk, ordered on the caller's stream.
#include <mutex>

#include "ceos_core.h"

namespace {

constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint4 scratch[kStaticScratchBytes / sizeof(uint4)];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (looked up on the first call on a device).
unsigned char* scratch_base() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw DevError("device ordinal out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw DevError("hipGetSymbolAddress returned no address for the static search scratch");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" int ceos_search_coceos_device_noalloc(ceos_index* h, const float* dQ, int nq, int d, int k,
                                                 int32_t* d_ids, float* d_ips, void* stream) {
    return guarded([&] {
        if (!h) throw ArgError("null handle");
        std::lock_guard<std::mutex> g(h->core.mu);
        Core& c = h->core;
        check_matrix(dQ, nq, d, c.d);
        c.validate_query(c.s, c.mp, c.c, k, c.DR(), "n_proj * n_repeats");
        if (nq <= 0 || k == 0) return;
        if (!d_ids || !d_ips) throw ArgError("null output array");
        hipStream_t st = c.stream(stream);
        ceos::QueryConfig cfg = c.search_config_with(c.s, c.c, k, c.wg_per_cu, c.table_log2);
        // layout: header (work, status, estimate) | projections | probes | per-workgroup parts
        const size_t DR = (size_t)c.DR();
        const size_t per_q = (DR + (size_t)cfg.s_max) * 4;
        const size_t cand_b = al256((size_t)cfg.cand_stride * 8);
        const size_t tile_b = al256((size_t)ceos::search_tile_slots_nbk(cfg.tile_buckets) * 4);
        const size_t stash_b = (size_t)ceos::kSearchStashList * 16;
        const size_t per_wg = 2 * cand_b + stash_b + tile_b;
        const size_t avail = kStaticScratchBytes - 256 - 2 * 256;  // header and alignment of two arrays
        // at most a quarter of the array for per-query data, the rest for workgroups
        const int64_t nqc = std::min<int64_t>(nq, std::max<int64_t>(1, (int64_t)(avail / 4 / per_q)));
        const size_t used_q = al256((size_t)nqc * DR * 4) + al256((size_t)nqc * cfg.s_max * 4);
        const int64_t fit = used_q < avail ? (int64_t)((avail - used_q) / per_wg) : 0;
        if (fit < 1)
            throw MemError("the search needs " + std::to_string(per_wg + per_q + 768) +
                           " bytes of scratch for one workgroup and one query, above the " +
                           std::to_string(kStaticScratchBytes) +
                           " bytes of the static scratch (n_probed_vectors * n_probed_points too large)");
        cfg.grid = (int)std::min<int64_t>(cfg.grid, fit);
        unsigned char* base = scratch_base();
        ceos::QueryScratch sc{};
        sc.work = reinterpret_cast<int32_t*>(base);
        sc.status = reinterpret_cast<int32_t*>(base + 4);
        sc.est = reinterpret_cast<float*>(base + 8);
        size_t off = 256;
        sc.qproj = reinterpret_cast<float*>(base + off);
        off += al256((size_t)nqc * DR * 4);
        sc.probes = reinterpret_cast<int32_t*>(base + off);
        off += al256((size_t)nqc * cfg.s_max * 4);
        sc.cand = reinterpret_cast<uint64_t*>(base + off);
        off += (size_t)cfg.grid * cand_b;
        sc.pairs = reinterpret_cast<uint2*>(base + off);
        off += (size_t)cfg.grid * cand_b;
        sc.stash = reinterpret_cast<uint4*>(base + off);
        off += (size_t)cfg.grid * stash_b;
        sc.keys = reinterpret_cast<uint32_t*>(base + off);
        off += (size_t)cfg.grid * tile_b;
        if (off > kStaticScratchBytes) throw DevError("static scratch layout overflow");
        // the status flag starts clear and the distinct-id estimate at one half (fill
        // kernels: no memset, so a caller may capture this call); the work counter is
        // cleared by every chunk's launch sequence
        HIP_CHECK(ceos::launch_fill_i32(sc.status, 0, 1, st));
        HIP_CHECK(ceos::launch_fill_i32(reinterpret_cast<int32_t*>(sc.est), 0x3f000000, 1, st));
        for (int64_t q0 = 0; q0 < nq; q0 += nqc) {
            const int nb = (int)std::min<int64_t>(nqc, nq - q0);
            ceos::QueryBatch qb{dQ + q0 * c.d, nb, nullptr, nullptr, nullptr, k, c.s, c.c, c.mp, nullptr,
                                d_ids + q0 * k, d_ips + q0 * k, nullptr, 0};
            HIP_CHECK(ceos::launch_coceos_search(c.dims(), c.signs.p, c.bkt.p, c.bsize.p, c.rows(), qb, cfg, sc, st));
        }
    });
}
