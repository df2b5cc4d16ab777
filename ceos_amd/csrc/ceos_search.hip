// coCEOs search on gfx950: three launches per batch of queries.
//
// 1. k_project_rows (grid-wide): every query's D*R projections, row-major [nq][DR]
//    (sign + FWHT, 3 rounds per repeat; src/CEOs.cpp:495-513).
// 2. k_probe_select (one 256-thread block per query): exact top-s of |p| by radix
//    select, sorted to ascending |p| -- the reference's pop order -- and encoded as
//    list ids 2c (close, p >= 0) / 2c+1 (far, p < 0) (src/CEOs.cpp:517-548, 562-564).
// 3. k_search_fused (persistent, one 512-thread workgroup per query at a time, next
//    query from a global work counter so uneven k / s / c balance across the chip):
//    a. score accumulation in two phases over an LDS tile of 4 * tile_buckets hashed slots
//       plus a small overflow region.
//       Phase 1 gives every distinct id of the probed lists a slot (two-choice
//       buckets of four 4-byte ids, or one bucket per id when the pass is planned
//       at most 70% full; claims by compare-and-swap, with no barrier; a small
//       linear-probing stash for the rare leftovers) and records each bucket
//       position's slot. Phase 2 turns the tile into FP32 sums and adds the values
//       probe by probe, a barrier between probes, so each point's sum is formed in
//       probe order exactly as the reference's robin-map accumulation
//       (src/CEOs.cpp:558-610). Bucket entries are prefetched into registers a
//       chunk of probes at a time, double buffered. When the distinct ids exceed
//       the tile, ids are split into hash partitions done in successive passes;
//       the partition count adapts per workgroup and doubles on overflow;
//    b. exact top-c of the estimates by (score desc, id asc) (src/CEOs.cpp:619-632);
//       with fewer distinct candidates than c all are kept. In a single pass the
//       select reads the tile itself: a value histogram over equal-width bins finds
//       the boundary bin, which a radix select resolves exactly; with several passes
//       the (score, id) keys of all passes are selected from scratch;
//    c. rerank: FP32 inner products of the candidates' rows with the query, 16 rows
//       in flight per lane group (8 for the largest bucket depths) (src/CEOs.cpp:638-657).
//       With int8 row copies (k_quantize_rows) and k small against c, a bound pass
//       first reads every candidate's int8 row (a quarter of the bytes) and brackets
//       its FP32 inner product in [lo, hi]; a candidate whose hi is below the k-th
//       largest lo cannot reach the top-k, and only the others (typically a little
//       over k) get the exact FP32 rerank. The top-k is the one reranking every
//       candidate gives;
//    d. top-k by (inner product desc, id asc), written descending, zero padded
//       (src/CEOs.cpp:659-666).
#include <algorithm>
#include <cfloat>

#include "ceos_device.h"
#include "ceos_kernels.h"

namespace ceos {

constexpr int kSNT = kSearchThreads;   // 512
constexpr int kSWaves = kSNT / kWave;
constexpr int kPNT = 256;              // probe-select block
constexpr uint32_t kEmptyKey = 0xffffffffu;
constexpr int kBufEntries = 8;         // bucket entries per thread per chunk
constexpr float kOneChoiceLoad = 0.70f;  // planned load of a single-bucket pass (phase 1)
constexpr uint32_t kUnset = 0xffffffffu;       // sum bits of a slot before its first (ordered) add
typedef const float __attribute__((address_space(1)))* GlobalF;
typedef float F4v __attribute__((ext_vector_type(4)));
typedef const F4v __attribute__((address_space(1)))* GlobalF4;
typedef uint32_t U4v __attribute__((ext_vector_type(4)));
typedef const U4v __attribute__((address_space(1)))* GlobalU4;
typedef const uint32_t __attribute__((address_space(1)))* GlobalU1;
constexpr uint32_t kStashMark = 0xfffffffeu;   // phase-1 slot value: entry sent to the stash
constexpr uint32_t kToPlace = 0xfffffffdu;     // phase-1 slot value: entry not placed yet
constexpr int kEmitU = 8;                      // tile slots per thread per emission step
constexpr int kSelBins = 2048;                 // candidate select: value histogram bins
constexpr int kBoundMax = 512;                 // boundary-bin keys selected in LDS (else in scratch)
constexpr int kStashList = kSearchStashList;   // pending stash entries per pass (global list)
constexpr int kPruneU = 16;                    // bound pass: int8 rows in flight per lane group
constexpr int kSurvivorMax = kSelBins;         // bound pass: exact-rerank list capacity (hist region)

// Step J of the reduce-scatter of 16 per-lane (h, l) int32 pairs over groups of 2^LG
// lanes: the first min(LG, 4) steps exchange half of the pairs still carried (the lane
// whose offset bit is set keeps the upper half; across lane bits 5 and 4 a
// v_permlane32/16_swap of the pair (v[k], v[HALF + k]) leaves each lane its own half and
// the partner's copy of it, with no selects), the rest are plain butterflies. Afterwards
// h[t], l[t] (t < 16 >> min(LG, 4)) are the group totals of pair rb + t, where rb sums
// 8 >> j over the steps j whose offset bit is set in the lane's group index. Integer
// sums are exact, so the order of the exchanges does not matter.
template <int LG, int J>
__device__ __forceinline__ void reduce_scatter16i(int (&h)[16], int (&l)[16], int gl) {
    if constexpr (J < LG) {
        constexpr int O = (1 << LG) >> (J + 1);
        if constexpr (J < (LG < 4 ? LG : 4) && (O == 32 || O == 16)) {
            constexpr int HALF = 16 >> (J + 1);
#pragma unroll
            for (int k = 0; k < HALF; ++k) {
                const auto a = O == 32 ? __builtin_amdgcn_permlane32_swap((uint32_t)h[k], (uint32_t)h[HALF + k], false, false)
                                       : __builtin_amdgcn_permlane16_swap((uint32_t)h[k], (uint32_t)h[HALF + k], false, false);
                const auto b = O == 32 ? __builtin_amdgcn_permlane32_swap((uint32_t)l[k], (uint32_t)l[HALF + k], false, false)
                                       : __builtin_amdgcn_permlane16_swap((uint32_t)l[k], (uint32_t)l[HALF + k], false, false);
                h[k] = (int)(a[0] + a[1]);
                l[k] = (int)(b[0] + b[1]);
            }
        } else if constexpr (J < (LG < 4 ? LG : 4)) {
            constexpr int HALF = 16 >> (J + 1);
            const bool up = (gl & O) != 0;
#pragma unroll
            for (int k = 0; k < HALF; ++k) {
                const int kh = up ? h[HALF + k] : h[k], sh = up ? h[k] : h[HALF + k];
                const int kl = up ? l[HALF + k] : l[k], sl = up ? l[k] : l[HALF + k];
                h[k] = kh + __shfl_xor(sh, O, kWave);
                l[k] = kl + __shfl_xor(sl, O, kWave);
            }
        } else {
            h[0] += __shfl_xor(h[0], O, kWave);
            l[0] += __shfl_xor(l[0], O, kWave);
        }
        reduce_scatter16i<LG, J + 1>(h, l, gl);
    }
}

// v_dot4_i32_i8 with a zero accumulator (the builtin with 0 becomes a move and v_dot4c).
__device__ __forceinline__ int sdot4_first(uint32_t a, uint32_t b) {
    int r;
    asm("v_dot4_i32_i8 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// FP32 sum of four int8 lanes of w (byte j = element j) times q, added to acc.
__device__ __forceinline__ float dot_i8x4(uint32_t w, float4 q, float acc) {
    acc = fmaf((float)((int32_t)(w << 24) >> 24), q.x, acc);
    acc = fmaf((float)((int32_t)(w << 16) >> 24), q.y, acc);
    acc = fmaf((float)((int32_t)(w << 8) >> 24), q.z, acc);
    acc = fmaf((float)((int32_t)w >> 24), q.w, acc);
    return acc;
}

// One multiplicative hash per id (a single 32-bit multiply) serves both bucket choices
// (multiply-high of the hash, and of the hash rotated by 16 bits, by the bucket count nbk),
// the claim rotation (bits 4-5) and the pass partition (high half folded onto the low half).
__device__ __forceinline__ uint32_t idhash(uint32_t id) { return id * 0x9e3779b1u; }
__device__ __forceinline__ uint32_t bucket1(uint32_t h, uint32_t nbk) {
    return (uint32_t)(((uint64_t)h * nbk) >> 32);
}
__device__ __forceinline__ uint32_t bucket2(uint32_t h, uint32_t b1, uint32_t nbk) {
    const uint32_t b2 = (uint32_t)(((uint64_t)((h << 16) | (h >> 16)) * nbk) >> 32);
    return b2 != b1 ? b2 : (b1 + 1u == nbk ? 0u : b1 + 1u);
}
__device__ __forceinline__ uint32_t pass_of(uint32_t h, uint32_t P) {
    return (((h ^ (h >> 15)) & 0xffffu) * P) >> 16;
}

template <typename K>
static void allow_lds_s(K kernel, size_t bytes) {
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

__host__ __device__ inline int al16(int b) { return (b + 15) & ~15; }

struct SLayout {
    int q_off, probe_off, lim_off, misc_off, cid_off, hist_off, u_off;
    int tab_off, rk_off, tk_off, total;
};

__host__ __device__ inline SLayout s_layout(int d, int nbk, int s_max, int c_max, int k_max) {
    SLayout o;
    int off = 0;
    o.q_off = off; off += al16(d * 4);
    o.probe_off = off; off += al16(s_max * 4);
    o.lim_off = off; off += al16(s_max * 4);
    o.misc_off = off; off += 64;
    o.cid_off = off; off += al16(c_max * 4);
    o.hist_off = off; off += kSelBins * 4;
    o.u_off = off;
    o.tab_off = off;
    const int endA = off + 4 * search_tile_slots_nbk(nbk);
    o.rk_off = off;
    o.tk_off = off + al16(c_max * 8);
    const int endB = o.tk_off + next_pow2(k_max) * 8;
    o.total = endA > endB ? endA : endB;
    return o;
}

int search_lds_bytes(int d, int tile_buckets, int s_max, int c_max, int k_max) {
    return s_layout(d, tile_buckets, s_max, c_max, k_max).total;
}

// ---------------------------------------------------------------------------------
// STAGE: the query's DR projections are copied to LDS once (after the keys) and the radix
// passes read them there; without it (DR above kProbeStageMax) they read global memory.
// With DR >= 2048 and 16 s <= DR, a sampled bound comes first: T = the value at rank
// r = 2 ceil(s * 256 / DR) + 4 among the |p| of 256 strided columns (a radix select over
// those 256; T keeps the value bits the select resolved, so it is at most that value).
// One pass gathers every column with |p| >= T (NaN included) into a list of at most
// kProbeList keys, and the exact top-s is selected from the list: the columns left out
// all rank below every gathered one, so with at least s gathered the result is the exact
// top-s. Otherwise (a rare tail, or a list overflow) the select runs over all DR columns.
constexpr int kProbeStageMax = 12288;
constexpr int kProbeList = 640;  // with DR = 8192 staged, 4 workgroups still fit a CU
template <bool STAGE>
__global__ __launch_bounds__(kPNT) void k_probe_select(int DR, const float* __restrict__ qproj, int nq,
                                                       const int32_t* __restrict__ nq_dev,
                                                       const int32_t* __restrict__ s_arr, int s0, int s_stride,
                                                       int32_t* __restrict__ probes) {
    __shared__ SelectShared sh;
    // next_pow2(s_stride) keys (at least 2), the gathered list, then (STAGE) the DR projections
    extern __shared__ uint64_t pkey[];
    uint64_t* plist = pkey + max(2, next_pow2(s_stride));
    float* pl = reinterpret_cast<float*>(plist + kProbeList);  // 16-byte aligned
    if (nq_dev) nq = min(nq, nq_dev[0]);
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const int sq = min(max(s_arr ? s_arr[q] : s0, 0), min(s_stride, DR));
        if (sq == 0) continue;  // uniform
        const float* pq = qproj + (int64_t)q * DR;
        if (STAGE) {  // all of a thread's loads in flight together
            constexpr int U = 8;
            if ((DR & 3) == 0) {
                const int n4 = DR >> 2;
                for (int c0 = threadIdx.x; c0 < n4; c0 += U * kPNT) {
                    float4 t[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        t[u] = c0 + u * kPNT < n4 ? reinterpret_cast<const float4*>(pq)[c0 + u * kPNT]
                                                  : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (c0 + u * kPNT < n4) reinterpret_cast<float4*>(pl)[c0 + u * kPNT] = t[u];
                }
            } else {
                for (int c0 = threadIdx.x; c0 < DR; c0 += U * kPNT) {
                    float t[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) t[u] = c0 + u * kPNT < DR ? pq[c0 + u * kPNT] : 0.0f;
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (c0 + u * kPNT < DR) pl[c0 + u * kPNT] = t[u];
                }
            }
            __syncthreads();
        }
        const float* src = STAGE ? pl : pq;
        auto get = [&](int c) -> uint64_t { return ckey(fabsf(src[c]), (uint32_t)c); };
        int nlist = -1;  // keys gathered above the sampled bound; -1: select over all columns
        if (DR >= 8 * kPNT && 16 * sq <= DR) {  // uniform
            const int stride = DR / kPNT;
            const int r = 2 * ((sq * kPNT + DR - 1) / DR) + 4;
            uint64_t sp;
            int ssh;
            radix_select<kPNT>(kPNT, r, [&](int i) -> uint64_t { return get(i * stride); }, sh, sp, ssh);
            const float T = fkey_inv((uint32_t)(sp >> 32));
            if (threadIdx.x == 0) sh.counter = 0;
            __syncthreads();
            for (int c0 = 0; c0 < DR; c0 += kPNT) {
                const int c = c0 + (int)threadIdx.x;
                const bool take = c < DR && !(fabsf(src[c]) < T);
                const uint32_t pos = wave_append(take, &sh.counter);
                if (take && pos < (uint32_t)kProbeList) plist[pos] = get(c);
            }
            __syncthreads();
            const int cnt = (int)sh.counter;
            nlist = (cnt >= sq && cnt <= kProbeList) ? cnt : -1;
            __syncthreads();  // sh is reused below
        }
        const int N = next_pow2(sq);
        auto select_top = [&](int n, auto g) __attribute__((always_inline)) {
            uint64_t prefix;
            int shift;
            radix_select<kPNT>(n, sq, g, sh, prefix, shift);
            for (int i = threadIdx.x; i < N; i += kPNT) pkey[i] = 0;
            if (threadIdx.x == 0) sh.counter = 0;
            __syncthreads();
            compact_selected<kPNT>(n, g, prefix, shift, &sh.counter, [&](uint32_t pos, uint64_t k) { pkey[pos] = k; });
            __syncthreads();
        };
        if (nlist >= 0)
            select_top(nlist, [&](int i) -> uint64_t { return plist[i]; });
        else
            select_top(DR, get);
        bitonic_sort_desc<kPNT>(pkey, N);
        for (int j = threadIdx.x; j < sq; j += kPNT) {
            const int c = (int)ckey_id(pkey[sq - 1 - j]);
            probes[(int64_t)q * s_stride + j] = 2 * c + (src[c] < 0.0f ? 1 : 0);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <int EPT>
__global__ __launch_bounds__(kSNT) void k_search_fused(IndexDims dm, const int32_t* __restrict__ probes_g,
                                                       int s_stride, const Entry* __restrict__ bkt,
                                                       const int32_t* __restrict__ bsize, RowRing rows,
                                                       QueryBatch qb, QueryConfig cfg, QueryScratch scr) {
    constexpr int CH = kBufEntries / EPT;  // probes per prefetch chunk
    constexpr int kRerankU = EPT >= 8 ? 8 : 16;  // rows in flight per lane group (register budget)
    constexpr int B = kBufEntries;         // entries per thread per chunk
    __shared__ SelectShared sh;
    extern __shared__ __align__(16) unsigned char lds[];
    const SLayout lo = s_layout(dm.d, cfg.tile_buckets, cfg.s_max, cfg.c_max, cfg.k_max);
    float* qv = reinterpret_cast<float*>(lds + lo.q_off);
    int* probes = reinterpret_cast<int*>(lds + lo.probe_off);
    int* lims = reinterpret_cast<int*>(lds + lo.lim_off);
    int* misc = reinterpret_cast<int*>(lds + lo.misc_off);
    uint32_t* cidl = reinterpret_cast<uint32_t*>(lds + lo.cid_off);
    uint32_t* keys = reinterpret_cast<uint32_t*>(lds + lo.tab_off);  // phase 1: ids; phase 2: FP32 sums
    uint64_t* tile64 = reinterpret_cast<uint64_t*>(lds + lo.tab_off);
    uint64_t* rk = reinterpret_cast<uint64_t*>(lds + lo.rk_off);
    uint64_t* tk = reinterpret_cast<uint64_t*>(lds + lo.tk_off);
    const uint32_t nbk = (uint32_t)cfg.tile_buckets;  // hashed slots: nbk buckets of 4 keys (16 bytes)
    const int Tmain = 4 * (int)nbk;
    const int T = search_tile_slots_nbk((int)nbk);    // plus the overflow (stash) region at the end
    const int sslots = T - Tmain;
    uint64_t* cand = scr.cand + (int64_t)blockIdx.x * cfg.cand_stride;
    uint2* pairs = scr.pairs + (int64_t)blockIdx.x * cfg.cand_stride;
    uint32_t* keys_g = scr.keys + (int64_t)blockIdx.x * T;
    uint4* stash_list = scr.stash + (int64_t)blockIdx.x * kStashList;
    const int tid = threadIdx.x, wave = tid / kWave, ln = tid % kWave;
    // running estimate of distinct ids / entries: starts from the value the last launch on
    // this scratch left (small batches, as in serving, then size their passes from what
    // earlier queries saw) and is written back when the workgroup finishes
    float frac = scr.est ? fminf(1.0f, fmaxf(0.01f, scr.est[0])) : 0.5f;
    bool did_query = false;
    int nq_done = 0;  // queries this workgroup finished in this launch
    const int nq_run = qb.nq_dev ? min(qb.nq, qb.nq_dev[0]) : qb.nq;
    // CEOs-Hash: candidates are all distinct ids (no scores); their ids go to idlist
    // (the `pairs` scratch, unused in this mode) and their rerank keys to `cand`
    const bool hash_mode = qb.hash != 0;
    uint32_t* idlist = reinterpret_cast<uint32_t*>(pairs);

    for (;;) {
        if (tid == 0) misc[0] = atomicAdd(scr.work, 1);
        __syncthreads();
        const int q = misc[0];
        __syncthreads();
        if (q >= nq_run) break;
        did_query = true;
        // per-query parameters, clamped to the bounds the launch was sized for
        const int kq = min(max(qb.k ? qb.k[q] : qb.k0, 0), cfg.k_max);
        const int sq = min(max(qb.s ? qb.s[q] : qb.s0, 0), cfg.s_max);
        const int cq = min(max(qb.c ? qb.c[q] : qb.c0, 0), cfg.c_max);
        const float* qrow = qb.Q + (int64_t)q * dm.d;
        for (int i = tid; i < dm.d; i += kSNT) qv[i] = qrow[i];
        for (int j = tid; j < sq; j += kSNT) {
            const int l = probes_g[(int64_t)q * s_stride + j];
            probes[j] = l;
            lims[j] = min(qb.mp, bsize[l]);
        }
        __syncthreads();

        // ---- a. score accumulation ----
        // Phase 1 (any order): every id of the probed lists gets a tile slot; nothing is
        //   removed within a pass, so an id found once keeps its slot. Each bucket
        //   position's (slot, value) is kept in per-workgroup scratch.
        // Phase 2 (probe order, one barrier per probe): the ids move to scratch and the tile
        //   holds FP32 sums; every position adds its value to its slot in probe order, the
        //   first value assigned, which is the reference's per-id summation order
        //   (src/CEOs.cpp:576-606).
        if (tid < 64) {
            uint32_t e = 0;
            for (int j = tid; j < sq; j += 64) e += (uint32_t)lims[j];
            e = wave_sum(e);
            if (tid == 0) misc[1] = (int)e;
        }
        __syncthreads();
        const int E = misc[1];
        // hash partitions: the expected distinct ids of a pass fill at most 90% of the tile
        // (an overflow restarts the pass with twice the partitions). A workgroup's first
        // query in a launch plans for 25% more ids and 70% load: with one query per
        // workgroup (small batches) the launch takes as long as its slowest query, and a
        // restart would double it.
        const bool first = nq_done == 0;
        uint32_t P = (uint32_t)max(1, (int)ceilf((float)E * fminf(1.0f, frac * (first ? 1.25f : 1.0f)) /
                                                 ((first ? 0.7f : 0.9f) * (float)T)));
        const int mp = qb.mp;
        const int jlast = sq - 1;

        auto load_chunk = [&](Entry (&buf)[CH][EPT], int j0) {  // unconditional, clamped in bounds
#pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const Entry* bl = bkt + (int64_t)probes[min(j0 + jj, jlast)] * dm.m;
#pragma unroll
                for (int e = 0; e < EPT; ++e) buf[jj][e] = bl[min(tid + e * kSNT, dm.m - 1)];
            }
        };
        auto load_pairs = [&](uint2 (&buf)[CH][EPT], int j0) {
#pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const uint2* pl = pairs + (int64_t)min(j0 + jj, jlast) * mp;
#pragma unroll
                for (int e = 0; e < EPT; ++e) buf[jj][e] = pl[min(tid + e * kSNT, mp - 1)];
            }
        };
        // One chunk of phase 1, no barrier. Every entry reads its bucket (both buckets in
        // a two-bucket pass); an id already present takes its slot; otherwise it claims an
        // empty slot of the first bucket with room by compare-and-swap. A claim that finds
        // the same id (a copy of it claimed first) has its slot; losers, and ids whose
        // buckets are full, go to the stash list, whose pass looks for each id in its
        // buckets first. All reads of a chunk are issued before any of them is used.
        auto phase1_chunk = [&](const Entry (&buf)[CH][EPT], int j0, uint32_t pass, bool one) {
            uint32_t id[B], slot[B], b1[B], hh[B];
            float v[B];
            // the chunk's probe lengths (0 past the last probe), read once
            int limc[CH];
#pragma unroll
            for (int jj = 0; jj < CH; ++jj) limc[jj] = j0 + jj < sq ? lims[j0 + jj] : 0;
            uint32_t inpos = 0;  // bit b: entry b is a position of a probed list
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int t = tid + (b % EPT) * kSNT;
                id[b] = buf[b / EPT][b % EPT].x - rows.base;
                v[b] = __uint_as_float(buf[b / EPT][b % EPT].y);
                hh[b] = idhash(id[b]);
                const bool pos_ok = t < limc[b / EPT];
                inpos |= (pos_ok ? 1u : 0u) << b;
                const bool valid = pos_ok && (P == 1 || pass_of(hh[b], P) == pass);
                slot[b] = valid ? kToPlace : kEmptyKey;  // kEmptyKey: not in this pass
                b1[b] = bucket1(hh[b], nbk);
            }
            uint32_t claim = 0;
            if (one) {
                // One bucket per id (a pass planned well below the tile's capacity): an id
                // present in its bucket takes its slot, a new one claims an empty slot of
                // it; a full bucket sends the entry to the stash list.
                uint4 k1[B];
#pragma unroll
                for (int b = 0; b < B; ++b) k1[b] = reinterpret_cast<const uint4*>(keys)[b1[b]];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t x = id[b];
                    const uint4 p = k1[b];
                    const uint32_t m1 = (p.x == x ? 1u : 0u) | (p.y == x ? 2u : 0u) | (p.z == x ? 4u : 0u) |
                                        (p.w == x ? 8u : 0u);
                    const uint32_t e1 = (p.x >> 31) | ((p.y >> 30) & 2u) | ((p.z >> 29) & 4u) | ((p.w >> 28) & 8u);
                    const uint32_t r = (hh[b] >> 4) & 3u;
                    const uint32_t rot = ((e1 | (e1 << 4)) >> r) & 15u;
                    const uint32_t within = m1 != 0u ? (uint32_t)__builtin_ctz(m1 | 16u)
                                                     : (((uint32_t)__builtin_ctz(rot | 16u) + r) & 3u);
                    const uint32_t pos = 4u * b1[b] + within;
                    const bool hit = m1 != 0u;
                    const bool room = e1 != 0u;
                    const bool place = slot[b] == kToPlace;
                    slot[b] = place ? (hit || room ? pos : kStashMark) : slot[b];
                    claim |= (place && !hit && room ? 1u : 0u) << b;
                }
            } else {
                uint32_t b2[B];
                uint4 k1[B], k2[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    b2[b] = bucket2(hh[b], b1[b], nbk);
                    k1[b] = reinterpret_cast<const uint4*>(keys)[b1[b]];
                    k2[b] = reinterpret_cast<const uint4*>(keys)[b2[b]];
                }
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t x = id[b];
                    const uint4 p = k1[b], q2 = k2[b];
                    // one-hot match masks (ids within a bucket are distinct)
                    const uint32_t m1 = (p.x == x ? 1u : 0u) | (p.y == x ? 2u : 0u) | (p.z == x ? 4u : 0u) |
                                        (p.w == x ? 8u : 0u);
                    const uint32_t m2 = (q2.x == x ? 1u : 0u) | (q2.y == x ? 2u : 0u) | (q2.z == x ? 4u : 0u) |
                                        (q2.w == x ? 8u : 0u);
                    // empty masks: ids are below 2^31, the empty key has the top bit set
                    const uint32_t e1 = (p.x >> 31) | ((p.y >> 30) & 2u) | ((p.z >> 29) & 4u) | ((p.w >> 28) & 8u);
                    const uint32_t e2 = (q2.x >> 31) | ((q2.y >> 30) & 2u) | ((q2.z >> 29) & 4u) | ((q2.w >> 28) & 8u);
                    const bool in1 = m1 != 0u || (m2 == 0u && e1 != 0u);  // target bucket: b1 or b2
                    const uint32_t hm = in1 ? m1 : m2;
                    // a new id takes the first empty slot at or after a per-id start position
                    // (cyclically), so that different ids rarely claim the same slot
                    const uint32_t r = (hh[b] >> 4) & 3u;
                    const uint32_t em = in1 ? e1 : e2;
                    const uint32_t rot = ((em | (em << 4)) >> r) & 15u;
                    const uint32_t within = hm != 0u ? (uint32_t)__builtin_ctz(hm | 16u)
                                                     : (((uint32_t)__builtin_ctz(rot | 16u) + r) & 3u);
                    const uint32_t pos = 4u * (in1 ? b1[b] : b2[b]) + within;
                    const bool hit = (m1 | m2) != 0u;
                    const bool room = (e1 | e2) != 0u;
                    const bool place = slot[b] == kToPlace;
                    slot[b] = place ? (hit || room ? pos : kStashMark) : slot[b];
                    claim |= (place && !hit && room ? 1u : 0u) << b;
                }
            }
            {
                // Claims by compare-and-swap, with no barrier: a claim that finds the same
                // id has its slot; one that finds another id goes to the stash. Slots only
                // fill within a pass, so an id never gets two bucket slots (with two
                // buckets, an id goes to its second only while its first is full), and
                // the stash pass finds in the buckets any id whose copy won a slot there.
                uint32_t got[B];
#pragma unroll
                for (int b = 0; b < B; ++b)
                    got[b] = ((claim >> b) & 1u) ? atomicCAS(&keys[slot[b]], kEmptyKey, id[b]) : kEmptyKey;
#pragma unroll
                for (int b = 0; b < B; ++b)
                    if (got[b] != kEmptyKey && got[b] != id[b]) slot[b] = kStashMark;  // taken by another id
            }
            // the chunk's stash entries: one append per wave (a returning LDS atomic per
            // entry would serialise up to B round trips per wave)
            uint32_t smask = 0;
#pragma unroll
            for (int b = 0; b < B; ++b) smask |= (slot[b] == kStashMark ? 1u : 0u) << b;
            static_assert(B < 16, "stash entries per thread fit 4 bits");
            uint32_t sp = wave_append_n<4>((uint32_t)__builtin_popcount(smask), reinterpret_cast<uint32_t*>(&misc[5]));
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = j0 + b / EPT, t = tid + (b % EPT) * kSNT;
                if ((smask >> b) & 1u) {
                    if (sp < (uint32_t)kStashList)
                        stash_list[sp] = make_uint4(id[b], __float_as_uint(v[b]), (uint32_t)(j * mp + t), 0u);
                    else
                        misc[4] = 1;  // stash list full: the pass restarts with more partitions
                    ++sp;
                }
                if (!hash_mode && ((inpos >> b) & 1u))
                    pairs[(int64_t)j * mp + t] = make_uint2(slot[b] == kStashMark ? kEmptyKey : slot[b],
                                                            __float_as_uint(v[b]));
            }
        };
        // Stash: linear probing with CAS over the last sslots keys; every entry first looks
        // for its id in its two main buckets (the main tile no longer changes).
        auto stash_insert = [&]() {
            const int ns = min(misc[5], kStashList);
            for (int i = tid; i < ns; i += kSNT) {
                const uint4 e = stash_list[i];
                const uint32_t idv = e.x;
                const uint32_t hv = idhash(idv);
                const uint32_t b1 = bucket1(hv, nbk), b2 = bucket2(hv, b1, nbk);
                const uint4 p = reinterpret_cast<const uint4*>(keys)[b1];
                const uint4 q2 = reinterpret_cast<const uint4*>(keys)[b2];
                const uint32_t m1 = (p.x == idv ? 1u : 0u) | (p.y == idv ? 2u : 0u) | (p.z == idv ? 4u : 0u) |
                                    (p.w == idv ? 8u : 0u);
                const uint32_t m2 = (q2.x == idv ? 1u : 0u) | (q2.y == idv ? 2u : 0u) | (q2.z == idv ? 4u : 0u) |
                                    (q2.w == idv ? 8u : 0u);
                uint32_t found = m1 ? 4u * b1 + (uint32_t)__builtin_ctz(m1)
                                    : m2 ? 4u * b2 + (uint32_t)__builtin_ctz(m2) : kEmptyKey;
                if (found == kEmptyKey) {
                    uint32_t h = (hv ^ (hv >> 13)) & (uint32_t)(sslots - 1);
                    for (int probe = 0; probe < sslots; ++probe) {
                        const uint32_t si = (uint32_t)(T - sslots) + h;
                        const uint32_t k = keys[si];
                        if (k == idv) { found = si; break; }
                        if (k == kEmptyKey) {
                            const uint32_t o = atomicCAS(&keys[si], kEmptyKey, idv);
                            if (o == kEmptyKey || o == idv) { found = si; break; }
                            continue;  // taken by another id: look at the same slot again
                        }
                        h = (h + 1) & (uint32_t)(sslots - 1);
                    }
                    if (found == kEmptyKey) misc[4] = 1;  // stash full
                }
                if (found != kEmptyKey && !hash_mode) pairs[e.z].x = found;
            }
        };
        uint32_t kmin = 0xffffffffu, kmax = 0u;  // fkey range of every sum written in phase 2
        auto phase2_chunk = [&](const uint2 (&buf)[CH][EPT], int j0) {
#pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const int j = j0 + jj;
                if (j >= sq) continue;  // uniform across the block
                const int lim = lims[j];
                uint32_t cur[EPT];
#pragma unroll
                for (int e = 0; e < EPT; ++e) {
                    const int t = tid + e * kSNT;
                    const bool ok = t < lim && buf[jj][e].x != kEmptyKey;
                    cur[e] = keys[ok ? buf[jj][e].x : 0u];  // unconditional read, in bounds
                }
#pragma unroll
                for (int e = 0; e < EPT; ++e) {
                    const int t = tid + e * kSNT;
                    if (t < lim && buf[jj][e].x != kEmptyKey) {
                        const float v = __uint_as_float(buf[jj][e].y);
                        const float nv = cur[e] == kUnset ? v : __uint_as_float(cur[e]) + v;
                        keys[buf[jj][e].x] = __float_as_uint(nv);
                        const uint32_t fk = fkey(nv);
                        kmin = min(kmin, fk);
                        kmax = max(kmax, fk);
                    }
                }
                __syncthreads();
            }
        };

        // Tile sweeps, kEmitU = 8 slots per thread per step of 4096: the quads tid and
        // tid + kSNT of the step, read 16 bytes at a time (T % 4 == 0); item u of a thread
        // is slot tile_item(i0, u), slots past T read as `fill`.
        static_assert(kEmitU == 8, "a sweep step reads two quads per thread");
        auto tile_item = [&](int i0, int u) -> uint32_t {
            return (uint32_t)(i0 + 4 * (tid + (u >> 2) * kSNT) + (u & 3));
        };
        auto tile_load8 = [&](const uint32_t* arr, int i0, uint32_t fill, uint32_t (&v)[kEmitU]) __attribute__((always_inline)) {
            const int q0 = (i0 >> 2) + tid, q1 = q0 + kSNT, nq4 = T >> 2;
            const uint4 f4 = make_uint4(fill, fill, fill, fill);
            const uint4 a = q0 < nq4 ? reinterpret_cast<const uint4*>(arr)[q0] : f4;
            const uint4 b = q1 < nq4 ? reinterpret_cast<const uint4*>(arr)[q1] : f4;
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        };

        int ncand = 0;
        // expected distinct ids, as the partitions were planned
        const float exp_ids = (float)E * fminf(1.0f, frac * (first ? 1.25f : 1.0f));
        for (; sq > 0;) {
            // a single pass expected to fill at most kOneChoiceLoad of the main slots uses
            // one bucket per id: ids over a full bucket are few there, and the stash takes them
            const bool one = P == 1 && exp_ids <= kOneChoiceLoad * (float)Tmain;
            if (tid == 0) { misc[2] = 0; misc[4] = 0; }
            bool overflow = false;
            for (uint32_t pass = 0; pass < P && !overflow; ++pass) {
                for (int i = tid; i < (T >> 2); i += kSNT)  // T % 4 == 0: 16-byte stores
                    reinterpret_cast<uint4*>(keys)[i] = make_uint4(kEmptyKey, kEmptyKey, kEmptyKey, kEmptyKey);
                if (tid == 0) {
                    misc[5] = 0;
                    misc[6] = 0;
                    misc[7] = -1;  // min fkey of the phase-2 sums (as unsigned)
                    misc[8] = 0;   // max fkey
                }
                __syncthreads();
                // phase 1
                {
                    Entry bufA[CH][EPT], bufB[CH][EPT];
                    load_chunk(bufA, 0);
                    for (int j0 = 0; j0 < sq; j0 += 2 * CH) {
                        load_chunk(bufB, j0 + CH);
                        phase1_chunk(bufA, j0, pass, one);
                        if (j0 + CH >= sq) break;
                        load_chunk(bufA, j0 + 2 * CH);
                        phase1_chunk(bufB, j0 + CH, pass, one);
                    }
                }
                __syncthreads();
                if (misc[5] > 0) {  // uniform
                    stash_insert();
                    __syncthreads();
                }
                if (misc[4] != 0) { overflow = true; break; }
                if (hash_mode) {  // every id of this pass is a candidate (src/CEOs.cpp:1005-1026)
                    for (int i0 = 0; i0 < T; i0 += kEmitU * kSNT) {  // uniform trip count (wave-level append)
                        uint32_t kv[kEmitU], cnt = 0;
                        tile_load8(keys, i0, kEmptyKey, kv);
#pragma unroll
                        for (int u = 0; u < kEmitU; ++u) cnt += kv[u] != kEmptyKey ? 1u : 0u;
                        uint32_t pos = wave_append_n<4>(cnt, reinterpret_cast<uint32_t*>(&misc[2]));
#pragma unroll
                        for (int u = 0; u < kEmitU; ++u)
                            if (kv[u] != kEmptyKey) idlist[pos++] = kv[u];
                    }
                    __syncthreads();
                    continue;
                }
                // ids to scratch; the tile becomes the sums ("unset" until the first value)
                {
                    uint32_t occ = 0;
                    for (int i = tid; i < (T >> 2); i += kSNT) {  // 16 bytes at a time
                        const uint4 kv = reinterpret_cast<const uint4*>(keys)[i];
                        reinterpret_cast<uint4*>(keys_g)[i] = kv;
                        reinterpret_cast<uint4*>(keys)[i] = make_uint4(kUnset, kUnset, kUnset, kUnset);
                        occ += (kv.x != kEmptyKey ? 1u : 0u) + (kv.y != kEmptyKey ? 1u : 0u) +
                               (kv.z != kEmptyKey ? 1u : 0u) + (kv.w != kEmptyKey ? 1u : 0u);
                    }
                    occ = wave_sum(occ);
                    if (ln == 0) atomicAdd(reinterpret_cast<uint32_t*>(&misc[6]), occ);
                }
                __syncthreads();
                // phase 2
                {
                    uint2 bufA[CH][EPT], bufB[CH][EPT];
                    load_pairs(bufA, 0);
                    for (int j0 = 0; j0 < sq; j0 += 2 * CH) {
                        load_pairs(bufB, j0 + CH);
                        phase2_chunk(bufA, j0);
                        if (j0 + CH >= sq) break;
                        load_pairs(bufA, j0 + 2 * CH);
                        phase2_chunk(bufB, j0 + CH);
                    }
                }
                if (P == 1) {  // single pass: the candidates are selected from the tile itself
                    // every final sum is one of the sums written, so [kmin, kmax] covers them
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        kmin = min(kmin, (uint32_t)__shfl_xor(kmin, o, kWave));
                        kmax = max(kmax, (uint32_t)__shfl_xor(kmax, o, kWave));
                    }
                    if (ln == 0) {
                        atomicMin(reinterpret_cast<uint32_t*>(&misc[7]), kmin);
                        atomicMax(reinterpret_cast<uint32_t*>(&misc[8]), kmax);
                    }
                    break;
                }
                // candidates of this pass: (sum, id) keys to scratch
                for (int i0 = 0; i0 < T; i0 += kEmitU * kSNT) {  // uniform trip count (wave-level append)
                    uint32_t idv[kEmitU], sv[kEmitU];
                    uint32_t cnt = 0;
                    tile_load8(keys_g, i0, kEmptyKey, idv);
                    tile_load8(keys, i0, kUnset, sv);
#pragma unroll
                    for (int u = 0; u < kEmitU; ++u) cnt += idv[u] != kEmptyKey ? 1u : 0u;
                    uint32_t pos = wave_append_n<4>(cnt, reinterpret_cast<uint32_t*>(&misc[2]));
#pragma unroll
                    for (int u = 0; u < kEmitU; ++u)
                        if (idv[u] != kEmptyKey) cand[pos++] = ckey(__uint_as_float(sv[u]), idv[u]);
                }
                __syncthreads();
            }
            if (!overflow) {
                ncand = (P == 1 && !hash_mode) ? misc[6] : misc[2];
                break;
            }
            __syncthreads();
            if (P >= 1024) {  // unreachable for valid inputs; report and drop the query
                if (tid == 0) atomicOr(scr.status, 1);
                ncand = 0;
                break;
            }
            P *= 2;
        }
        if (E > 0) frac = 0.75f * frac + 0.25f * ((float)ncand / (float)E);

        // ---- b. candidates: exact top-c ----
        const int ce = hash_mode ? ncand : min(cq, ncand);
        // Exact top-ce of nk (sum, id) keys in scratch (cand) by key: the ids go to cidl. The
        // keys are moved to the (no longer needed) tile first when they fit.
        auto select_from_cand = [&](int nk, bool in_tile) __attribute__((always_inline)) {
            const bool lds_keys = in_tile || 2 * nk <= T;  // 8-byte keys fit the tile
            if (lds_keys && !in_tile) {
                for (int i = tid; i < nk; i += kSNT) tile64[i] = cand[i];
                __syncthreads();
            }
            auto msel = [&](auto get) __attribute__((always_inline)) {
                uint64_t prefix;
                int shift;
                radix_select<kSNT>(nk, ce, get, sh, prefix, shift);
                if (tid == 0) sh.counter = 0;
                __syncthreads();
                compact_selected<kSNT>(nk, get, prefix, shift, &sh.counter,
                                       [&](uint32_t pos, uint64_t k) { cidl[pos] = ckey_id(k); });
            };
            if (lds_keys)
                msel([&](int i) -> uint64_t { return tile64[i]; });
            else
                msel([&](int i) -> uint64_t { return cand[i]; });
            __syncthreads();
        };
        // Single pass with few candidates taken (ce well below the distinct ids): a lower
        // bound L of the ce-th largest sum is read off a strided sample of kSNT tile slots
        // (the sample value at 1.5 times the expected rank, plus 4); every slot whose
        // sum is at least L goes to scratch as a (sum, id) key in one pass over the tile.
        // When at least ce keys pass, the top-ce keys are all among them and are selected
        // exactly; otherwise (a rare event) the histogram select below runs. The tile pass
        // writes (sum key, slot) pairs only; the ids are then fetched for the pairs kept, all
        // loads in flight together, and the keys go to the tile (when they fit) or stay in
        // scratch. Returns the number of keys, or -1.
        auto sampled_compaction = [&]() __attribute__((always_inline)) -> int {
            // 2 * kSNT strided sample keys (the hist region holds 1024 keys)
            constexpr int NS = 2 * kSNT;
            static_assert(NS * 8 <= kSelBins * 4, "sample keys fit the histogram region");
            uint64_t* skey = reinterpret_cast<uint64_t*>(lds + lo.hist_off);
            const int stepS = T / NS;
            const uint32_t sb0 = keys[tid * stepS], sb1 = keys[(tid + kSNT) * stepS];
            skey[tid] = sb0 != kUnset ? (uint64_t)fkey(__uint_as_float(sb0)) << 32 : 0ull;
            skey[tid + kSNT] = sb1 != kUnset ? (uint64_t)fkey(__uint_as_float(sb1)) << 32 : 0ull;
            if (tid == 0) { misc[14] = 0; misc[15] = 0; }
            __syncthreads();
            const uint32_t occ = (uint32_t)__popcll(__ballot(sb0 != kUnset)) + (uint32_t)__popcll(__ballot(sb1 != kUnset));
            if (ln == 0) atomicAdd(reinterpret_cast<uint32_t*>(&misc[14]), occ);
            __syncthreads();
            const int nso = misc[14];
            const int rs = (int)((3ll * ce * nso + 2ll * ncand - 1) / (2ll * max(ncand, 1))) + 4;
            if (2 * rs > nso) return -1;  // uniform: ce is a large share of the ids
            uint64_t prefix;
            int shift;
            radix_select<kSNT>(NS, rs, [&](int i) -> uint64_t { return skey[i]; }, sh, prefix, shift, 32);
            const uint32_t lkey = (uint32_t)(prefix >> 32);
            // the bound as a float: x >= Lf keeps every sum whose key reaches lkey (and -0
            // against a +0 bound, a superset that the exact select below does not mind);
            // kUnset is a NaN and compares false. A bound at or below the key of -inf keeps
            // every sum.
            const float Lf = lkey <= fkey(-INFINITY) ? -INFINITY : fkey_inv(lkey);
            for (int i0 = 0; i0 < T; i0 += kEmitU * kSNT) {  // uniform trip count (wave-level append)
                uint32_t tm = 0, sb[kEmitU];
                tile_load8(keys, i0, kUnset, sb);
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) tm |= (__uint_as_float(sb[u]) >= Lf ? 1u : 0u) << u;
                uint32_t pos = wave_append_n<4>((uint32_t)__popc(tm), reinterpret_cast<uint32_t*>(&misc[15]));
#pragma unroll
                for (int u = 0; u < kEmitU; ++u)
                    if ((tm >> u) & 1u)
                        cand[pos++] = ((uint64_t)fkey(__uint_as_float(sb[u])) << 32) | tile_item(i0, u);
            }
            __syncthreads();
            const int nk = misc[15];
            if (nk < ce) return -1;  // uniform
            // (sum key, slot) -> (sum key, ~id); into the tile when the keys fit it
            uint64_t* dst = 2 * nk <= T ? tile64 : cand;
            for (int i0 = 0; i0 < nk; i0 += kEmitU * kSNT) {
                uint64_t e[kEmitU];
                uint32_t idv[kEmitU];
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) {
                    const int i = i0 + u * kSNT + tid;
                    e[u] = i < nk ? cand[i] : 0ull;
                }
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) {
                    const int i = i0 + u * kSNT + tid;
                    idv[u] = i < nk ? keys_g[(uint32_t)e[u]] : 0u;
                }
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) {
                    const int i = i0 + u * kSNT + tid;
                    if (i < nk) dst[i] = (e[u] & 0xffffffff00000000ull) | (uint64_t)(~idv[u]);
                }
            }
            __syncthreads();
            return nk;
        };
        int nsampled = -1;
        if (!hash_mode && ce > 0 && P == 1 && ce < ncand) nsampled = sampled_compaction();
        if (hash_mode) {
            // no selection: every candidate is reranked
        } else if (nsampled >= 0) {
            select_from_cand(nsampled, 2 * nsampled <= T);
        } else if (ce > 0 && P == 1) {
            // Single pass: the tile holds each occupied slot's sum (kUnset: empty) and
            // keys_g its id. Keys are (sum desc, id asc). A histogram of the sums over
            // kSelBins equal-width bins between the smallest and largest sum (a monotone
            // map of the value) finds the bin holding the ce-th largest: every slot above
            // it is taken; the boundary bin is resolved exactly by a radix select on its
            // full keys. cidl first gets slot indices, then ids.
            uint32_t* hist = reinterpret_cast<uint32_t*>(lds + lo.hist_off);
            uint64_t* bkeys = reinterpret_cast<uint64_t*>(lds + lo.hist_off);             // kBoundMax keys
            uint32_t* bslots = reinterpret_cast<uint32_t*>(lds + lo.hist_off) + 2 * kBoundMax;  // after them
            constexpr int BPT = kSelBins / kSNT;  // bins per thread in the boundary scan
            if (tid == 0) {
                misc[12] = 0;  // boundary-bin count (appends)
                misc[13] = 0;  // boundary-bin appends to scratch (crowded bin)
            }
            for (int i = tid; i < kSelBins; i += kSNT) hist[i] = 0;
            __syncthreads();
            const float vlo = fkey_inv((uint32_t)misc[7]);
            const float range = fkey_inv((uint32_t)misc[8]) - vlo;
            const float scale = (range > 0.0f && range <= 3.0e38f) ? (float)kSelBins / range : 0.0f;
            auto vbin = [&](uint32_t sb) -> int {  // monotone in the value
                if (scale == 0.0f) return 0;
                const float t = (__uint_as_float(sb) - vlo) * scale;
                return t >= (float)(kSelBins - 1) ? kSelBins - 1 : (int)t;
            };
            for (int q4 = tid; q4 < (T >> 2); q4 += kSNT) {  // 16 bytes at a time
                const uint4 sb = reinterpret_cast<const uint4*>(keys)[q4];
                if (sb.x != kUnset) atomicAdd(&hist[vbin(sb.x)], 1u);
                if (sb.y != kUnset) atomicAdd(&hist[vbin(sb.y)], 1u);
                if (sb.z != kUnset) atomicAdd(&hist[vbin(sb.z)], 1u);
                if (sb.w != kUnset) atomicAdd(&hist[vbin(sb.w)], 1u);
            }
            __syncthreads();
            {  // boundary bin: thread t scans bins [kSelBins - BPT*(t+1), kSelBins - BPT*t) downwards
                uint32_t c[BPT], tot = 0;
#pragma unroll
                for (int j = 0; j < BPT; ++j) {
                    c[j] = hist[kSelBins - 1 - BPT * tid - j];
                    tot += c[j];
                }
                uint32_t ex = wave_excl_scan(tot, ln);
                if (ln == kWave - 1) sh.hist[wave] = ex + tot;
                __syncthreads();
                for (int w = 0; w < wave; ++w) ex += sh.hist[w];
                const uint32_t need = (uint32_t)ce;
                if (ex < need && need <= ex + tot) {
                    for (int j = 0; j < BPT; ++j) {
                        if (need <= ex + c[j]) {
                            misc[9] = kSelBins - 1 - BPT * tid - j;  // boundary bin
                            misc[10] = (int)ex;                      // slots above it
                            break;
                        }
                        ex += c[j];
                    }
                }
                if (tid == 0) sh.counter = 0;
            }
            __syncthreads();
            const int bbin = misc[9], above = misc[10];
            // emission: slots above the boundary bin to cidl, boundary-bin slots to bslots
            for (int i0 = 0; i0 < T; i0 += kEmitU * kSNT) {  // uniform trip count (wave-level append)
                uint32_t am = 0, bm = 0, ca = 0, cb = 0, sbv[kEmitU];
                tile_load8(keys, i0, kUnset, sbv);
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) {
                    const uint32_t sb = sbv[u];
                    const int bn = sb != kUnset ? vbin(sb) : -1;
                    am |= (bn > bbin ? 1u : 0u) << u;
                    bm |= (bn == bbin ? 1u : 0u) << u;
                }
                ca = (uint32_t)__popc(am);
                cb = (uint32_t)__popc(bm);
                uint32_t pa = wave_append_n<4>(ca, &sh.counter);
                uint32_t pb = wave_append_n<4>(cb, reinterpret_cast<uint32_t*>(&misc[12]));
#pragma unroll
                for (int u = 0; u < kEmitU; ++u) {
                    const uint32_t i = tile_item(i0, u);
                    if ((am >> u) & 1u) cidl[pa++] = i;
                    if ((bm >> u) & 1u) {
                        if (pb < (uint32_t)kBoundMax) bslots[pb] = i;
                        ++pb;
                    }
                }
            }
            __syncthreads();
            const int nb = misc[12];
            const int needb = ce - above;
            if (nb <= kBoundMax) {
                for (int i = tid; i < nb; i += kSNT) {
                    const uint32_t sl = bslots[i];
                    bkeys[i] = ckey(__uint_as_float(keys[sl]), keys_g[sl]);
                }
            } else {  // rare: a crowded boundary bin goes through scratch
                for (int i0 = 0; i0 < T; i0 += kSNT) {
                    const int i = i0 + tid;
                    const uint32_t sb = i < T ? keys[i] : kUnset;
                    const bool tk2 = sb != kUnset && vbin(sb) == bbin;
                    const uint32_t pos = wave_append(tk2, reinterpret_cast<uint32_t*>(&misc[13]));
                    if (tk2) cand[pos] = ckey(__uint_as_float(sb), keys_g[i]);
                }
            }
            // ids of the slots above the boundary (one batch of gathers)
            for (int p2 = tid; p2 < above; p2 += kSNT) cidl[p2] = keys_g[cidl[p2]];
            __syncthreads();
            {
                // one instantiation per key array, so each reads one address space
                auto bsel = [&](auto get) __attribute__((always_inline)) {
                    uint64_t prefix;
                    int shift;
                    radix_select<kSNT>(nb, needb, get, sh, prefix, shift);
                    if (tid == 0) sh.counter = (uint32_t)above;
                    __syncthreads();
                    compact_selected<kSNT>(nb, get, prefix, shift, &sh.counter,
                                           [&](uint32_t pos, uint64_t k) { cidl[pos] = ckey_id(k); });
                };
                if (nb <= kBoundMax)
                    bsel([&](int i) -> uint64_t { return bkeys[i]; });
                else
                    bsel([&](int i) -> uint64_t { return cand[i]; });
            }
            __syncthreads();
        } else if (ce > 0) {
            select_from_cand(ncand, false);
        }

        // ---- c. rerank ----
        // instantiated once per mode: ids and keys are in LDS (cidl, rk) or in global
        // scratch (idlist, cand), and each copy reads one address space
        // RU rows in flight per lane group; the vector path issues UNR = 16 / RU of a
        // lane's float4 loads per row together. A row's sum is formed the same way for
        // every RU: lane gl of its G-lane group adds units gl, gl + G, ... in order, then
        // the group's butterfly.
        auto rerank = [&](const uint32_t* idsrc, uint64_t* keydst, const int nc, auto ruc) __attribute__((always_inline)) {
            constexpr int RU = decltype(ruc)::value;
            constexpr int UNR = RU >= 16 ? 1 : 16 / RU;
            const int d = dm.d;
            const bool vec = (d & 3) == 0;
            const int units = vec ? (d >> 2) : d;
            int G = 1;
            while (2 * G < units && G < kWave) G <<= 1;
            const int groups = kWave / G;
            const int sub = ln / G, gl = ln % G;
            const int per_iter = kSWaves * groups * RU;
            for (int base = 0; base < nc; base += per_iter) {
                float acc[RU];
                GlobalF rp[RU];  // rows are in global memory: plain global loads, never flat
                uint32_t ids[RU];
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    const int i = base + ((u * kSWaves + wave) * groups + sub);
                    acc[u] = 0.0f;
                    rp[u] = nullptr;
                    ids[u] = 0;
                    if (i < nc) {
                        ids[u] = idsrc[i];
                        int64_t slot = (int64_t)ids[u] + rows.head;
                        if (slot >= rows.cap) slot -= rows.cap;
                        rp[u] = (GlobalF)(rows.X + slot * d);
                    }
                }
                if (vec) {
                    const float4* q4 = reinterpret_cast<const float4*>(qv);
                    for (int uu0 = gl; uu0 < units; uu0 += UNR * G) {
                        float4 a[UNR][RU];
#pragma unroll
                        for (int v = 0; v < UNR; ++v) {
                            const int uu = uu0 + v * G;
#pragma unroll
                            for (int u = 0; u < RU; ++u) {
                                const F4v t = (rp[u] && uu < units) ? reinterpret_cast<GlobalF4>(rp[u])[uu]
                                                                    : F4v{0.0f, 0.0f, 0.0f, 0.0f};
                                a[v][u] = make_float4(t.x, t.y, t.z, t.w);
                            }
                        }
#pragma unroll
                        for (int v = 0; v < UNR; ++v) {
                            const int uu = uu0 + v * G;
                            if (uu < units) {
                                const float4 b = q4[uu];
#pragma unroll
                                for (int u = 0; u < RU; ++u) {
                                    acc[u] = fmaf(a[v][u].x, b.x, acc[u]);
                                    acc[u] = fmaf(a[v][u].y, b.y, acc[u]);
                                    acc[u] = fmaf(a[v][u].z, b.z, acc[u]);
                                    acc[u] = fmaf(a[v][u].w, b.w, acc[u]);
                                }
                            }
                        }
                    }
                } else {
                    for (int uu = gl; uu < d; uu += G) {
                        float a[RU];
#pragma unroll
                        for (int u = 0; u < RU; ++u) a[u] = rp[u] ? rp[u][uu] : 0.0f;
                        const float b = qv[uu];
#pragma unroll
                        for (int u = 0; u < RU; ++u) acc[u] = fmaf(a[u], b, acc[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    float v = acc[u];
                    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
                    const int i = base + ((u * kSWaves + wave) * groups + sub);
                    if (i < nc && gl == 0) keydst[i] = ckey(v, ids[u]);
                }
            }
            __syncthreads();
        };
        // Bound pass over the int8 rows: keydst[i] (rk, or cand in hash mode) becomes the
        // float pair (lo, hi) around candidate i's FP32 inner product v. A non-finite query
        // or row gives infinite bounds. Per row, k_quantize_rows stores x ~ s c
        // (c int8) with r1 = |x - s c| + g |x| and r2 = s |c| (g = (d + 64) 2^-24, the
        // rounding of any FP32 sum of depth below d + 64).
        //  * Integer path (d % 16 == 0, 64 <= d <= 1024; one 16-byte unit per lane): the
        //    query is coded as q^ = t1 b1 + t2 b2 (int8 codes of q and of its residual);
        //    I1, I2 = exact int32 dot products of the codes with c (v_dot4 per lane, then an
        //    integer reduce-scatter; |I| <= 127^2 d < 2^24, so each converts exactly) and
        //    a = s * fl(t1 I1 + t2 I2). |v - a| <= |q| r1 + (|q - q^| + 12u (t1|b1| +
        //    t2|b2|)) r2, u = 2^-24 (the two roundings of t1 I1 + t2 I2 take 2u of the 12u).
        //  * Float path (other d): a = s * fl(q . c), |v - a| <= |q| (r1 + g r2).
        // The margin added to the bound covers the rounding of the bound, lo and hi.
        auto bound_pass = [&](const uint32_t* idsrc, uint64_t* keydst) __attribute__((always_inline)) {
            constexpr int U = kPruneU;
            const int d = dm.d;
            const bool v16 = (d & 15) == 0;
            const int units = v16 ? (d >> 4) : (d >> 2);  // 16 or 4 int8 per lane load
            const bool idot = v16 && units >= 4 && units <= kWave;  // one 16-byte unit per lane
            int G = 1;
            if (idot) {
                while (G < units) G <<= 1;
            } else {
                while (2 * G < units && G < kWave) G <<= 1;
            }
            const int groups = kWave / G;
            const int sub = ln / G, gl = ln % G;
            const int per_iter = kSWaves * groups * U;
            // |q|_2 rounded up: the FP32 sum of squares is within 2^-12 for d <= 4096
            float q2 = 0.0f;
            for (int i = ln; i < d; i += kWave) q2 = fmaf(qv[i], qv[i], q2);
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) q2 += __shfl_xor(q2, o, kWave);
            const float qn = sqrtf(q2) * (1.0f + 0x1p-10f);
            float2* lh = reinterpret_cast<float2*>(keydst);
            const float4* q4 = reinterpret_cast<const float4*>(qv);
            auto slot_of = [&](int i) -> int64_t {
                int64_t slot = (int64_t)idsrc[i] + rows.head;
                return slot >= rows.cap ? slot - rows.cap : slot;
            };
            const int qs = quant_stride(d), qpo = quant_param_offset(d);
            auto row_params = [&](int64_t slot) -> float4 {
                const F4v t = *(GlobalF4)(rows.Xq + slot * qs + qpo);
                return make_float4(t.x, t.y, t.z, t.w);
            };
            auto emit = [&](int i, float a, float df) {
                float2 b = make_float2(-INFINITY, INFINITY);
                if (fabsf(a) <= FLT_MAX && df <= FLT_MAX) {
                    const float dd = df + (fabsf(a) + df) * 0x1p-20f;
                    b = make_float2(a - dd, a + dd);
                }
                lh[i] = b;
            };
            if (idot) {
                // this lane's 16 query elements (unit gl; idle lanes hold zeros)
                const bool act = gl < units;
                float qf[16];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float4 t = act ? q4[4 * gl + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    qf[4 * k] = t.x; qf[4 * k + 1] = t.y; qf[4 * k + 2] = t.z; qf[4 * k + 3] = t.w;
                }
                float m1 = 0.0f;
#pragma unroll
                for (int e = 0; e < 16; ++e) m1 = fmaxf(m1, fabsf(qf[e]));
                for (int o = G >> 1; o > 0; o >>= 1) m1 = fmaxf(m1, __shfl_xor(m1, o, kWave));
                const bool qfin = m1 <= FLT_MAX;
                const float t1 = (qfin && m1 > 0.0f) ? m1 / 127.0f : 0.0f;
                const float i1 = t1 > 0.0f ? 127.0f / m1 : 0.0f;
                int b1[16];
                float m2 = 0.0f;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    b1[e] = t1 > 0.0f ? min(127, max(-127, (int)rintf(qf[e] * i1))) : 0;
                    m2 = fmaxf(m2, fabsf(fmaf(-t1, (float)b1[e], qf[e])));
                }
                for (int o = G >> 1; o > 0; o >>= 1) m2 = fmaxf(m2, __shfl_xor(m2, o, kWave));
                const float t2 = (qfin && m2 > 0.0f) ? m2 / 127.0f : 0.0f;
                const float i2 = t2 > 0.0f ? 127.0f / m2 : 0.0f;
                double e2 = 0.0;
                int n1 = 0, n2 = 0;
                uint32_t w1[4] = {0u, 0u, 0u, 0u}, w2[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int b2 = t2 > 0.0f ? min(127, max(-127, (int)rintf(fmaf(-t1, (float)b1[e], qf[e]) * i2))) : 0;
                    const double dv = (double)qf[e] - (double)t1 * (double)b1[e] - (double)t2 * (double)b2;
                    e2 += dv * dv;
                    n1 += b1[e] * b1[e];
                    n2 += b2 * b2;
                    w1[e >> 2] |= ((uint32_t)b1[e] & 0xffu) << (8 * (e & 3));
                    w2[e >> 2] |= ((uint32_t)b2 & 0xffu) << (8 * (e & 3));
                }
                for (int o = G >> 1; o > 0; o >>= 1) {
                    e2 += __shfl_xor(e2, o, kWave);
                    n1 += __shfl_xor(n1, o, kWave);
                    n2 += __shfl_xor(n2, o, kWave);
                }
                float beta = INFINITY;
                if (qfin) {
                    const double bb = (sqrt(e2) + 12.0 * 0x1p-24 * ((double)t1 * sqrt((double)n1) +
                                                                    (double)t2 * sqrt((double)n2))) * (1.0 + 0x1p-20);
                    beta = (float)bb;
                    if ((double)beta < bb) beta = nextafterf(beta, INFINITY);
                }
                // Lane gl of a G-lane group loads 16 bytes (unit gl) of each of the group's
                // U rows (idle lanes re-read unit 0: their query codes are zero; rows past
                // ce re-read row ce - 1). The lane partial sums are then reduced by a
                // reduce-scatter: each of the first min(LG, 4) xor steps halves the rows a
                // lane carries, the rest are plain butterflies; a lane ends with the full
                // sums of NO = 16 >> min(LG, 4) consecutive rows starting at `rb`.
                static_assert(kPruneU == 16, "the reduce-scatter below sums 16 rows per group");
                const int lu = act ? gl : 0;
                // Row offsets: 32-bit, with a 24-bit multiply, when the ring has at most 2^24
                // slots and every row lies within 4 GiB of the copies' base (loads then take
                // the base in scalar registers); else 64-bit.
                const bool off32 = rows.cap <= (1 << 24) && (uint64_t)rows.cap * (uint64_t)qs <= 0xffffffffull;
                auto run = [&](auto lgc, auto o32c) __attribute__((always_inline)) {
                    constexpr int LG = decltype(lgc)::value;
                    constexpr bool O32 = decltype(o32c)::value;
                    constexpr int GS = 1 << LG;
                    constexpr int H = LG < 4 ? LG : 4;
                    constexpr int NO = U >> H;
                    int rb = 0;
#pragma unroll
                    for (int j = 0; j < H; ++j)
                        if ((gl >> (LG - 1 - j)) & 1) rb += (U / 2) >> j;
                    // emitting lanes: one per row copy
                    const bool emitter = (gl & ((GS >> H) - 1)) == 0;
                    const uint32_t head32 = (uint32_t)rows.head, cap32 = (uint32_t)rows.cap;
                    const int step = kSWaves * groups;
                    for (int base = 0; base < ce; base += per_iter) {
                        const int i0 = base + wave * groups + sub;
                        U4v w[U];
                        float4 prm[NO];
                        if constexpr (O32) {
                            uint32_t off[U];
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                const uint32_t sl = idsrc[min(i0 + u * step, ce - 1)] + head32;
                                off[u] = __umul24(min(sl, sl - cap32), (uint32_t)qs);  // ring wrap: slot < 2 cap
                            }
                            uint32_t poff[NO];
#pragma unroll
                            for (int t = 0; t < NO; ++t) {
                                const uint32_t sl = idsrc[min(i0 + (rb + t) * step, ce - 1)] + head32;
                                poff[t] = __umul24(min(sl, sl - cap32), (uint32_t)qs) + (uint32_t)qpo;
                            }
#pragma unroll
                            for (int u = 0; u < U; ++u) w[u] = *(GlobalU4)(rows.Xq + (off[u] + 16u * (uint32_t)lu));
#pragma unroll
                            for (int t = 0; t < NO; ++t) {
                                const F4v x = *(GlobalF4)(rows.Xq + poff[t]);
                                prm[t] = make_float4(x.x, x.y, x.z, x.w);
                            }
                        } else {
                            uint32_t id[U];
#pragma unroll
                            for (int u = 0; u < U; ++u) id[u] = idsrc[min(i0 + u * step, ce - 1)];
                            uint32_t pid[NO];
#pragma unroll
                            for (int t = 0; t < NO; ++t) pid[t] = idsrc[min(i0 + (rb + t) * step, ce - 1)];
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                int64_t slot = (int64_t)id[u] + rows.head;
                                if (slot >= rows.cap) slot -= rows.cap;
                                w[u] = ((GlobalU4)(rows.Xq + slot * qs))[lu];
                            }
#pragma unroll
                            for (int t = 0; t < NO; ++t) {
                                int64_t slot = (int64_t)pid[t] + rows.head;
                                if (slot >= rows.cap) slot -= rows.cap;
                                prm[t] = row_params(slot);
                            }
                        }
                        int hs[U], ls[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            int h = sdot4_first(w[u].x, w1[0]);
                            int l = sdot4_first(w[u].x, w2[0]);
                            h = __builtin_amdgcn_sdot4((int)w[u].y, (int)w1[1], h, false);
                            l = __builtin_amdgcn_sdot4((int)w[u].y, (int)w2[1], l, false);
                            h = __builtin_amdgcn_sdot4((int)w[u].z, (int)w1[2], h, false);
                            l = __builtin_amdgcn_sdot4((int)w[u].z, (int)w2[2], l, false);
                            h = __builtin_amdgcn_sdot4((int)w[u].w, (int)w1[3], h, false);
                            l = __builtin_amdgcn_sdot4((int)w[u].w, (int)w2[3], l, false);
                            hs[u] = h;
                            ls[u] = l;
                        }
                        reduce_scatter16i<LG, 0>(hs, ls, gl);
#pragma unroll
                        for (int t = 0; t < NO; ++t) {
                            const int i = i0 + (rb + t) * step;
                            if (emitter && i < ce)
                                emit(i, prm[t].x * fmaf(t1, (float)hs[t], t2 * (float)ls[t]),
                                     fmaf(qn, prm[t].y, beta * prm[t].z) * (1.0f + 0x1p-20f));
                        }
                    }
                };
                auto run_g = [&](auto o32c) __attribute__((always_inline)) {
                    switch (G) {
                        case 64: run(std::integral_constant<int, 6>{}, o32c); break;
                        case 32: run(std::integral_constant<int, 5>{}, o32c); break;
                        case 16: run(std::integral_constant<int, 4>{}, o32c); break;
                        case 8: run(std::integral_constant<int, 3>{}, o32c); break;
                        default: run(std::integral_constant<int, 2>{}, o32c); break;
                    }
                };
                if (off32)
                    run_g(std::true_type{});
                else
                    run_g(std::false_type{});
            } else {
                constexpr int UF = kPruneU / 2;
                const int per_iter_f = kSWaves * groups * UF;
                const float g = (float)(d + 64) * 0x1p-24f * (1.0f + 0x1p-10f);
                for (int base = 0; base < ce; base += per_iter_f) {
                    float acc[UF];
                    const int8_t* rp[UF];
                    float4 prm[UF];
#pragma unroll
                    for (int u = 0; u < UF; ++u) {
                        const int i = base + ((u * kSWaves + wave) * groups + sub);
                        acc[u] = 0.0f;
                        rp[u] = nullptr;
                        prm[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (i < ce) {
                            const int64_t slot = slot_of(i);
                            rp[u] = rows.Xq + slot * qs;
                            prm[u] = row_params(slot);
                        }
                    }
                    if (v16) {
                        for (int uu = gl; uu < units; uu += G) {
                            U4v w[UF];
#pragma unroll
                            for (int u = 0; u < UF; ++u)
                                w[u] = rp[u] ? ((GlobalU4)rp[u])[uu] : U4v{0u, 0u, 0u, 0u};
                            const float4 qa = q4[4 * uu], qb4 = q4[4 * uu + 1], qc = q4[4 * uu + 2], qd = q4[4 * uu + 3];
#pragma unroll
                            for (int u = 0; u < UF; ++u) {
                                acc[u] = dot_i8x4(w[u].x, qa, acc[u]);
                                acc[u] = dot_i8x4(w[u].y, qb4, acc[u]);
                                acc[u] = dot_i8x4(w[u].z, qc, acc[u]);
                                acc[u] = dot_i8x4(w[u].w, qd, acc[u]);
                            }
                        }
                    } else {
                        for (int uu = gl; uu < units; uu += G) {
                            uint32_t w[UF];
#pragma unroll
                            for (int u = 0; u < UF; ++u) w[u] = rp[u] ? ((GlobalU1)rp[u])[uu] : 0u;
                            const float4 qa = q4[uu];
#pragma unroll
                            for (int u = 0; u < UF; ++u) acc[u] = dot_i8x4(w[u], qa, acc[u]);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UF; ++u) {
                        float v = acc[u];
                        for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
                        const int i = base + ((u * kSWaves + wave) * groups + sub);
                        if (i < ce && gl == 0) emit(i, prm[u].x * v, qn * fmaf(g, prm[u].z, prm[u].y));
                    }
                }
            }
            __syncthreads();
        };
        // Rerank of the ce candidates whose ids are in `ids` into keys `keys` (LDS cidl/rk,
        // or global idlist/cand in hash mode); nr = keys the top-k reads.
        int nr = ce;
        auto finish = [&](const uint32_t* ids, uint64_t* keys) __attribute__((always_inline)) {
            const int kk0 = min(kq, ce);
            const bool prune = rows.Xq != nullptr && cfg.prune != 0 && kk0 > 0 && kk0 <= 256 && 4 * kk0 <= ce &&
                               (dm.d & 3) == 0;
            if (!prune) {
                rerank(ids, keys, ce, std::integral_constant<int, kRerankU>{});
                return;
            }
            uint32_t* surv = reinterpret_cast<uint32_t*>(lds + lo.hist_off);
            uint32_t* miscu = reinterpret_cast<uint32_t*>(misc);
            if (tid == 0) miscu[11] = 0u;  // survivors
            bound_pass(ids, keys);
            const float2* lh = reinterpret_cast<const float2*>(keys);
            auto lkey = [&](int i) -> uint64_t { return ((uint64_t)fkey(lh[i].x) << 32) | (uint32_t)i; };
            uint64_t prefix;
            int shift;
            radix_select<kSNT>(ce, kk0, lkey, sh, prefix, shift);
            // L: the value bits of the select's prefix. Every selected key is at least the
            // prefix (its bits below `shift` are zero), so L is at most the kk-th largest lo,
            // and equal to it when the select resolved all 32 value bits (shift < 32).
            const float L = shift >= 63 ? -INFINITY : fkey_inv((uint32_t)(prefix >> 32));
            for (int i0 = 0; i0 < ce; i0 += kSNT) {
                const int i = i0 + tid;
                const bool keep = i < ce && lh[i].y >= L;
                const uint32_t pos = wave_append(keep, &miscu[11]);
                if (keep && pos < (uint32_t)kSurvivorMax) surv[pos] = ids[i];
            }
            __syncthreads();
            const int ns = (int)miscu[11];
            if (ns <= kSurvivorMax) {
                nr = ns;
                rerank(surv, keys, ns, std::integral_constant<int, 4>{});
            } else {
                rerank(ids, keys, ce, std::integral_constant<int, kRerankU>{});
            }
        };
        if (hash_mode)
            finish(idlist, cand);
        else
            finish(cidl, rk);

        // ---- d. top-k ----
        const int64_t off = qb.out_off ? qb.out_off[q] : (int64_t)q * qb.k0;
        const int kk = min(kq, ce);
        if (kk > 0) {
            // one instantiation per key array (LDS rk, or global cand in hash mode)
            auto topk = [&](auto get) __attribute__((always_inline)) {
                if (nr <= kSNT) {  // uniform: one key per thread, placed by its rank (keys are unique)
                    uint64_t k = 0;
                    int r = 0;
                    if (tid < nr) {
                        k = get(tid);
                        for (int j = 0; j < nr; ++j) r += get(j) > k ? 1 : 0;
                    }
                    if (tid < nr && r < kk) tk[r] = k;
                    __syncthreads();
                    return;
                }
                uint64_t prefix;
                int shift;
                radix_select<kSNT>(nr, kk, get, sh, prefix, shift);
                const int N = next_pow2(kk);
                for (int i = tid; i < N; i += kSNT) tk[i] = 0;
                if (tid == 0) sh.counter = 0;
                __syncthreads();
                compact_selected<kSNT>(nr, get, prefix, shift, &sh.counter,
                                       [&](uint32_t pos, uint64_t k) { tk[pos] = k; });
                __syncthreads();
                bitonic_sort_desc<kSNT>(tk, next_pow2(kk));
            };
            if (hash_mode)
                topk([&](int i) -> uint64_t { return cand[i]; });
            else
                topk([&](int i) -> uint64_t { return rk[i]; });
        }
        const int kw = qb.out_off ? kq : qb.k0;  // fixed-stride rows are zero padded to the stride
        for (int t = tid; t < kw; t += kSNT) {
            qb.out_id[off + t] = t < kk ? (int32_t)ckey_id(tk[t]) : 0;
            qb.out_ip[off + t] = t < kk ? ckey_val(tk[t]) : 0.0f;
        }
        __syncthreads();
        ++nq_done;
    }
    if (tid == 0 && did_query && scr.est) scr.est[0] = frac;  // any workgroup's value will do
}

int search_ept(int mp) {
    int e = 1;
    while (e * kSNT < mp && e < kBufEntries) e <<= 1;
    return e;
}

hipError_t launch_coceos_search(const IndexDims& dm, const uint32_t* signs, const Entry* bkt,
                                const int32_t* bsize, RowRing rows, const QueryBatch& qb,
                                const QueryConfig& cfg, const QueryScratch& scr, hipStream_t st) {
    if (qb.nq <= 0) return hipSuccess;
    const int DR = dm.D * dm.R;
    const int mp = std::max(1, qb.mp);
    if (mp > kSNT * kBufEntries) return hipErrorInvalidValue;
    if (cfg.tile_buckets < 64 || search_tile_slots_nbk(cfg.tile_buckets) % 4 != 0) return hipErrorInvalidValue;  // 16-byte tile passes
    hipError_t err = hipSuccess;
    if (cfg.ev[0] && (err = hipEventRecord(cfg.ev[0], st)) != hipSuccess) return err;
    err = launch_project_rows(dm, qb.Q, 0, 1, qb.nq, signs, scr.qproj, DR, false, st, cfg.proj_waves, qb.nq_dev);
    if (err != hipSuccess) return err;
    {
        const int pgrid = std::min(qb.nq, 4096);
        const size_t klds = ((size_t)std::max(2, next_pow2(cfg.s_max)) + kProbeList) * 8;
        if (DR <= kProbeStageMax) {
            const size_t plds = klds + (size_t)DR * 4;
            allow_lds_s(k_probe_select<true>, plds);
            hipLaunchKernelGGL(k_probe_select<true>, dim3(pgrid), dim3(kPNT), plds, st, DR, scr.qproj, qb.nq,
                               qb.nq_dev, qb.s, qb.s0, cfg.s_max, scr.probes);
        } else {
            allow_lds_s(k_probe_select<false>, klds);
            hipLaunchKernelGGL(k_probe_select<false>, dim3(pgrid), dim3(kPNT), klds, st, DR, scr.qproj, qb.nq,
                               qb.nq_dev, qb.s, qb.s0, cfg.s_max, scr.probes);
        }
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    err = launch_fill_i32(scr.work, 0, 1, st);
    if (err != hipSuccess) return err;
    const int grid = std::max(1, std::min(cfg.grid, qb.nq));
    if (cfg.ev[1] && (err = hipEventRecord(cfg.ev[1], st)) != hipSuccess) return err;
    switch (search_ept(mp)) {
#define CEOS_EPT(E)                                                                                            \
    case E:                                                                                                    \
        allow_lds_s(k_search_fused<E>, cfg.lds_bytes);                                                        \
        hipLaunchKernelGGL(k_search_fused<E>, dim3(grid), dim3(kSNT), cfg.lds_bytes, st, dm, scr.probes,        \
                           cfg.s_max, bkt, bsize, rows, qb, cfg, scr);                                         \
        break;
        CEOS_EPT(1)
        CEOS_EPT(2)
        CEOS_EPT(4)
        CEOS_EPT(8)

#undef CEOS_EPT
        default:
            return hipErrorInvalidValue;
    }
    err = hipGetLastError();
    if (err == hipSuccess && cfg.ev[2]) err = hipEventRecord(cfg.ev[2], st);
    return err;
}

}  // namespace ceos
