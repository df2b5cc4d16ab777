// coCEOs search on gfx950: three launches per batch of queries.
//
// 1. k_project_rows (grid-wide): every query's D*R projections, row-major [nq][DR]
//    (sign + FWHT, 3 rounds per repeat; src/CEOs.cpp:495-513).
// 2. k_probe_select (one 256-thread block per query): exact top-s of |p| by radix
//    select, sorted to ascending |p| -- the reference's pop order -- and encoded as
//    list ids 2c (close, p >= 0) / 2c+1 (far, p < 0) (src/CEOs.cpp:517-548, 562-564).
// 3. k_search_fused (persistent, one 512-thread workgroup per query at a time, next
//    query from a global work counter so uneven k / s / c balance across the chip):
//    a. score accumulation into an LDS hash tile of 8-byte (id, FP32 sum) slots.
//       Bucket entries are prefetched into registers a chunk of probes at a time,
//       double buffered, and applied probe by probe with a barrier between probes.
//       Ids within one bucket list are distinct, so each point's sum is formed in
//       probe order exactly as the reference's robin-map accumulation
//       (src/CEOs.cpp:558-610). When the distinct ids exceed the tile, ids are split
//       into hash partitions done in successive passes; the partition count adapts
//       per query and restarts doubled on overflow;
//    b. exact top-c of the estimates by (score desc, id asc) (src/CEOs.cpp:619-632);
//       with fewer distinct candidates than c all are kept;
//    c. rerank: FP32 inner products of the candidates' rows with the query, eight
//       rows in flight per lane group (src/CEOs.cpp:638-657);
//    d. top-k by (inner product desc, id asc), written descending, zero padded
//       (src/CEOs.cpp:659-666).
#include <algorithm>

#include "ceos_device.h"
#include "ceos_kernels.h"

namespace ceos {

constexpr int kSNT = kSearchThreads;   // 512
constexpr int kSWaves = kSNT / kWave;
constexpr int kPNT = 256;              // probe-select block
constexpr uint32_t kEmptyKey = 0xffffffffu;
constexpr uint64_t kEmptySlot = 0xffffffffull;  // key = empty, value bits = 0
constexpr int kBufEntries = 8;         // bucket entries per thread per chunk
constexpr uint32_t kMulti = 0x80000000u;  // key flag: the id occurs in more than one probed list
constexpr uint32_t kUnset = 0xffffffffu;  // value bits of a multi slot before its first ordered add
constexpr uint64_t kNeverSlot = 0x7fffffffull;  // key 0x7fffffff: no slot ever holds it (ids < 2^31-1)
constexpr uint32_t kStashMark = 0xfffffffeu;   // phase-1 slot value: entry sent to the stash
constexpr int kStashSlots = 1024;              // linear-probing stash at the end of the tile
constexpr int kMaxRounds = 12;                 // claim rounds per chunk before the stash
constexpr int kStashList = kSearchStashList;   // pending stash entries per pass (global list)

__device__ __forceinline__ uint32_t bucket1(uint32_t id, uint32_t nbk) {
    return (uint32_t)(((uint64_t)(id * 0x9e3779b1u) * nbk) >> 32);
}
__device__ __forceinline__ uint32_t bucket2(uint32_t id, uint32_t b1, uint32_t nbk) {
    const uint32_t b2 = (uint32_t)(((uint64_t)((id ^ 0x5bd1e995u) * 0x85ebca6bu) * nbk) >> 32);
    return b2 == b1 ? (b1 + 1 == nbk ? 0u : b1 + 1) : b2;
}
constexpr int kRerankU = 8;            // rows in flight per lane group

template <typename K>
static void allow_lds_s(K kernel, size_t bytes) {
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

__host__ __device__ inline int al16(int b) { return (b + 15) & ~15; }

struct SLayout {
    int q_off, probe_off, lim_off, misc_off, cid_off, u_off;
    int tab_off, rk_off, tk_off, total;
};

__host__ __device__ inline SLayout s_layout(int d, int tlog2, int s_max, int c_max, int k_max) {
    SLayout o;
    int off = 0;
    o.q_off = off; off += al16(d * 4);
    o.probe_off = off; off += al16(s_max * 4);
    o.lim_off = off; off += al16(s_max * 4);
    o.misc_off = off; off += 64;
    o.cid_off = off; off += al16(c_max * 4);
    o.u_off = off;
    o.tab_off = off;
    const int endA = off + (4 << tlog2);
    o.rk_off = off;
    o.tk_off = off + al16(c_max * 8);
    const int endB = o.tk_off + next_pow2(k_max) * 8;
    o.total = endA > endB ? endA : endB;
    return o;
}

int search_lds_bytes(int d, int table_log2, int s_max, int c_max, int k_max) {
    return s_layout(d, table_log2, s_max, c_max, k_max).total;
}

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kPNT) void k_probe_select(int DR, const float* __restrict__ qproj, int nq,
                                                       const int32_t* __restrict__ s_arr, int s0, int s_stride,
                                                       int32_t* __restrict__ probes) {
    __shared__ SelectShared sh;
    extern __shared__ uint64_t pkey[];
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const int sq = min(max(s_arr ? s_arr[q] : s0, 0), min(s_stride, DR));
        if (sq == 0) continue;  // uniform
        const float* pq = qproj + (int64_t)q * DR;
        auto get = [&](int c) -> uint64_t { return ckey(fabsf(pq[c]), (uint32_t)c); };
        uint64_t prefix;
        int shift;
        radix_select<kPNT>(DR, sq, get, sh, prefix, shift);
        const int N = next_pow2(sq);
        for (int i = threadIdx.x; i < N; i += kPNT) pkey[i] = 0;
        if (threadIdx.x == 0) sh.counter = 0;
        __syncthreads();
        compact_selected<kPNT>(DR, get, prefix, shift, &sh.counter, [&](uint32_t pos, uint64_t k) { pkey[pos] = k; });
        __syncthreads();
        bitonic_sort_desc<kPNT>(pkey, N);
        for (int j = threadIdx.x; j < sq; j += kPNT) {
            const int c = (int)ckey_id(pkey[sq - 1 - j]);
            probes[(int64_t)q * s_stride + j] = 2 * c + (pq[c] < 0.0f ? 1 : 0);
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
    constexpr int B = kBufEntries;         // entries per thread per chunk
    __shared__ SelectShared sh;
    extern __shared__ __align__(16) unsigned char lds[];
    const SLayout lo = s_layout(dm.d, cfg.table_log2, cfg.s_max, cfg.c_max, cfg.k_max);
    float* qv = reinterpret_cast<float*>(lds + lo.q_off);
    int* probes = reinterpret_cast<int*>(lds + lo.probe_off);
    int* lims = reinterpret_cast<int*>(lds + lo.lim_off);
    int* misc = reinterpret_cast<int*>(lds + lo.misc_off);
    uint32_t* cidl = reinterpret_cast<uint32_t*>(lds + lo.cid_off);
    uint32_t* keys = reinterpret_cast<uint32_t*>(lds + lo.tab_off);  // phase 1: ids; phase 2: FP32 sums
    uint64_t* tile64 = reinterpret_cast<uint64_t*>(lds + lo.tab_off);
    uint64_t* rk = reinterpret_cast<uint64_t*>(lds + lo.rk_off);
    uint64_t* tk = reinterpret_cast<uint64_t*>(lds + lo.tk_off);
    const int T = 1 << cfg.table_log2;
    const int sslots = min(kStashSlots, T / 8);          // stash keys at the end of the tile
    const uint32_t nbk = (uint32_t)((T - sslots) / 4);  // main buckets of 4 keys (16 bytes)
    uint64_t* cand = scr.cand + (int64_t)blockIdx.x * cfg.cand_stride;
    uint2* pairs = scr.pairs + (int64_t)blockIdx.x * cfg.cand_stride;
    uint32_t* keys_g = scr.keys + ((int64_t)blockIdx.x << cfg.table_log2);
    uint4* stash_list = scr.stash + (int64_t)blockIdx.x * kStashList;
    const int tid = threadIdx.x, wave = tid / kWave, ln = tid % kWave;
    float frac = 0.5f;  // running estimate of distinct ids / entries for this workgroup

    for (;;) {
        if (tid == 0) misc[0] = atomicAdd(scr.work, 1);
        __syncthreads();
        const int q = misc[0];
        __syncthreads();
        if (q >= qb.nq) break;
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
        // Phase 1 (any order): every id of the probed lists gets a tile slot. An id claims an
        //   empty slot of its bucket with a non-returning atomicMin between barriers; the next
        //   round's read of the bucket shows whether it holds the slot (reads and claims of a
        //   round are separated by a barrier, so a read never sees a claim of its own round).
        //   An id lives in the first of its two buckets that had room when it arrived, and
        //   nothing is removed within a pass. Ids that find both buckets full go to the stash.
        //   Each bucket position's (slot, value) is kept in per-workgroup scratch.
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
        uint32_t P = (uint32_t)max(1, (int)ceilf((float)E * fminf(1.0f, frac * 1.25f) / (0.7f * (float)T)));
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
        auto phase1_chunk = [&](const Entry (&buf)[CH][EPT], int j0, uint32_t pass) {
            uint32_t id[B], bk[B], slot[B];
            float v[B];
            uint32_t pend = 0, second = 0;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = j0 + b / EPT, t = tid + (b % EPT) * kSNT;
                id[b] = buf[b / EPT][b % EPT].x - rows.base;
                v[b] = __uint_as_float(buf[b / EPT][b % EPT].y);
                bk[b] = bucket1(id[b], nbk);
                slot[b] = kEmptyKey;
                const bool valid = j < sq && t < lims[min(j, jlast)] &&
                                   (P == 1 || (uint32_t)(((uint64_t)(id[b] * 0x85ebca77u) * P) >> 32) == pass);
                pend |= (valid ? 1u : 0u) << b;
            }
            for (int round = 0;; ++round) {
                uint4 kk[B];
#pragma unroll
                for (int b = 0; b < B; ++b) kk[b] = reinterpret_cast<const uint4*>(keys)[bk[b]];
                uint32_t claim = 0, cslot[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t k0 = kk[b].x, k1 = kk[b].y, k2 = kk[b].z, k3 = kk[b].w;
                    const int hit = k0 == id[b] ? 0 : k1 == id[b] ? 1 : k2 == id[b] ? 2 : k3 == id[b] ? 3 : -1;
                    const int emp = k0 == kEmptyKey ? 0 : k1 == kEmptyKey ? 1 : k2 == kEmptyKey ? 2 : k3 == kEmptyKey ? 3 : -1;
                    cslot[b] = 4 * bk[b] + (uint32_t)(emp < 0 ? 0 : emp);
                    if ((pend >> b) & 1u) {
                        if (hit >= 0) {
                            slot[b] = 4 * bk[b] + (uint32_t)hit;
                            pend &= ~(1u << b);
                        } else if (emp >= 0 && round < kMaxRounds) {
                            claim |= 1u << b;
                        } else if (!((second >> b) & 1u) && round < kMaxRounds) {
                            second |= 1u << b;
                            bk[b] = bucket2(id[b], bk[b], nbk);
                        } else {
                            slot[b] = kStashMark;  // both buckets full (or out of rounds): stash
                            pend &= ~(1u << b);
                        }
                    }
                }
                __syncthreads();  // every read of this round precedes every claim of this round
#pragma unroll
                for (int b = 0; b < B; ++b)
                    if ((claim >> b) & 1u) atomicMin(&keys[cslot[b]], id[b]);
                if (!__syncthreads_or(pend != 0u)) break;
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = j0 + b / EPT, t = tid + (b % EPT) * kSNT;
                if (slot[b] == kStashMark) {
                    const uint32_t pos = atomicAdd(reinterpret_cast<uint32_t*>(&misc[5]), 1u);
                    if (pos < (uint32_t)kStashList)
                        stash_list[pos] = make_uint4(id[b], __float_as_uint(v[b]), (uint32_t)(j * mp + t), 0u);
                    else
                        misc[4] = 1;  // stash list full: the pass restarts with more partitions
                }
                if (j < sq && t < lims[min(j, jlast)])
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
                uint32_t found = kEmptyKey;
                const uint32_t b1 = bucket1(idv, nbk), b2 = bucket2(idv, b1, nbk);
                for (int q2 = 0; q2 < 8 && found == kEmptyKey; ++q2) {
                    const uint32_t si = 4 * (q2 < 4 ? b1 : b2) + (q2 & 3);
                    if (keys[si] == idv) found = si;
                }
                if (found == kEmptyKey) {
                    uint32_t h = (idv * 0x9e3779b1u) & (uint32_t)(sslots - 1);
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
                if (found != kEmptyKey) pairs[e.z].x = found;
            }
        };
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
                    }
                }
                __syncthreads();
            }
        };

        int ncand = 0;
        for (; sq > 0;) {
            if (tid == 0) { misc[2] = 0; misc[4] = 0; }
            bool overflow = false;
            for (uint32_t pass = 0; pass < P && !overflow; ++pass) {
                for (int i = tid; i < T; i += kSNT) keys[i] = kEmptyKey;
                if (tid == 0) misc[5] = 0;
                __syncthreads();
                // phase 1
                {
                    Entry bufA[CH][EPT], bufB[CH][EPT];
                    load_chunk(bufA, 0);
                    for (int j0 = 0; j0 < sq; j0 += 2 * CH) {
                        load_chunk(bufB, j0 + CH);
                        phase1_chunk(bufA, j0, pass);
                        if (j0 + CH >= sq) break;
                        load_chunk(bufA, j0 + 2 * CH);
                        phase1_chunk(bufB, j0 + CH, pass);
                    }
                }
                __syncthreads();
                if (misc[5] > 0) {  // uniform
                    stash_insert();
                    __syncthreads();
                }
                if (misc[4] != 0) { overflow = true; break; }
                // ids to scratch; the tile becomes the sums ("unset" until the first value)
                for (int i = tid; i < T; i += kSNT) {
                    keys_g[i] = keys[i];
                    keys[i] = kUnset;
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
                // candidates of this pass: (sum, id) keys to scratch
                for (int i0 = 0; i0 < T; i0 += kSNT) {  // uniform trip count: wave_append needs all lanes
                    const int i = i0 + tid;
                    const uint32_t idv = i < T ? keys_g[i] : kEmptyKey;
                    const bool take = idv != kEmptyKey;
                    const uint32_t pos = wave_append(take, reinterpret_cast<uint32_t*>(&misc[2]));
                    if (take) cand[pos] = ckey(__uint_as_float(keys[i]), idv);
                }
                __syncthreads();
            }
            if (!overflow) {
                ncand = misc[2];
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
        const int ce = min(cq, ncand);
        if (ce > 0) {
            const bool lds_keys = 2 * ncand <= T;  // 8-byte keys fit the (now free) tile
            if (lds_keys) {
                for (int i = tid; i < ncand; i += kSNT) tile64[i] = cand[i];
                __syncthreads();
            }
            auto get = [&](int i) -> uint64_t { return lds_keys ? tile64[i] : cand[i]; };
            uint64_t prefix;
            int shift;
            radix_select<kSNT>(ncand, ce, get, sh, prefix, shift);
            if (tid == 0) sh.counter = 0;
            __syncthreads();
            compact_selected<kSNT>(ncand, get, prefix, shift, &sh.counter,
                                   [&](uint32_t pos, uint64_t k) { cidl[pos] = ckey_id(k); });
            __syncthreads();
        }

        // ---- c. rerank ----
        {
            const int d = dm.d;
            const bool vec = (d & 3) == 0;
            const int units = vec ? (d >> 2) : d;
            int G = 1;
            while (2 * G < units && G < kWave) G <<= 1;
            const int groups = kWave / G;
            const int sub = ln / G, gl = ln % G;
            const int per_iter = kSWaves * groups * kRerankU;
            for (int base = 0; base < ce; base += per_iter) {
                float acc[kRerankU];
                const float* rp[kRerankU];
                uint32_t ids[kRerankU];
#pragma unroll
                for (int u = 0; u < kRerankU; ++u) {
                    const int i = base + ((u * kSWaves + wave) * groups + sub);
                    acc[u] = 0.0f;
                    rp[u] = nullptr;
                    ids[u] = 0;
                    if (i < ce) {
                        ids[u] = cidl[i];
                        int64_t slot = (int64_t)ids[u] + rows.head;
                        if (slot >= rows.cap) slot -= rows.cap;
                        rp[u] = rows.X + slot * d;
                    }
                }
                if (vec) {
                    const float4* q4 = reinterpret_cast<const float4*>(qv);
                    for (int uu = gl; uu < units; uu += G) {
                        float4 a[kRerankU];
#pragma unroll
                        for (int u = 0; u < kRerankU; ++u)
                            a[u] = rp[u] ? reinterpret_cast<const float4*>(rp[u])[uu] : make_float4(0, 0, 0, 0);
                        const float4 b = q4[uu];
#pragma unroll
                        for (int u = 0; u < kRerankU; ++u) {
                            acc[u] = fmaf(a[u].x, b.x, acc[u]);
                            acc[u] = fmaf(a[u].y, b.y, acc[u]);
                            acc[u] = fmaf(a[u].z, b.z, acc[u]);
                            acc[u] = fmaf(a[u].w, b.w, acc[u]);
                        }
                    }
                } else {
                    for (int uu = gl; uu < d; uu += G) {
                        float a[kRerankU];
#pragma unroll
                        for (int u = 0; u < kRerankU; ++u) a[u] = rp[u] ? rp[u][uu] : 0.0f;
                        const float b = qv[uu];
#pragma unroll
                        for (int u = 0; u < kRerankU; ++u) acc[u] = fmaf(a[u], b, acc[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < kRerankU; ++u) {
                    float v = acc[u];
                    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
                    const int i = base + ((u * kSWaves + wave) * groups + sub);
                    if (i < ce && gl == 0) rk[i] = ckey(v, ids[u]);
                }
            }
            __syncthreads();
        }

        // ---- d. top-k ----
        const int64_t off = qb.out_off ? qb.out_off[q] : (int64_t)q * qb.k0;
        const int kk = min(kq, ce);
        if (kk > 0) {
            auto get = [&](int i) -> uint64_t { return rk[i]; };
            uint64_t prefix;
            int shift;
            radix_select<kSNT>(ce, kk, get, sh, prefix, shift);
            const int N = next_pow2(kk);
            for (int i = tid; i < N; i += kSNT) tk[i] = 0;
            if (tid == 0) sh.counter = 0;
            __syncthreads();
            compact_selected<kSNT>(ce, get, prefix, shift, &sh.counter,
                                   [&](uint32_t pos, uint64_t k) { tk[pos] = k; });
            __syncthreads();
            bitonic_sort_desc<kSNT>(tk, N);
        }
        for (int t = tid; t < kq; t += kSNT) {
            qb.out_id[off + t] = t < kk ? (int32_t)ckey_id(tk[t]) : 0;
            qb.out_ip[off + t] = t < kk ? ckey_val(tk[t]) : 0.0f;
        }
        __syncthreads();
    }
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
    hipError_t err = launch_project_rows(dm, qb.Q, 0, 1, qb.nq, signs, scr.qproj, DR, false, st);
    if (err != hipSuccess) return err;
    {
        const int pgrid = std::min(qb.nq, 4096);
        const size_t plds = (size_t)next_pow2(cfg.s_max) * 8;
        allow_lds_s(k_probe_select, plds);
        hipLaunchKernelGGL(k_probe_select, dim3(pgrid), dim3(kPNT), plds, st, DR, scr.qproj, qb.nq, qb.s, qb.s0,
                           cfg.s_max, scr.probes);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    err = launch_fill_i32(scr.work, 0, 1, st);
    if (err != hipSuccess) return err;
    const int grid = std::max(1, std::min(cfg.grid, qb.nq));
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
    return hipGetLastError();
}

}  // namespace ceos
