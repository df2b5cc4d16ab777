// CEOs-Est query kernel (gfx950); the coCEOs search lives in ceos_search.hip.
//
// One workgroup (4 waves) serves one query at a time and takes the next query from
// a global work counter (persistent grid), so queries with different k, s and c in
// one launch balance across the chip. Per query:
//   1. project the query: one wave per repeat (sign + FWHT, 3 rounds);
//      src/CEOs.cpp:495-513
//   2. probe selection: exact top-s of |p| (radix select), sorted to ascending |p|,
//      the reference's pop order; close when p >= 0, far when p < 0;
//      src/CEOs.cpp:517-548, 562-564
//   3. score accumulation into an LDS hash tile (ids -> FP32 sums). Probes are
//      applied one at a time with a barrier between them, and ids within one bucket
//      list are distinct, so every point's sum is formed in probe order exactly as
//      the reference's robin-map accumulation (src/CEOs.cpp:558-610). When the
//      entries exceed half the tile, ids are split into hash partitions handled in
//      successive passes;
//   4. exact top-c of the estimates (radix select on (score, id) keys); with fewer
//      distinct candidates than c all of them are kept; src/CEOs.cpp:619-632
//   5. rerank: FP32 inner products of the candidates' rows with the query;
//      src/CEOs.cpp:638-657
//   6. top-k by (inner product desc, id asc), written descending and zero padded;
//      src/CEOs.cpp:659-666
#include <algorithm>

#include "ceos_device.h"
#include "ceos_kernels.h"

namespace ceos {

// Kernels whose dynamic LDS exceeds 64 KiB must opt in to the larger size.
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

#define CEOS_SWITCH_L(Lv, ...)                           \
    switch (Lv) {                                         \
        case 1: { constexpr int LL = 1; __VA_ARGS__; } break;   \
        case 2: { constexpr int LL = 2; __VA_ARGS__; } break;   \
        case 3: { constexpr int LL = 3; __VA_ARGS__; } break;   \
        case 4: { constexpr int LL = 4; __VA_ARGS__; } break;   \
        case 5: { constexpr int LL = 5; __VA_ARGS__; } break;   \
        case 6: { constexpr int LL = 6; __VA_ARGS__; } break;   \
        case 7: { constexpr int LL = 7; __VA_ARGS__; } break;   \
        case 8: { constexpr int LL = 8; __VA_ARGS__; } break;   \
        case 9: { constexpr int LL = 9; __VA_ARGS__; } break;   \
        case 10: { constexpr int LL = 10; __VA_ARGS__; } break; \
        case 11: { constexpr int LL = 11; __VA_ARGS__; } break; \
        case 12: { constexpr int LL = 12; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;                  \
    }

constexpr int kQNT = 256;
constexpr int kQWaves = kQNT / kWave;
constexpr uint32_t kEmpty = 0xffffffffu;

__host__ __device__ inline int align16(int b) { return (b + 15) & ~15; }
__host__ __device__ inline int tile_words(int L) {
    return L >= 8 ? (1 << L) + 64 : (1 << L);
}

// CEOs-Est candidate bound: estimates of kEstBlocks blocks of 256 points spread over the
// data give a lower bound of the c-th largest estimate
constexpr int kEstBlocks = 16;
constexpr int kEstSample = kEstBlocks * kQNT;
constexpr int kEstFlight = 8;  // columns loaded together per point group

struct QLayout {
    int q_off, probe_off, pkey_off, misc_off, u_off;
    int tile_off, proj_off, tab_off, rk_off, tk_off, total;
};

__host__ __device__ inline QLayout q_layout(int d, int L, int DR, int tlog2, int s_max, int c_max, int k_max) {
    QLayout o;
    int off = 0;
    o.q_off = off; off += align16(d * 4);
    o.probe_off = off; off += align16(s_max * 4);
    o.pkey_off = off; off += next_pow2(s_max) * 8;
    o.misc_off = off; off += 64;
    o.u_off = off;
    o.tile_off = off;
    o.proj_off = off + align16(kQWaves * tile_words(L) * 4);
    const int endA = o.proj_off + align16(DR * 4);
    o.tab_off = off;
    const int endB = tlog2 > 0 ? o.tab_off + (8 << tlog2) : off;
    o.rk_off = off;
    o.tk_off = off + align16(c_max * 8);
    const int endD = o.tk_off + next_pow2(k_max) * 8;
    const int endS = off + kEstSample * 4;  // CEOs-Est sample keys (after the probe selection)
    o.total = endA > endB ? endA : endB;
    if (endD > o.total) o.total = endD;
    if (endS > o.total) o.total = endS;
    return o;
}

int est_search_lds_bytes(const IndexDims& dm, int s_max, int c_max, int k_max) {
    return q_layout(dm.d, dm.L, dm.D * dm.R, 0, s_max, c_max, k_max).total;
}

// ---------------------------------------------------------------------------------
// Shared phases
// ---------------------------------------------------------------------------------
template <int L>
__device__ void project_query(const IndexDims& dm, const float* __restrict__ qrow,
                              const uint32_t* __restrict__ signs, float* tile, float* proj, float* qv) {
    using P = Proj<L>;
    const int tid = threadIdx.x, wave = tid / kWave, ln = tid % kWave;
    for (int i = tid; i < dm.d; i += kQNT) qv[i] = qrow[i];
    const int W = ((1 << L) + 31) / 32;
    for (int r = wave; r < dm.R; r += kQWaves) {
        float xr[P::XE];
        P::load_x(qrow, dm.d, xr, ln);
        float y[P::OE];
        P::run(xr, signs + (size_t)r * 3 * W, tile + wave * P::kLdsWords, ln, y);
#pragma unroll
        for (int reg = 0; reg < P::OE; ++reg) {
            const int j = reg * kWave + ln;
            if (j < dm.D) proj[r * dm.D + j] = y[reg];
        }
    }
    __syncthreads();
}

// Top-s columns by |p| into probes[0..s) in ascending |p| order, as list ids 2c+far.
__device__ void select_probes(int DR, int s, const float* proj, uint64_t* pkey, int* probes, SelectShared& sh) {
    const int tid = threadIdx.x;
    auto get = [&](int c) -> uint64_t { return ckey(fabsf(proj[c]), (uint32_t)c); };
    uint64_t prefix;
    int shift;
    radix_select<kQNT>(DR, s, get, sh, prefix, shift);
    const int N = next_pow2(s);
    for (int i = tid; i < N; i += kQNT) pkey[i] = 0;
    if (tid == 0) sh.counter = 0;
    __syncthreads();
    for (int c = tid; c < DR; c += kQNT) {
        const uint64_t k = get(c);
        if (selected(k, prefix, shift)) pkey[atomicAdd(&sh.counter, 1u)] = k;
    }
    __syncthreads();
    bitonic_sort_desc<kQNT>(pkey, N);
    for (int j = tid; j < s; j += kQNT) {
        const int c = (int)ckey_id(pkey[s - 1 - j]);
        probes[j] = 2 * c + (proj[c] < 0.0f ? 1 : 0);
    }
    __syncthreads();
}

// Exact top-c of keys[0..n) (global) compacted into sel[0..min(c,n)).
__device__ int select_candidates(const uint64_t* keys, int n, int c, uint64_t* sel, SelectShared& sh) {
    const int tid = threadIdx.x;
    const int ce = min(c, n);
    if (ce <= 0) return 0;
    auto get = [&](int i) -> uint64_t { return keys[i]; };
    uint64_t prefix;
    int shift;
    radix_select<kQNT>(n, ce, get, sh, prefix, shift);
    if (tid == 0) sh.counter = 0;
    __syncthreads();
    for (int i = tid; i < n; i += kQNT) {
        const uint64_t k = keys[i];
        if (selected(k, prefix, shift)) sel[atomicAdd(&sh.counter, 1u)] = k;
    }
    __syncthreads();
    return ce;
}

// FP32 inner products of the candidates' rows with the query (query in LDS).
__device__ void rerank(const float* qv, int d, const RowRing& rows, const uint64_t* sel, int ce, uint64_t* rk) {
    const int tid = threadIdx.x, wave = tid / kWave, ln = tid % kWave;
    const bool vec = (d & 3) == 0;
    const int units = vec ? (d >> 2) : d;
    int G = 1;
    while (G < units && G < kWave) G <<= 1;
    const int per_wave = kWave / G;
    const int sub = ln / G, gl = ln % G;
    for (int i0 = wave * per_wave; i0 < ce; i0 += kQWaves * per_wave) {
        const int i = i0 + sub;
        float acc = 0.0f;
        uint32_t id = 0;
        if (i < ce) {
            id = ckey_id(sel[i]);
            int64_t slot = (int64_t)id + rows.head;
            if (slot >= rows.cap) slot -= rows.cap;
            const float* x = rows.X + slot * d;
            if (vec) {
                const float4* x4 = reinterpret_cast<const float4*>(x);
                const float4* q4 = reinterpret_cast<const float4*>(qv);
                for (int u = gl; u < units; u += G) {
                    const float4 a = x4[u], b = q4[u];
                    acc = fmaf(a.x, b.x, acc);
                    acc = fmaf(a.y, b.y, acc);
                    acc = fmaf(a.z, b.z, acc);
                    acc = fmaf(a.w, b.w, acc);
                }
            } else {
                for (int u = gl; u < d; u += G) acc = fmaf(x[u], qv[u], acc);
            }
        }
        for (int o = G >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        if (i < ce && gl == 0) rk[i] = ckey(acc, id);
    }
    __syncthreads();
}

__device__ void write_topk(uint64_t* rk, int ce, int kq, uint64_t* tk, SelectShared& sh, int32_t* out_id,
                           float* out_ip) {
    const int tid = threadIdx.x;
    const int kk = min(kq, ce);
    if (kk > 0) {
        auto get = [&](int i) -> uint64_t { return rk[i]; };
        uint64_t prefix;
        int shift;
        radix_select<kQNT>(ce, kk, get, sh, prefix, shift);
        const int N = next_pow2(kk);
        for (int i = tid; i < N; i += kQNT) tk[i] = 0;
        if (tid == 0) sh.counter = 0;
        __syncthreads();
        for (int i = tid; i < ce; i += kQNT) {
            const uint64_t k = rk[i];
            if (selected(k, prefix, shift)) tk[atomicAdd(&sh.counter, 1u)] = k;
        }
        __syncthreads();
        bitonic_sort_desc<kQNT>(tk, N);
    }
    for (int t = tid; t < kq; t += kQNT) {
        if (t < kk) {
            out_id[t] = (int32_t)ckey_id(tk[t]);
            out_ip[t] = ckey_val(tk[t]);
        } else {
            out_id[t] = 0;
            out_ip[t] = 0.0f;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ int next_query(int32_t* work, int* misc) {
    if (threadIdx.x == 0) misc[0] = atomicAdd(work, 1);
    __syncthreads();
    const int q = misc[0];
    __syncthreads();
    return q;
}

// ---------------------------------------------------------------------------------
// CEOs-Est dense search (src/CEOs.cpp:107-297)
// ---------------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(kQNT) void k_est_search(IndexDims dm, const uint32_t* __restrict__ signs,
                                                     const float* __restrict__ Pm, int64_t n,
                                                     const float* __restrict__ X, QueryBatch qb, int s_max,
                                                     int c_max, int k_max, uint64_t* scratch,
                                                     int64_t scratch_stride, int32_t* work) {
    __shared__ SelectShared sh;
    extern __shared__ __align__(16) unsigned char lds[];
    const int DR = dm.D * dm.R;
    const QLayout lo = q_layout(dm.d, L, DR, 0, s_max, c_max, k_max);
    float* qv = reinterpret_cast<float*>(lds + lo.q_off);
    int* probes = reinterpret_cast<int*>(lds + lo.probe_off);
    uint64_t* pkey = reinterpret_cast<uint64_t*>(lds + lo.pkey_off);
    int* misc = reinterpret_cast<int*>(lds + lo.misc_off);
    float* tile = reinterpret_cast<float*>(lds + lo.tile_off);
    float* proj = reinterpret_cast<float*>(lds + lo.proj_off);
    uint64_t* rk = reinterpret_cast<uint64_t*>(lds + lo.rk_off);
    uint64_t* tk = reinterpret_cast<uint64_t*>(lds + lo.tk_off);
    uint64_t* est = scratch + (int64_t)blockIdx.x * scratch_stride;
    uint64_t* sel = est + n;
    const RowRing rows{X, n, 0, 0u, (uint32_t)n};
    const int tid = threadIdx.x;

    for (;;) {
        const int q = next_query(work, misc);
        if (q >= qb.nq) break;
        const int kq = qb.k ? qb.k[q] : qb.k0;
        const int sq = qb.s ? qb.s[q] : qb.s0;
        const int cq = qb.c ? qb.c[q] : qb.c0;
        const float* qrow = qb.Q + (int64_t)q * dm.d;
        project_query<L>(dm, qrow, signs, tile, proj, qv);
        select_probes(DR, sq, proj, pkey, probes, sh);
        // Estimates in probe order (the reference's vecEst += / -= per probe,
        // src/CEOs.cpp:208-219), kEstFlight column loads in flight per point (group).
        auto est1 = [&](int64_t i) -> float {
            float e = 0.0f;
            for (int j0 = 0; j0 < sq; j0 += kEstFlight) {
                float v[kEstFlight];
#pragma unroll
                for (int u = 0; u < kEstFlight; ++u)
                    v[u] = j0 + u < sq ? Pm[(int64_t)(probes[j0 + u] >> 1) * n + i] : 0.0f;
#pragma unroll
                for (int u = 0; u < kEstFlight; ++u)
                    if (j0 + u < sq) e = (probes[j0 + u] & 1) ? e - v[u] : e + v[u];
            }
            return e;
        };
        auto est4 = [&](int64_t i) -> float4 {  // points i..i+3; n % 4 == 0, i % 4 == 0
            float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int j0 = 0; j0 < sq; j0 += kEstFlight) {
                float4 v[kEstFlight];
#pragma unroll
                for (int u = 0; u < kEstFlight; ++u)
                    v[u] = j0 + u < sq ? *reinterpret_cast<const float4*>(Pm + (int64_t)(probes[j0 + u] >> 1) * n + i)
                                       : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int u = 0; u < kEstFlight; ++u) {
                    if (j0 + u >= sq) continue;
                    if (probes[j0 + u] & 1) {
                        e.x -= v[u].x; e.y -= v[u].y; e.z -= v[u].z; e.w -= v[u].w;
                    } else {
                        e.x += v[u].x; e.y += v[u].y; e.z += v[u].z; e.w += v[u].w;
                    }
                }
            }
            return e;
        };
        const bool vec4 = (n & 3) == 0;
        // lower bound of the cq-th largest estimate from the sample: the sample value of
        // about twice the expected rank, plus 8 (none when the sample is not small against
        // n, or cq is a large share of n)
        uint32_t lkey = 0;
        bool bounded = false;
        const int64_t rs = 2 * (((int64_t)cq * kEstSample + n - 1) / n) + 8;
        if (n >= 8 * (int64_t)kEstSample && 2 * rs <= kEstSample) {
            uint32_t* skey = reinterpret_cast<uint32_t*>(lds + lo.u_off);
            for (int b = 0; b < kEstBlocks; ++b) {
                const int64_t start = ((int64_t)b * (n - kQNT) / (kEstBlocks - 1)) & ~(int64_t)3;
                skey[b * kQNT + tid] = fkey(est1(start + tid));
            }
            __syncthreads();
            uint64_t prefix;
            int shift;
            radix_select<kQNT>(kEstSample, (int)rs, [&](int i) -> uint64_t { return (uint64_t)skey[i] << 32; }, sh,
                               prefix, shift, 32);
            lkey = (uint32_t)(prefix >> 32);
            bounded = true;
        }
        int nk = 0;
        if (bounded) {  // keys of the points at or above the bound, appended
            if (tid == 0) misc[1] = 0;
            __syncthreads();
            for (int64_t i0 = 0; i0 < n; i0 += 4 * kQNT) {  // uniform trip count (wave-level append)
                const int64_t i = i0 + 4 * tid;
                float e[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                int cnt = 0;
                if (i < n) {
                    if (vec4) {
                        const float4 e4 = est4(i);
                        e[0] = e4.x; e[1] = e4.y; e[2] = e4.z; e[3] = e4.w;
                    } else {
                        for (int u = 0; u < 4; ++u)
                            if (i + u < n) e[u] = est1(i + u);
                    }
                }
                uint32_t tm = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (i + u < n && fkey(e[u]) >= lkey) tm |= 1u << u;
                cnt = __popc(tm);
                uint32_t pos = wave_append_n<3>((uint32_t)cnt, reinterpret_cast<uint32_t*>(&misc[1]));
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if ((tm >> u) & 1u) est[pos++] = ckey(e[u], (uint32_t)(i + u));
            }
            __syncthreads();
            nk = misc[1];
            __syncthreads();
        }
        if (!bounded || nk < cq) {  // every point's key, in place (rare when bounded)
            for (int64_t i = 4 * (int64_t)tid; i < n; i += 4 * kQNT) {
                if (vec4) {
                    const float4 e4 = est4(i);
                    est[i] = ckey(e4.x, (uint32_t)i);
                    est[i + 1] = ckey(e4.y, (uint32_t)(i + 1));
                    est[i + 2] = ckey(e4.z, (uint32_t)(i + 2));
                    est[i + 3] = ckey(e4.w, (uint32_t)(i + 3));
                } else {
                    for (int u = 0; u < 4; ++u)
                        if (i + u < n) est[i + u] = ckey(est1(i + u), (uint32_t)(i + u));
                }
            }
            nk = (int)n;
        }
        __syncthreads();
        const int ce = select_candidates(est, nk, cq, sel, sh);
        rerank(qv, dm.d, rows, sel, ce, rk);
        const int64_t off = qb.out_off ? qb.out_off[q] : (int64_t)q * qb.k0;
        write_topk(rk, ce, kq, tk, sh, qb.out_id + off, qb.out_ip + off);
    }
}

hipError_t launch_est_search(const IndexDims& dm, const uint32_t* signs, const float* P, int64_t n,
                             const float* X, const QueryBatch& qb, int grid, uint64_t* scratch,
                             int64_t scratch_stride, int32_t* work, hipStream_t st) {
    if (qb.nq <= 0) return hipSuccess;
    hipError_t err = launch_fill_i32(work, 0, 1, st);
    if (err != hipSuccess) return err;
    int s_max = qb.s0, c_max = qb.c0, k_max = qb.k0;
    const int lds = est_search_lds_bytes(dm, s_max, c_max, k_max);
    grid = std::max(1, std::min(grid, qb.nq));
    CEOS_SWITCH_L(dm.L, {
        allow_lds(k_est_search<LL>, lds);
        hipLaunchKernelGGL(k_est_search<LL>, dim3(grid), dim3(kQNT), lds, st, dm, signs, P, n, X, qb, s_max,
                           c_max, k_max, scratch, scratch_stride, work);
    });
    return hipGetLastError();
}

}  // namespace ceos
