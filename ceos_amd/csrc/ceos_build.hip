// coCEOs index construction kernels (gfx950).
//
// Build = sample thresholds + one streaming scan + exact per-list selection, without
// materialising the n x (D*R) projection matrix the reference builds
// (src/CEOs.cpp:335-372):
//   1. project a strided sample of S points, column-major [DR][S];
//   2. per list (close/far of each column) take the sample value at a rank chosen so
//      that about 2.75*top_m of all n points pass it;
//   3. project all n points once more and append (id, value) of every point that passes
//      a list's threshold into that list's candidate segment;
//   4. per list, exact top-m of the candidates by (value desc, id asc), sorted.
// Step 4 returns exactly the reference's bucket (src/CEOs.cpp:377-432) whenever the
// candidate segment holds at least top_m entries, since every point outside it is
// below the threshold; the host verifies the counts and rescans lists that missed.
#include <algorithm>

#include <cfloat>

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

constexpr int kBuildNT = 256;
constexpr int kBuildWaves = kBuildNT / kWave;

__host__ __device__ inline int sign_words(int L) { return ((1 << L) + 31) / 32; }

// NW waves per workgroup, each projecting one (row, repeat) at a time.
template <int L, int NW>
__global__ __launch_bounds__(NW * kWave) void k_project_rows(IndexDims dm, const float* __restrict__ X,
                                                             int64_t first, int64_t stride, int64_t nrows,
                                                             const uint32_t* __restrict__ signs,
                                                             float* __restrict__ out, int64_t ld_out,
                                                             int col_major, const int32_t* __restrict__ nrows_dev) {
    using P = Proj<L>;
    extern __shared__ float lds_all[];
    const int wave = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    float* lds = lds_all + wave * P::kLdsWords;
    const int W = sign_words(L);
    // captured plans size the grid for their largest batch and read the actual row count here
    if (nrows_dev) nrows = min(nrows, (int64_t)max(0, nrows_dev[0]));
    const int64_t total = nrows * dm.R;
    for (int64_t p = (int64_t)blockIdx.x * NW + wave; p < total; p += (int64_t)gridDim.x * NW) {
        const int64_t k = p / dm.R;
        const int r = (int)(p - k * dm.R);
        const float* x = X + (first + k * stride) * dm.d;
        float xr[P::XE];
        P::load_x(x, dm.d, xr, ln);
        float y[P::OE];
        P::run(xr, signs + (size_t)r * 3 * W, lds, ln, y);
#pragma unroll
        for (int reg = 0; reg < P::OE; ++reg) {
            const int j = reg * kWave + ln;
            if (j < dm.D) {
                const int64_t col = (int64_t)r * dm.D + j;
                if (col_major) out[col * ld_out + k] = y[reg];
                else out[k * ld_out + col] = y[reg];
            }
        }
    }
}

template <int NW>
static hipError_t launch_project_rows_nw(const IndexDims& dm, const float* X, int64_t first, int64_t stride,
                                         int64_t nrows, const uint32_t* signs, float* out, int64_t ld_out,
                                         bool col_major, hipStream_t st, const int32_t* nrows_dev) {
    const int64_t waves = nrows * dm.R;
    const int grid = (int)std::min<int64_t>((waves + NW - 1) / NW, 256 * 64 / NW);
    CEOS_SWITCH_L(dm.L, {
        const size_t lds = (size_t)NW * Proj<LL>::kLdsWords * sizeof(float);
        allow_lds(k_project_rows<LL, NW>, lds);
        hipLaunchKernelGGL((k_project_rows<LL, NW>), dim3(grid), dim3(NW * kWave), lds, st, dm, X, first, stride,
                           nrows, signs, out, ld_out, (int)col_major, nrows_dev);
    });
    return hipGetLastError();
}

hipError_t launch_project_rows(const IndexDims& dm, const float* X, int64_t first, int64_t stride,
                               int64_t nrows, const uint32_t* signs, float* out, int64_t ld_out,
                               bool col_major, hipStream_t st, int waves_per_wg, const int32_t* nrows_dev) {
    if (nrows <= 0) return hipSuccess;
    switch (waves_per_wg) {
        case 2:
            return launch_project_rows_nw<2>(dm, X, first, stride, nrows, signs, out, ld_out, col_major, st, nrows_dev);
        case 8:
            return launch_project_rows_nw<8>(dm, X, first, stride, nrows, signs, out, ld_out, col_major, st, nrows_dev);
        default:
            return launch_project_rows_nw<4>(dm, X, first, stride, nrows, signs, out, ld_out, col_major, st, nrows_dev);
    }
}

// ---------------------------------------------------------------------------------
// out[c * ld_out + r] = in[r * cols + c], through 64 x 64 LDS tiles (coalesced both ways).
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in, int64_t rows, int64_t cols,
                                                   float* __restrict__ out, int64_t ld_out) {
    __shared__ float tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int i = ty; i < 64; i += 4) {
        const int64_t r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) tile[i][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int64_t c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) out[c * ld_out + r] = tile[tx][i];
    }
}

hipError_t launch_transpose(const float* in, int64_t rows, int64_t cols, float* out, hipStream_t st,
                            int64_t ld_out) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    if ((rows + 63) / 64 > 65535 || (cols + 63) / 64 > 0x7fffffff) return hipErrorInvalidValue;
    if (ld_out <= 0) ld_out = rows;
    if (ld_out < rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64)), dim3(256), 0,
                       st, in, rows, cols, out, ld_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Per column c of the column-major sample (S values): thr[2c] = the rank-th largest value
// (close list) and thr[2c+1] = the rank-th largest negated value (far list), duplicates
// counted. Each thread's largest value is one sample value, so the rank-th largest of the
// kBuildNT per-thread maxima is at most the rank-th largest value (rank <= kBuildNT);
// likewise for the negated values with the per-thread minima. A second read of the column
// collects every value at or above the first bound and every negated value at or above the
// second (a few hundred each) in LDS, and the rank-th largest of each collection is exact
// (a collection of exactly rank values: its smallest). A side whose collection exceeds
// kThrCollect falls back to a radix select over the column.
constexpr int kThrCollect = 2048;

__global__ __launch_bounds__(kBuildNT) void k_sample_thresholds(const float* __restrict__ sample, int64_t S,
                                                                int rank, float* __restrict__ thr) {
    __shared__ SelectShared sh;
    __shared__ float ext[2][kBuildNT];       // per-thread maxima, negated minima
    __shared__ float coll[2][kThrCollect];   // collected values, negated values
    __shared__ uint32_t counter[2];
    __shared__ uint32_t kmin[2];
    __shared__ float bound[2];
    const int col = blockIdx.x;
    const float* v = sample + (int64_t)col * S;
    float mx = -INFINITY, nmn = -INFINITY;  // largest value, largest negated value
    for (int64_t i = threadIdx.x; i < S; i += kBuildNT) {
        const float x = v[i];
        mx = fmaxf(mx, x);
        nmn = fmaxf(nmn, -x);
    }
    ext[0][threadIdx.x] = mx;
    ext[1][threadIdx.x] = nmn;
    if (threadIdx.x == 0) counter[0] = counter[1] = 0;
    if (threadIdx.x == 0) kmin[0] = kmin[1] = 0xffffffffu;
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        uint64_t prefix;
        int shift;
        const float* e = ext[side];
        radix_select<kBuildNT>(kBuildNT, rank, [&](int i) -> uint64_t { return (uint64_t)fkey(e[i]) << 32; }, sh,
                               prefix, shift, 32);
        if (threadIdx.x == 0) bound[side] = (shift >= 63) ? -INFINITY : fkey_inv((uint32_t)(prefix >> 32));
    }
    __syncthreads();
    const float b0 = bound[0], b1 = bound[1];
    for (int64_t i0 = 0; i0 < S; i0 += kBuildNT) {  // uniform trip count (wave-level appends)
        const int64_t i = i0 + threadIdx.x;
        const float x = i < S ? v[i] : 0.0f;
        const bool t0 = i < S && x >= b0;
        const bool t1 = i < S && -x >= b1;
        const uint32_t p0 = wave_append(t0, &counter[0]);
        const uint32_t p1 = wave_append(t1, &counter[1]);
        if (t0 && p0 < (uint32_t)kThrCollect) coll[0][p0] = x;
        if (t1 && p1 < (uint32_t)kThrCollect) coll[1][p1] = -x;
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        const int cnt = (int)counter[side];
        const float* c = coll[side];
        const bool neg = side == 1;
        uint64_t prefix;
        int shift;
        if (cnt == rank) {  // every collected value is in the top rank (cnt >= rank always)
            uint32_t mn = 0xffffffffu;
            for (int i = threadIdx.x; i < cnt; i += kBuildNT) mn = min(mn, fkey(c[i]));
            atomicMin(&kmin[side], mn);
            __syncthreads();
            prefix = (uint64_t)kmin[side] << 32;
            shift = 32;
        } else if (cnt <= kThrCollect) {
            radix_select<kBuildNT>(cnt, rank, [&](int i) -> uint64_t { return (uint64_t)fkey(c[i]) << 32; }, sh,
                                   prefix, shift, 32);
        } else {  // rare: too many values at the bound
            radix_select<kBuildNT>((int)S, rank,
                                   [&](int i) -> uint64_t { return (uint64_t)fkey(neg ? -v[i] : v[i]) << 32; }, sh,
                                   prefix, shift, 32);
        }
        if (threadIdx.x == 0) thr[2 * col + side] = (shift >= 63) ? -INFINITY : fkey_inv((uint32_t)(prefix >> 32));
    }
}

// Generic path (rank above kBuildNT): one workgroup per list, radix select over the column.
__global__ __launch_bounds__(kBuildNT) void k_sample_thresholds_radix(const float* __restrict__ sample, int64_t S,
                                                                      int rank, float* __restrict__ thr) {
    __shared__ SelectShared sh;
    const int list = blockIdx.x;
    const int col = list >> 1;
    const bool far = list & 1;
    const float* v = sample + (int64_t)col * S;
    auto get = [&](int i) -> uint64_t {
        const float x = far ? -v[i] : v[i];
        return (uint64_t)fkey(x) << 32;
    };
    uint64_t prefix;
    int shift;
    radix_select<kBuildNT>((int)S, rank, get, sh, prefix, shift, 32);
    if (threadIdx.x == 0) thr[list] = (shift >= 63) ? -INFINITY : fkey_inv((uint32_t)(prefix >> 32));
}

hipError_t launch_sample_thresholds(const float* sample, int64_t S, int DR, int rank, float* thr,
                                    hipStream_t st) {
    if (rank <= kBuildNT)
        hipLaunchKernelGGL(k_sample_thresholds, dim3(DR), dim3(kBuildNT), 0, st, sample, S, rank, thr);
    else
        hipLaunchKernelGGL(k_sample_thresholds_radix, dim3(2 * DR), dim3(kBuildNT), 0, st, sample, S, rank, thr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Hits (list, point, value) are staged per wave in LDS, placed by ballot + lane prefix
// (no atomics), and flushed when the stage may overflow: every lane then takes staged
// hits and reserves their positions with global atomics issued back to back, so one
// round trip serves up to kStage hits instead of one divergent round trip per hit.
constexpr int kStage = 256;  // staged hits per wave (5 workgroups of 4 waves fit a CU's LDS)

template <int L>
__global__ __launch_bounds__(kBuildNT) __attribute__((amdgpu_waves_per_eu(L <= 10 ? 5 : 1))) void k_scan_append(IndexDims dm, const float* __restrict__ X, int64_t n,
                                                          const uint32_t* __restrict__ signs,
                                                          const float* __restrict__ thr,
                                                          const uint8_t* __restrict__ active,
                                                          const int64_t* __restrict__ base,
                                                          const int32_t* __restrict__ cap,
                                                          int32_t* __restrict__ count, Entry* __restrict__ ent,
                                                          float* __restrict__ copy_to) {
    using P = Proj<L>;
    extern __shared__ float lds_all[];
    const int wave = threadIdx.x / kWave, ln = threadIdx.x % kWave;
    float* lds = lds_all + wave * P::kLdsWords;
    // per-wave stage after the FWHT tiles: list, point id, value bits
    uint32_t* st_list = reinterpret_cast<uint32_t*>(lds_all + kBuildWaves * P::kLdsWords) + wave * 3 * kStage;
    uint32_t* st_id = st_list + kStage;
    uint32_t* st_val = st_id + kStage;
    // Workgroup blockIdx.x = block * R + repeat: the R workgroups of a point block are
    // dispatched together (on different XCDs) and read the same rows at about the same
    // time, so the memory-side cache serves the repeats after the first instead of R
    // sweeps over the points.
    const int r = (int)(blockIdx.x % (unsigned)dm.R);
    const int64_t bx = blockIdx.x / (unsigned)dm.R, gxr = gridDim.x / (unsigned)dm.R;
    const int W = sign_words(L);
    const uint32_t* sg = signs + (size_t)r * 3 * W;
    const typename P::Signs sw = P::load_signs(sg, ln);

    // an inactive list (or a column past n_proj) gets a NaN threshold: no value passes it
    float th_c[P::OE], th_f[P::OE];
    bool any_active = false;
#pragma unroll
    for (int reg = 0; reg < P::OE; ++reg) {
        const int j = reg * kWave + ln;
        th_c[reg] = th_f[reg] = __builtin_nanf("");
        if (j < dm.D) {
            const int c = r * dm.D + j;
            if (!active || active[2 * c]) { th_c[reg] = thr[2 * c]; any_active = true; }
            if (!active || active[2 * c + 1]) { th_f[reg] = thr[2 * c + 1]; any_active = true; }
        }
    }
    if (!__any(any_active) && !copy_to) return;

    uint32_t staged = 0;  // wave-uniform
    auto flush = [&]() {
        wave_sync();
        constexpr int U = kStage / kWave;
        int32_t pos[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {  // reserve: atomics in flight together
            const uint32_t i = (uint32_t)(u * kWave + ln);
            pos[u] = i < staged ? atomicAdd(&count[st_list[i]], 1) : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)(u * kWave + ln);
            if (i < staged) {
                const uint32_t list = st_list[i];
                if (pos[u] < cap[list]) ent[base[list] + pos[u]] = make_uint2(st_id[i], st_val[i]);
            }
        }
        wave_sync();
        staged = 0;
    };

    // pr = p mod R, kept up to date without 64-bit divisions
    const int64_t pstep = gxr * kBuildWaves;
    const int rstep = (int)(pstep % dm.R);
    int pr = (int)((bx * kBuildWaves + wave) % dm.R);
    for (int64_t p = bx * kBuildWaves + wave; p < n; p += pstep) {
        float xr[P::XE];
        P::load_x(X + p * dm.d, dm.d, xr, ln);
        if (copy_to && pr == r) P::store_x(copy_to + p * dm.d, dm.d, xr, ln);  // wave-uniform
        pr += rstep;
        if (pr >= dm.R) pr -= dm.R;
        float y[P::OE];
        P::run(xr, sw, lds, ln, y);
#pragma unroll
        for (int reg = 0; reg < P::OE; ++reg) {
            const float v = y[reg];
            const uint32_t c = (uint32_t)(r * dm.D + reg * kWave + ln);
            const bool hc = v >= th_c[reg];
            const bool hf = -v >= th_f[reg];
            const uint64_t mc = __ballot(hc), mf = __ballot(hf);
            if (mc | mf) {  // wave-uniform
                if (staged + (uint32_t)(__popcll(mc) + __popcll(mf)) > (uint32_t)kStage) flush();  // <= 128 fit after
                if (hc) {
                    const uint32_t i = staged + lanes_below(mc);
                    st_list[i] = 2 * c;
                    st_id[i] = (uint32_t)p;
                    st_val[i] = __float_as_uint(v);
                }
                staged += (uint32_t)__popcll(mc);
                if (hf) {
                    const uint32_t i = staged + lanes_below(mf);
                    st_list[i] = 2 * c + 1;
                    st_id[i] = (uint32_t)p;
                    st_val[i] = __float_as_uint(-v);
                }
                staged += (uint32_t)__popcll(mf);
            }
        }
    }
    if (staged > 0) flush();
}

hipError_t launch_scan_append(const IndexDims& dm, const float* X, int64_t n, const uint32_t* signs,
                              const float* thr, const uint8_t* active, const int64_t* base,
                              const int32_t* cap, int32_t* count, Entry* ent, hipStream_t st, float* copy_to) {
    if (n <= 0) return hipSuccess;
    int gx = (int)std::min<int64_t>((n + kBuildWaves - 1) / kBuildWaves, std::max(1, 2048 / dm.R) * 4);
    CEOS_SWITCH_L(dm.L, {
        const size_t lds = (size_t)kBuildWaves * (Proj<LL>::kLdsWords * sizeof(float) + 3 * kStage * 4);
        allow_lds(k_scan_append<LL>, lds);
        hipLaunchKernelGGL(k_scan_append<LL>, dim3((unsigned)(gx * dm.R)), dim3(kBuildNT), lds, st, dm, X, n, signs,
                           thr, active, base, cap, count, ent, copy_to);
    });
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kBuildNT) void k_list_select(int m, const int64_t* __restrict__ base,
                                                          const int32_t* __restrict__ cap,
                                                          const int32_t* __restrict__ count,
                                                          const Entry* __restrict__ ent, Entry* __restrict__ bkt,
                                                          int32_t* __restrict__ size) {
    __shared__ SelectShared sh;
    extern __shared__ uint64_t keys[];
    const int list = blockIdx.x;
    const int avail = min(count[list], cap[list]);
    const Entry* e = ent + base[list];
    const int K = min(m, avail);
    auto get = [&](int i) -> uint64_t { return ckey(__uint_as_float(e[i].y), e[i].x); };
    uint64_t prefix;
    int shift;
    radix_select<kBuildNT>(avail, max(K, 1), get, sh, prefix, shift);
    const int N = next_pow2(max(K, 1));
    for (int i = threadIdx.x; i < N; i += kBuildNT) keys[i] = 0;
    if (threadIdx.x == 0) sh.counter = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < avail; i += kBuildNT) {
        const uint64_t k = get(i);
        if (selected(k, prefix, shift)) keys[atomicAdd(&sh.counter, 1u)] = k;
    }
    __syncthreads();
    bitonic_sort_desc<kBuildNT>(keys, N);
    for (int t = threadIdx.x; t < K; t += kBuildNT) {
        const uint64_t k = keys[t];
        bkt[(int64_t)list * m + t] = make_uint2(ckey_id(k), __float_as_uint(ckey_val(k)));
    }
    if (threadIdx.x == 0) size[list] = K;
}

hipError_t launch_list_select(int nlists, int m, const int64_t* base, const int32_t* cap,
                              const int32_t* count, const Entry* ent, Entry* bkt, int32_t* size,
                              hipStream_t st) {
    const size_t lds = (size_t)next_pow2(m) * sizeof(uint64_t);
    allow_lds(k_list_select, lds);
    hipLaunchKernelGGL(k_list_select, dim3(nlists), dim3(kBuildNT), lds, st, m, base, cap, count, ent, bkt,
                       size);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Stream update merge (src/streamCEOs.cpp:267-358 restated on device-resident buckets).
__global__ __launch_bounds__(kBuildNT) void k_stream_merge(int m, Entry* __restrict__ bkt,
                                                           int32_t* __restrict__ size,
                                                           const float* __restrict__ newproj, int64_t n_new,
                                                           uint32_t first_abs, uint32_t new_base,
                                                           uint32_t live) {
    __shared__ SelectShared sh;
    __shared__ uint32_t s_valid;
    extern __shared__ uint64_t keys[];
    const int list = blockIdx.x;
    const int col = list >> 1;
    const bool far = list & 1;
    Entry* b = bkt + (int64_t)list * m;
    const int sz = size[list];
    const float* np = newproj + (int64_t)col * n_new;
    if (threadIdx.x == 0) s_valid = 0;
    __syncthreads();
    uint32_t myvalid = 0;
    for (int t = threadIdx.x; t < sz; t += kBuildNT) myvalid += (b[t].x - new_base) < live;
    atomicAdd(&s_valid, myvalid);
    __syncthreads();
    const int64_t total = (int64_t)sz + n_new;
    const int K = (int)min<int64_t>(m, (int64_t)s_valid + n_new);
    auto get = [&](int i) -> uint64_t {
        if (i < sz) {
            const Entry e = b[i];
            const uint32_t lid = e.x - new_base;
            return lid < live ? ckey(__uint_as_float(e.y), lid) : 0ull;
        }
        const int64_t j = i - sz;
        const float v = np[j];
        const uint32_t lid = first_abs + (uint32_t)j - new_base;
        return ckey(far ? -v : v, lid);
    };
    uint64_t prefix;
    int shift;
    radix_select<kBuildNT>((int)total, max(K, 1), get, sh, prefix, shift);
    const int N = next_pow2(max(K, 1));
    for (int i = threadIdx.x; i < N; i += kBuildNT) keys[i] = 0;
    if (threadIdx.x == 0) sh.counter = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += kBuildNT) {
        const uint64_t k = get(i);
        if (k != 0ull && selected(k, prefix, shift)) keys[atomicAdd(&sh.counter, 1u)] = k;
    }
    __syncthreads();
    bitonic_sort_desc<kBuildNT>(keys, N);
    for (int t = threadIdx.x; t < K; t += kBuildNT) {
        const uint64_t k = keys[t];
        b[t] = make_uint2(ckey_id(k) + new_base, __float_as_uint(ckey_val(k)));
    }
    if (threadIdx.x == 0) size[list] = K;
}

hipError_t launch_stream_merge(int DR, int m, Entry* bkt, int32_t* size, const float* newproj,
                               int64_t n_new, uint32_t first_abs, uint32_t new_base, uint32_t live,
                               hipStream_t st) {
    const size_t lds = (size_t)next_pow2(m) * sizeof(uint64_t);
    allow_lds(k_stream_merge, lds);
    hipLaunchKernelGGL(k_stream_merge, dim3(2 * DR), dim3(kBuildNT), lds, st, m, bkt, size, newproj, n_new,
                       first_abs, new_base, live);
    return hipGetLastError();
}

// Stable in-order removal of entries whose ring slot is flagged dead (one wave per list).
__global__ __launch_bounds__(kWave) void k_bucket_purge(int m, Entry* __restrict__ bkt, int32_t* __restrict__ size,
                                                        const uint32_t* __restrict__ dead, RowRing ring) {
    const int list = blockIdx.x;
    const int ln = threadIdx.x;
    Entry* b = bkt + (int64_t)list * m;
    const int sz = size[list];
    int out = 0;
    for (int t0 = 0; t0 < sz; t0 += kWave) {
        const int t = t0 + ln;
        Entry e = make_uint2(0, 0);
        bool keep = false;
        if (t < sz) {
            e = b[t];
            int64_t slot = (int64_t)(e.x - ring.base) + ring.head;
            if (slot >= ring.cap) slot -= ring.cap;
            keep = !((dead[slot >> 5] >> (slot & 31)) & 1u);
        }
        const uint64_t mask = __ballot(keep);
        const int before = __popcll(mask & ((1ull << ln) - 1ull));
        wave_sync();
        if (keep) b[out + before] = e;
        out += __popcll(mask);
        wave_sync();
    }
    if (ln == 0) size[list] = out;
}

// Sets the dead bit of each listed id's ring slot (ids are live logical ids, checked by
// the caller; repeats are harmless).
__global__ __launch_bounds__(256) void k_mark_dead(const int32_t* __restrict__ ids, int64_t count, int64_t head,
                                                   int64_t cap, uint32_t* __restrict__ dead) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += step) {
        int64_t sl = head + ids[i];
        if (sl >= cap) sl -= cap;
        atomicOr(&dead[sl >> 5], 1u << (sl & 31));
    }
}

hipError_t launch_mark_dead(const int32_t* ids, int64_t count, int64_t head, int64_t cap, uint32_t* dead,
                            hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(k_mark_dead, dim3(grid), dim3(256), 0, st, ids, count, head, cap, dead);
    return hipGetLastError();
}

hipError_t launch_bucket_purge(int nlists, int m, Entry* bkt, int32_t* size, const uint32_t* dead,
                               RowRing ring, hipStream_t st) {
    hipLaunchKernelGGL(k_bucket_purge, dim3(nlists), dim3(kWave), 0, st, m, bkt, size, dead, ring);
    return hipGetLastError();
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Clears the bits of ring slots (slot0 + j) mod cap, j < count (bitmap of 32-bit words).
__global__ void k_clear_slot_bits(uint32_t* __restrict__ bits, int64_t slot0, int64_t count, int64_t cap) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t sl = slot0 + j;
        if (sl >= cap) sl -= cap;
        atomicAnd(&bits[sl >> 5], ~(1u << (sl & 31)));
    }
}

hipError_t launch_clear_slot_bits(uint32_t* bits, int64_t slot0, int64_t count, int64_t cap, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(k_clear_slot_bits, dim3(grid), dim3(256), 0, st, bits, slot0, count, cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Centering (src/streamCEOs.cpp:22-34): the column mean of rows [0, n). Pass 1: block b
// sums its rows [b rpb, (b+1) rpb) of every column in FP64, in row order (thread t takes
// columns t, t + 256, ...; a wave reads 64 consecutive floats of a row). Pass 2: one
// thread per column adds the block sums in block order and rounds sum / n to FP32 once.
// The order is fixed, so the centre does not depend on the launch.
__global__ __launch_bounds__(256) void k_column_partials(const float* __restrict__ X, int64_t n, int d,
                                                         int64_t rpb, double* __restrict__ partial) {
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    const int64_t r1 = min(n, r0 + rpb);
    for (int j = threadIdx.x; j < d; j += 256) {
        double s = 0.0;
        int64_t r = r0;
        for (; r + 4 <= r1; r += 4) {  // four independent loads in flight, added in row order
            const float a = X[r * d + j], b = X[(r + 1) * d + j], c = X[(r + 2) * d + j], e = X[(r + 3) * d + j];
            s += (double)a;
            s += (double)b;
            s += (double)c;
            s += (double)e;
        }
        for (; r < r1; ++r) s += (double)X[r * d + j];
        partial[(int64_t)blockIdx.x * d + j] = s;
    }
}

__global__ __launch_bounds__(256) void k_column_mean(const double* __restrict__ partial, int nblocks, int d,
                                                     int64_t n, float* __restrict__ center) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= d) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * d + j];
    center[j] = (float)(s / (double)n);
}

int column_mean_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 1023) / 1024)); }

hipError_t launch_column_mean(const float* X, int64_t n, int d, double* partial, float* center, hipStream_t st) {
    if (n <= 0 || d <= 0) return hipErrorInvalidValue;
    const int nb = column_mean_blocks(n);
    const int64_t rpb = (n + nb - 1) / nb;
    hipLaunchKernelGGL(k_column_partials, dim3(nb), dim3(256), 0, st, X, n, d, rpb, partial);
    hipLaunchKernelGGL(k_column_mean, dim3((d + 255) / 256), dim3(256), 0, st, partial, nb, d, n, center);
    return hipGetLastError();
}

// x -= center (FP32) for ring slots (slot0 + i) mod cap, i < count.
__global__ __launch_bounds__(256) void k_subtract_center(float* __restrict__ X, int64_t slot0, int64_t count,
                                                         int64_t cap, int d, const float* __restrict__ center) {
    const int64_t total = count * d;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / d;
        const int j = (int)(t - i * d);
        int64_t slot = slot0 + i;
        if (slot >= cap) slot -= cap;
        X[slot * d + j] -= center[j];
    }
}

hipError_t launch_subtract_center(float* X, int64_t slot0, int64_t count, int64_t cap, int d, const float* center,
                                  hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((count * d + 255) / 256, 8192);
    hipLaunchKernelGGL(k_subtract_center, dim3(grid), dim3(256), 0, st, X, slot0, count, cap, d, center);
    return hipGetLastError();
}

// Rerank bounds: an int8 copy c of each row and (s, r1, r2, 0) per slot, one wave per
// row. s = max_i |x_i| / 127 and c_i = round(x_i / s), clamped to [-127, 127];
//   r1 = |x - s c|_2 + g |x|_2,  r2 = s |c|_2,  g = n u / (1 - n u), u = 2^-24, n = d + 64,
// evaluated in FP64 and rounded up to FP32. For a query q, the FP32 inner product v the
// rerank forms (a summation of depth below n) satisfies |v - <q,x>| <= g |q||x|, and
// |<q,x> - s <q,c>| <= |q| |x - s c| (Cauchy-Schwarz), so v is within |q| r1 of s <q,c>;
// r2 scales the error of an approximation of <q,c> (rounding, or a coded query q^:
// |s <q - q^, c>| <= |q - q^| r2). A row with a non-finite entry gets s = 0, c = 0,
// r1 = +inf, r2 = 0.
// G = lanes per row (the power of two covering d/4 float4 units, at most 64); a lane
// keeps its first four units in registers between the two passes over the row.
__host__ __device__ inline int quantize_lanes(int d) {
    int G = 1;
    while (G < (d >> 2) && G < kWave) G <<= 1;
    return G;
}

__global__ __launch_bounds__(256) void k_quantize_rows(const float* __restrict__ X, int64_t slot0, int64_t count,
                                                       int64_t cap, int d, int8_t* __restrict__ Xq) {
    const int ln = threadIdx.x % kWave;
    const int units = d >> 2;
    const int G = quantize_lanes(d);
    const int gl = ln % G;
    const int64_t j = ((int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * (kWave / G) + ln / G;
    const bool ok = j < count;  // lanes of a missing row still join the group reductions
    int64_t sl = slot0 + (ok ? j : 0);
    if (sl >= cap) sl -= cap;
    const float4* x4 = reinterpret_cast<const float4*>(X + sl * d);
    int8_t* xrow = Xq + sl * quant_stride(d);
    uint32_t* c4 = reinterpret_cast<uint32_t*>(xrow);
    constexpr int KEEP = 4;
    float4 keep[KEEP];
    float mx = 0.0f;
    bool fin = true;
    auto see = [&](const float4 v) {
        const float a = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
        fin &= a <= FLT_MAX;  // false for inf and NaN
        mx = fmaxf(mx, a);
    };
#pragma unroll
    for (int k = 0; k < KEEP; ++k) {
        const int uu = gl + k * G;
        keep[k] = (ok && uu < units) ? x4[uu] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        see(keep[k]);
    }
    if (ok)
        for (int uu = gl + KEEP * G; uu < units; uu += G) see(x4[uu]);
    int bad = fin ? 0 : 1;
    for (int o = G >> 1; o > 0; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
        bad |= __shfl_xor(bad, o, kWave);
    }
    fin = bad == 0;
    const float s = (fin && mx > 0.0f) ? mx / 127.0f : 0.0f;
    const float inv = s > 0.0f ? 127.0f / mx : 0.0f;
    double e2 = 0.0, x2 = 0.0;
    int c2 = 0;
    auto code = [&](float v) {
        int ci = 0;
        if (s > 0.0f) ci = min(127, max(-127, (int)rintf(v * inv)));
        const double dv = (double)v - (double)s * (double)ci;
        e2 = fma(dv, dv, e2);
        x2 = fma((double)v, (double)v, x2);
        c2 += ci * ci;
        return (uint32_t)ci & 0xffu;
    };
    auto put = [&](int uu, const float4 v) {
        const uint32_t w = code(v.x) | (code(v.y) << 8) | (code(v.z) << 16) | (code(v.w) << 24);
        c4[uu] = w;
    };
    if (ok) {
#pragma unroll
        for (int k = 0; k < KEEP; ++k)
            if (gl + k * G < units) put(gl + k * G, keep[k]);
        for (int uu = gl + KEEP * G; uu < units; uu += G) put(uu, x4[uu]);
    }
    for (int o = G >> 1; o > 0; o >>= 1) {
        e2 += __shfl_xor(e2, o, kWave);
        x2 += __shfl_xor(x2, o, kWave);
        c2 += __shfl_xor(c2, o, kWave);
    }
    if (ok && gl == 0) {
        auto up = [](double v) {
            float f = (float)v;
            if ((double)f < v) f = nextafterf(f, INFINITY);
            return f;
        };
        float r1 = INFINITY, r2 = 0.0f;
        if (fin) {
            const double n = (double)d + 64.0, u = 0x1p-24;
            const double g = n * u / (1.0 - n * u);
            r1 = up((sqrt(e2) + g * sqrt(x2)) * (1.0 + 0x1p-30));
            r2 = up((double)s * sqrt((double)c2) * (1.0 + 0x1p-30));
        }
        *reinterpret_cast<float4*>(xrow + quant_param_offset(d)) = make_float4(s, r1, r2, 0.0f);
    }
}

hipError_t launch_quantize_rows(const float* X, int64_t slot0, int64_t count, int64_t cap, int d, int8_t* Xq,
                                hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (count > cap || slot0 < 0 || slot0 >= cap || d <= 0 || (d & 3) != 0) return hipErrorInvalidValue;
    const int64_t per_block = 4 * (kWave / quantize_lanes(d));
    const int64_t blocks = (count + per_block - 1) / per_block;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_quantize_rows, dim3((unsigned)blocks), dim3(256), 0, st, X, slot0, count, cap, d, Xq);
    return hipGetLastError();
}

// Streaming-read calibration: every thread sums 16-byte loads over a grid-stride loop
// (four in flight) and writes one float. n4 = number of float4 elements.
// Plain streaming read: grid-stride 16-byte loads, 8 in flight per thread, marked
// non-temporal (nothing is read twice). Non-temporal loads with 8 in flight read about
// 10% faster than ordinary ones with 4 in flight on gfx950.
__global__ __launch_bounds__(256) void k_stream_read(const float4* __restrict__ x, int64_t n4,
                                                     float* __restrict__ out) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    const F4* __restrict__ xv = reinterpret_cast<const F4*>(x);
    constexpr int U = 8;
    F4 s = {0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n4; i += U * stride) {
        F4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(xv + i + u * stride);
#pragma unroll
        for (int u = 0; u < U; ++u) s += v[u];
    }
    for (; i < n4; i += stride) s += __builtin_nontemporal_load(xv + i);
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = (s.x + s.y) + (s.z + s.w);
}

hipError_t launch_stream_read(const float* x, int64_t n, float* out, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, st, reinterpret_cast<const float4*>(x), n / 4, out);
    return hipGetLastError();
}

// Segment blockIdx.y; 16-byte words when the segment allows, else 4-byte words.
__global__ __launch_bounds__(256) void k_copy_segments(CopySegs c) {
    const int g = blockIdx.y;
    const uint32_t* __restrict__ src = c.src[g];
    uint32_t* __restrict__ dst = c.dst[g];
    const int64_t n = c.words[g];
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool v4 = (n & 3) == 0 && ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
    if (v4) {
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        for (int64_t i = t0; i < (n >> 2); i += step) d4[i] = s4[i];
    } else {
        for (int64_t i = t0; i < n; i += step) dst[i] = src[i];
    }
}

hipError_t launch_copy_segments(const CopySegs& c, hipStream_t st) {
    if (c.n <= 0) return hipSuccess;
    if (c.n > 4) return hipErrorInvalidValue;
    int64_t most = 0;
    for (int g = 0; g < c.n; ++g) {
        if (c.words[g] < 0 || (c.words[g] > 0 && (!c.src[g] || !c.dst[g]))) return hipErrorInvalidValue;
        most = std::max(most, c.words[g]);
    }
    if (most == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(1024, (most + 4 * 256 - 1) / (4 * 256));
    hipLaunchKernelGGL(k_copy_segments, dim3((unsigned)blocks, (unsigned)c.n), dim3(256), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

}  // namespace ceos
