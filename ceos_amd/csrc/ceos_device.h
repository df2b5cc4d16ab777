// Device building blocks shared by the coCEOs kernels (gfx950, wave64).
//
//  * Proj<L>: sign-flip + fast Walsh-Hadamard projection of one vector per wave,
//    F = 2^L coordinates, 3 rounds per repeat. Butterflies run in registers; the
//    data is re-laid out between register phases through a per-wave LDS tile so
//    that every butterfly stage is register-local. Stage order is ascending bit
//    index (stride 1, 2, 4, ...), the order of the reference's fht_float
//    (src/fht_avx.c:1012-1252), so the result is bit-identical to it.
//  * radix_select: block-wide exact top-K over 64-bit keys (8-bit digits, early
//    exit once the boundary digit bucket is taken whole).
//  * bitonic_sort_desc: block-wide descending sort of an LDS array (rank counting up
//    to the block size, a bitonic network above).
//  * Keys: fkey() maps an FP32 value to a uint32 whose unsigned order is the
//    float order; composite keys put the value in the high word and ~id in the
//    low word, so descending key order is (value desc, id asc) and keys are unique.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ceos {

constexpr int kWave = 64;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t fkey(float v) {
    uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}
__device__ __forceinline__ uint64_t ckey(float v, uint32_t id) {
    return ((uint64_t)fkey(v) << 32) | (uint64_t)(~id);
}
__device__ __forceinline__ uint32_t ckey_id(uint64_t k) { return ~(uint32_t)k; }
__device__ __forceinline__ float ckey_val(uint64_t k) { return fkey_inv((uint32_t)(k >> 32)); }

__device__ __forceinline__ float flip_if_zero_bit(float y, uint32_t bits, int r) {
    // reference multiplies by (2*bit - 1): bit 0 -> -1 (sign flip), bit 1 -> +1
    return __uint_as_float(__float_as_uint(y) ^ ((((~bits) >> r) & 1u) << 31));
}

// ----------------------------------------------------------------------------------
// Projection of one vector by one wave.
//
// Register path (L >= 8): E = F/64 values per lane. In the phase whose register bits
// start at bit b, lane `ln` register `r` holds element
//     elem(ln, r, b) = (ln & (2^b - 1)) | (r << b) | ((ln >> b) << (b + e)),  e = L - 6.
// Phases cover bits [0,e), [e,2e), ... and a last phase with b = L - e, whose layout
// is element = r*64 + ln: lane-contiguous, so stores of the output are coalesced.
// The first phase (b = 0) gives each lane E consecutive input coordinates.
//
// LDS path (L <= 7): the vector lives in LDS, one butterfly per lane per stage.
// ----------------------------------------------------------------------------------
template <int L, bool kRegs = (L >= 8)>
struct Proj;

template <int L>
struct Proj<L, true> {
    static constexpr int F = 1 << L;
    static constexpr int e = L - 6;
    static constexpr int E = 1 << e;
    static constexpr int XE = E;        // input registers per lane
    static constexpr int OE = E;        // output registers per lane (element r*64+lane)
    static constexpr int kLdsWords = F + F / E;  // per-wave tile with one pad word per E

    __device__ static __forceinline__ int pad(int i) { return i + (i >> e); }
    __device__ static __forceinline__ int elem(int ln, int r, int b) {
        return (ln & ((1 << b) - 1)) | (r << b) | ((ln >> b) << (b + e));
    }

    // x (d <= F floats) -> registers in the b = 0 layout (zero padded).
    __device__ static void load_x(const float* __restrict__ x, int d, float (&xr)[XE], int ln) {
        const int i0 = ln * E;
        // uniform fast path: with d a multiple of E and x 16-byte aligned, each lane's E
        // coordinates are all inside the row or all past it (one branch for the wave)
        if (E % 4 == 0 && d % E == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0) {
            if (i0 < d) {
#pragma unroll
                for (int r = 0; r < E; r += 4) {
                    float4 v = *reinterpret_cast<const float4*>(x + i0 + r);
                    xr[r] = v.x; xr[r + 1] = v.y; xr[r + 2] = v.z; xr[r + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int r = 0; r < E; ++r) xr[r] = 0.0f;
            }
            return;
        }
        if (i0 + E <= d && ((reinterpret_cast<uintptr_t>(x + i0) & 15) == 0) && (E % 4 == 0)) {
#pragma unroll
            for (int r = 0; r < E; r += 4) {
                float4 v = *reinterpret_cast<const float4*>(x + i0 + r);
                xr[r] = v.x; xr[r + 1] = v.y; xr[r + 2] = v.z; xr[r + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int r = 0; r < E; ++r) xr[r] = (i0 + r < d) ? x[i0 + r] : 0.0f;
        }
    }

    // registers in the load_x layout -> x (the first d elements), non-temporal: a copy
    // that is not read back soon should not evict rows other waves are about to read
    __device__ static void store_x(float* __restrict__ x, int d, const float (&xr)[XE], int ln) {
        const int i0 = ln * E;
        if (E % 4 == 0 && d % E == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0) {
            if (i0 < d) {
#pragma unroll
                for (int r = 0; r < E; r += 4) {
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    const f4 v = {xr[r], xr[r + 1], xr[r + 2], xr[r + 3]};
                    __builtin_nontemporal_store(v, reinterpret_cast<f4*>(x + i0 + r));
                }
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < E; ++r)
            if (i0 + r < d) __builtin_nontemporal_store(xr[r], x + i0 + r);
    }

    template <int b, int t0, int t1>
    __device__ static __forceinline__ void stages(float (&y)[E]) {
#pragma unroll
        for (int t = t0; t < t1; ++t) {
            const int h = 1 << (t - b);
#pragma unroll
            for (int r = 0; r < E; ++r) {
                if (!(r & h)) {
                    const float u = y[r], v = y[r + h];
                    y[r] = u + v;
                    y[r + h] = u - v;
                }
            }
        }
    }

    // pad(elem(ln, r, b)) = pad(elem(ln, 0, b)) + r * step(b): the register index occupies
    // element bits [b, b + e), which are clear in elem(ln, 0, b), and for b < e the lane
    // bits sit at and above bit e, so the pad term (i >> e) splits the same way. One base
    // address per layout and immediate offsets per register.
    __device__ static __forceinline__ int step(int b) { return (1 << b) + ((1 << b) >> e); }
    __device__ static __forceinline__ void relayout(float (&y)[E], float* lds, int ln, int bf, int bt) {
        float* wf = lds + pad(elem(ln, 0, bf));
        const int sf = step(bf);
#pragma unroll
        for (int r = 0; r < E; ++r) wf[r * sf] = y[r];
        wave_sync();
        const float* rt = lds + pad(elem(ln, 0, bt));
        const int stt = step(bt);
#pragma unroll
        for (int r = 0; r < E; ++r) y[r] = rt[r * stt];
        wave_sync();
    }

    // Relayout from phase bp to the last phase (b = L - e = 6) in registers, for bp >= 4:
    // in layout bp, lane bit j in [bp, 6) holds element bit j + e and register bit j - bp
    // holds element bit j; in layout 6 those trade places. v_permlane16_swap exchanges
    // rows 1 and 3 of its first operand with rows 0 and 2 of its second (lane bit 4 against
    // the register pair's bit), v_permlane32_swap the upper half of the first with the lower
    // half of the second (lane bit 5). Afterwards register bit i < s = 6 - bp holds element
    // bit bp + e + i and register bit i >= s element bit bp + i, so the register index is
    // rotated right by s bits (a renaming, no instructions).
    template <int bp>
    __device__ static __forceinline__ void relayout_swap(float (&y)[E]) {
        constexpr int s = (L - e) - bp;
#pragma unroll
        for (int j = bp; j < L - e; ++j) {
            const int hb = 1 << (j - bp);
#pragma unroll
            for (int r = 0; r < E; ++r) {
                if (!(r & hb)) {
                    const auto t = j == 4
                        ? __builtin_amdgcn_permlane16_swap(__float_as_uint(y[r]), __float_as_uint(y[r + hb]), false, false)
                        : __builtin_amdgcn_permlane32_swap(__float_as_uint(y[r]), __float_as_uint(y[r + hb]), false, false);
                    y[r] = __uint_as_float(t[0]);
                    y[r + hb] = __uint_as_float(t[1]);
                }
            }
        }
        float t[E];
#pragma unroll
        for (int r = 0; r < E; ++r) t[(r >> s) | ((r & ((1 << s) - 1)) << (e - s))] = y[r];
#pragma unroll
        for (int r = 0; r < E; ++r) y[r] = t[r];
    }

    template <int covered, int bprev>
    struct Phase {
        static constexpr int b = (covered + e <= L) ? covered : (L - e);
        __device__ static __forceinline__ void run(float (&y)[E], float* lds, int ln) {
            if constexpr (covered > 0 && b == L - e && bprev >= 4)
                relayout_swap<bprev>(y);
            else if (covered > 0)
                relayout(y, lds, ln, bprev, b);
            stages<b, covered, b + e>(y);
            Phase<b + e, b>::run(y, lds, ln);
        }
    };
    template <int bprev>
    struct Phase<L, bprev> {
        __device__ static __forceinline__ void run(float (&)[E], float*, int) {}
    };

    // Lane's E sign bits of round `i` (words of 32 bits, bit j of word w = coordinate 32w+j).
    __device__ static __forceinline__ void sign_bits(const uint32_t* __restrict__ sg, int ln, uint32_t (&w)[2]) {
        if (E >= 64) {
            w[0] = sg[ln * 2]; w[1] = sg[ln * 2 + 1];
        } else {
            const int i0 = ln * E;
            w[0] = (sg[i0 >> 5] >> (i0 & 31));
            w[1] = 0;
        }
    }

    // The lane's sign words of the 3 rounds (loop-invariant for a repeat: load once).
    struct Signs {
        uint32_t w[3][2];
    };
    __device__ static __forceinline__ Signs load_signs(const uint32_t* __restrict__ sg, int ln) {
        Signs s;
#pragma unroll
        for (int rd = 0; rd < 3; ++rd) sign_bits(sg + rd * (F / 32), ln, s.w[rd]);
        return s;
    }

    // Output y[r] = coordinate r*64 + ln.
    __device__ static void run(const float (&xr)[XE], const Signs& sw, float* lds, int ln, float (&y)[OE]) {
#pragma unroll
        for (int r = 0; r < E; ++r) y[r] = xr[r];
#pragma unroll 1
        for (int rd = 0; rd < 3; ++rd) {
            if (rd > 0) relayout(y, lds, ln, L - e, 0);
            uint32_t w[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) w[k] = rd == 0 ? sw.w[0][k] : rd == 1 ? sw.w[1][k] : sw.w[2][k];
#pragma unroll
            for (int r = 0; r < E; ++r) y[r] = flip_if_zero_bit(y[r], w[r >> 5], r & 31);
            Phase<0, 0>::run(y, lds, ln);
        }
    }

    // sg: 3 rounds x F/32 words for this repeat.
    __device__ static void run(const float (&xr)[XE], const uint32_t* __restrict__ sg, float* lds, int ln,
                               float (&y)[OE]) {
        run(xr, load_signs(sg, ln), lds, ln, y);
    }
};

template <int L>
struct Proj<L, false> {
    static constexpr int F = 1 << L;
    static constexpr int OE = (F + 63) / 64;
    static constexpr int XE = OE;
    static constexpr int kLdsWords = F;

    __device__ static void load_x(const float* __restrict__ x, int d, float (&xr)[XE], int ln) {
#pragma unroll
        for (int r = 0; r < XE; ++r) {
            const int i = r * 64 + ln;
            xr[r] = (i < d) ? x[i] : 0.0f;
        }
    }

    __device__ static void store_x(float* __restrict__ x, int d, const float (&xr)[XE], int ln) {
#pragma unroll
        for (int r = 0; r < XE; ++r) {
            const int i = r * 64 + ln;
            if (i < d) __builtin_nontemporal_store(xr[r], x + i);
        }
    }

    // (the LDS path reads its sign words per round)
    struct Signs {
        const uint32_t* sg;
    };
    __device__ static __forceinline__ Signs load_signs(const uint32_t* __restrict__ sg, int) { return Signs{sg}; }
    __device__ static void run(const float (&xr)[XE], const Signs& sw, float* lds, int ln, float (&y)[OE]) {
        run(xr, sw.sg, lds, ln, y);
    }

    __device__ static void run(const float (&xr)[XE], const uint32_t* __restrict__ sg, float* lds, int ln,
                               float (&y)[OE]) {
#pragma unroll
        for (int r = 0; r < XE; ++r)
            if (r * 64 + ln < F) lds[r * 64 + ln] = xr[r];
        wave_sync();
#pragma unroll 1
        for (int rd = 0; rd < 3; ++rd) {
            const uint32_t* s = sg + rd * ((F + 31) / 32);
#pragma unroll
            for (int r = 0; r < XE; ++r) {
                const int i = r * 64 + ln;
                if (i < F) lds[i] = flip_if_zero_bit(lds[i], s[i >> 5], i & 31);
            }
            wave_sync();
#pragma unroll 1
            for (int h = 1; h < F; h <<= 1) {
                if (ln < F / 2) {
                    const int i = ((ln / h) * 2 * h) + (ln % h);
                    const float u = lds[i], v = lds[i + h];
                    lds[i] = u + v;
                    lds[i + h] = u - v;
                }
                wave_sync();
            }
        }
#pragma unroll
        for (int r = 0; r < OE; ++r) {
            const int i = r * 64 + ln;
            y[r] = (i < F) ? lds[i] : 0.0f;
        }
        wave_sync();
    }
};

// ----------------------------------------------------------------------------------
// Block-wide helpers (blockDim.x == NT, NT a multiple of 64).
// ----------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int ln) {
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if (ln >= o) incl += t;
    }
    return incl - v;
}

// Digit width of radix_select: 8 bits (256 bins); 4 bits can be built for comparison
// with -DCEOS_RADIX_BITS=4.
#ifndef CEOS_RADIX_BITS
#define CEOS_RADIX_BITS 8
#endif
constexpr int kRadixBits = CEOS_RADIX_BITS;
static_assert(kRadixBits == 4 || kRadixBits == 8, "radix digits of 4 or 8 bits");

struct SelectShared {
    uint32_t hist[256];
    uint32_t bin, need, cnt, pad_;
    uint32_t counter;
};

// Exact top-K (K >= 1) of n keys given by get(i), or by get(i, shift) where only the
// key bits at and above `shift` need to be valid. On return, key k is selected iff
// (k >> shift) >= (prefix >> shift). With unique keys exactly min(K, n) keys are
// selected; with repeated keys and min_shift > 0 the prefix holds the K-th largest
// key's bits above min_shift. With kSkipZero, zero keys are absent (not counted, never
// selected; K must not exceed the number of nonzero keys). Must be called by all threads of the block.
// Keys are read 8 per thread at a time (independent loads in flight) and equal
// consecutive digits of one thread are added to the histogram as one atomic.
template <int NT, typename GetKey, bool kSkipZero = false>
__device__ __forceinline__ void radix_select(int n, int K, GetKey get, SelectShared& sh, uint64_t& prefix_out, int& shift_out,
                             int min_shift = 0) {
    constexpr int UNR = 8;
    const int tid = threadIdx.x;
    if (K >= n) {  // everything selected
        prefix_out = 0;
        shift_out = 63;
        return;
    }
    constexpr int NB = 1 << kRadixBits;        // bins
    constexpr int PB = (NB + kWave - 1) / kWave;  // bins per lane in the boundary scan
    uint64_t prefix = 0, mask = 0;
    uint32_t need = (uint32_t)K;
    int shift = 64 - kRadixBits;
    for (;;) {
        for (int b = tid; b < NB; b += NT) sh.hist[b] = 0;
        __syncthreads();
        uint32_t run_bin = 0xffffffffu, run_cnt = 0;
        for (int i0 = tid; i0 < n; i0 += UNR * NT) {
            uint64_t k[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int i = i0 + u * NT;
                if constexpr (std::is_invocable_v<GetKey, int, int>)
                    k[u] = i < n ? get(i, shift) : 0ull;  // getter told which bits are needed
                else
                    k[u] = i < n ? get(i) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int i = i0 + u * NT;
                if (i < n && (k[u] & mask) == prefix && (!kSkipZero || k[u] != 0ull)) {
                    const uint32_t bn = (uint32_t)(k[u] >> shift) & (uint32_t)(NB - 1);
                    if (bn != run_bin) {
                        if (run_cnt) atomicAdd(&sh.hist[run_bin], run_cnt);
                        run_bin = bn;
                        run_cnt = 0;
                    }
                    ++run_cnt;
                }
            }
        }
        if (run_cnt) atomicAdd(&sh.hist[run_bin], run_cnt);
        __syncthreads();
        if (tid < 64) {
            const int ln = tid;
            uint32_t c[PB], tot = 0;
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                const int b = NB - 1 - PB * ln - j;
                c[j] = b >= 0 ? sh.hist[b] : 0u;
                tot += c[j];
            }
            uint32_t ex = wave_excl_scan(tot, ln);
            if (ex < need && need <= ex + tot) {
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    if (need <= ex + c[j]) {
                        sh.bin = NB - 1 - PB * ln - j;
                        sh.need = need - ex;
                        sh.cnt = c[j];
                        break;
                    }
                    ex += c[j];
                }
            }
        }
        __syncthreads();
        const uint32_t bin = sh.bin, nneed = sh.need, cnt = sh.cnt;
        __syncthreads();
        prefix |= (uint64_t)bin << shift;
        mask |= (uint64_t)(NB - 1) << shift;
        need = nneed;
        if (cnt == need || shift <= min_shift) break;
        shift -= kRadixBits;
    }
    prefix_out = prefix;
    shift_out = shift;
}

__device__ __forceinline__ bool selected(uint64_t k, uint64_t prefix, int shift) {
    return shift >= 63 ? true : ((k >> shift) >= (prefix >> shift));
}

// Number of set bits of a wave mask in the lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated append: lanes with `take` get consecutive positions from *counter
// (one atomic per wave). Must be called by every lane of the wave.
__device__ __forceinline__ uint32_t wave_append(bool take, uint32_t* counter) {
    const int ln = threadIdx.x % kWave;
    const uint64_t m = __ballot(take);
    uint32_t base = 0;
    if (ln == 0 && m) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __builtin_amdgcn_readfirstlane(base);
    return base + lanes_below(m);
}

// Reserves cnt (< 2^NB) consecutive positions per lane at *counter (one atomic per
// wave); returns the lane's first position. Lanes take their ranges in lane order.
// The lane prefix sum is formed bit plane by bit plane from ballots. Must be called
// by every lane of the wave.
template <int NB>
__device__ __forceinline__ uint32_t wave_append_n(uint32_t cnt, uint32_t* counter) {
    const int ln = threadIdx.x % kWave;
    uint32_t excl = 0, total = 0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const uint64_t m = __ballot((cnt >> b) & 1u);
        excl += lanes_below(m) << b;
        total += (uint32_t)__popcll(m) << b;
    }
    uint32_t base = 0;
    if (ln == 0 && total) base = atomicAdd(counter, total);
    base = __builtin_amdgcn_readfirstlane(base);
    return base + excl;
}

// Compacts the keys selected by (prefix, shift) into out[] (order unspecified).
template <int NT, typename GetKey, typename Emit>
__device__ __forceinline__ void compact_selected(int n, GetKey get, uint64_t prefix, int shift, uint32_t* counter, Emit emit) {
    for (int i0 = 0; i0 < n; i0 += NT) {
        const int i = i0 + (int)threadIdx.x;
        uint64_t k = 0;
        bool take = false;
        if (i < n) {
            k = get(i);
            take = selected(k, prefix, shift);
        }
        const uint32_t pos = wave_append(take, counter);
        if (take) emit(pos, k);
    }
}

// Descending bitonic sort of N (power of two) 64-bit keys in LDS.
template <int NT>
__device__ __forceinline__ void bitonic_sort_network(uint64_t* s, int N) {
    const int tid = threadIdx.x;
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < (N >> 1); t += NT) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool desc = (i & size) == 0;
                const uint64_t a = s[i], b = s[j];
                if ((a < b) == desc) { s[i] = b; s[j] = a; }
            }
            __syncthreads();
        }
    }
}

// Descending sort of N 64-bit keys in LDS. Up to NT keys, by rank: the key of thread i
// moves to the count of keys greater than it plus equal keys before it (a permutation
// for any keys, ties keep their order), two barriers instead of the network's
// log2(N)(log2(N)+1)/2. Larger N (a power of two) takes the bitonic network.
template <int NT>
__device__ __forceinline__ void bitonic_sort_desc(uint64_t* s, int N) {
    if (N > NT) {
        bitonic_sort_network<NT>(s, N);
        return;
    }
    const int tid = threadIdx.x;
    uint64_t k = 0;
    int r = 0;
    if (tid < N) {
        k = s[tid];
#pragma unroll 4
        for (int j = 0; j < N; ++j) {
            const uint64_t x = s[j];
            r += (x > k) | ((x == k) & (j < tid));
        }
    }
    __syncthreads();
    if (tid < N) s[r] = k;
    __syncthreads();
}

__host__ __device__ __forceinline__ int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace ceos
