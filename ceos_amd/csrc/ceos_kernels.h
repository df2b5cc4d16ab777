// Host-side launch interface of the coCEOs HIP kernels (internal to libceos_amd).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ceos {

// Per-index constants shared by every kernel.
struct IndexDims {
    int d;        // features
    int L;        // log2 of the FWHT size F
    int D;        // projections kept per repeat (n_proj)
    int R;        // repeats
    int m;        // bucket depth (top_m): stride of a bucket list
};

// Bucket entry: x = point id, y = float bits of the stored value
// (close lists: projection; far lists: -projection, src/CEOs.cpp:425-426).
using Entry = uint2;

// Row addressing for the stored points: logical id -> row slot. A static index uses
// head = 0, cap = n. A stream index keeps a ring of `cap` rows whose logical id 0 is
// at slot `head`; bucket ids are absolute and logical = abs - base.
struct RowRing {
    const float* X;
    int64_t cap;
    int64_t head;
    uint32_t base;   // absolute id of logical id 0
    uint32_t live;   // number of live rows
    // int8 copies of the rows (same slots, row stride quant_stride(d) bytes): the d codes,
    // then at byte quant_param_offset(d) the row's (s, r1, r2, 0) as float4; see
    // k_quantize_rows. Null: every candidate gets the exact rerank.
    const int8_t* Xq = nullptr;
};

// Layout of one int8 row copy: codes at [0, d), the float4 (s, r1, r2, 0) at the next
// 16-byte boundary, the row padded to a multiple of 128 bytes (one L2 line). A row then
// starts on a line and its bound terms share a line with its codes whenever the codes
// leave 16 bytes free in their last line (d = 960: 1024 bytes, 8 lines, where codes
// at stride d plus a separate 16-byte record touch 9).
__host__ __device__ inline int quant_param_offset(int d) { return (d + 15) & ~15; }
__host__ __device__ inline int quant_stride(int d) { return (quant_param_offset(d) + 16 + 127) & ~127; }

// ---------------- build ----------------
// Projects rows row(k) = first + k*stride (k < nrows) for all R repeats.
// col_major: out[col * ld_out + k]; else out[k * ld_out + col], col = r*D + j.
// waves_per_wg: rows projected together per workgroup (2, 4 or 8; the query projection's
// tuned launch shape). nrows_dev (optional): a row count read on the device, at most
// nrows (captured plans launch for their largest batch).
hipError_t launch_project_rows(const IndexDims& dm, const float* X, int64_t first, int64_t stride,
                               int64_t nrows, const uint32_t* signs, float* out, int64_t ld_out,
                               bool col_major, hipStream_t st, int waves_per_wg = 4,
                               const int32_t* nrows_dev = nullptr);

// out[c * ld_out + r] = in[r * cols + c] (rows <= 65535 * 64; ld_out >= rows, 0 = rows).
hipError_t launch_transpose(const float* in, int64_t rows, int64_t cols, float* out, hipStream_t st,
                            int64_t ld_out = 0);

// Per list (2*DR lists: list 2c = close, 2c+1 = far) threshold = rank-th largest of the
// column's sample values (far: of -value). sample is column-major [DR][S].
hipError_t launch_sample_thresholds(const float* sample, int64_t S, int DR, int rank, float* thr,
                                    hipStream_t st);

// Streams all n points, projects them and appends (id, value) to list L when
// value >= thr[L] (close) / -value >= thr[L] (far). Lists with active[L] == 0 are
// skipped (active may be null). Writes only while pos < cap[L]; count[L] counts all.
// copy_to (optional): rows 0..n-1 of X are also copied there (the workgroups of repeat r
// store the rows p with p mod R = r).
hipError_t launch_scan_append(const IndexDims& dm, const float* X, int64_t n, const uint32_t* signs,
                              const float* thr, const uint8_t* active, const int64_t* base,
                              const int32_t* cap, int32_t* count, Entry* ent, hipStream_t st,
                              float* copy_to = nullptr);

// Per list: top-m of ent[base[L] .. base[L]+min(count[L],cap[L])) by (value desc, id asc),
// written sorted into bkt[L*m ...]; size[L] = min(m, available).
hipError_t launch_list_select(int nlists, int m, const int64_t* base, const int32_t* cap,
                              const int32_t* count, const Entry* ent, Entry* bkt, int32_t* size,
                              hipStream_t st);

// Stream update: per list, top-m of {old entries whose logical id (abs - new_base) is
// below live} U {new points abs ids first_abs + j with values from newproj (col-major
// [DR][n_new])}; ties prefer the smaller logical id.
hipError_t launch_stream_merge(int DR, int m, Entry* bkt, int32_t* size, const float* newproj,
                               int64_t n_new, uint32_t first_abs, uint32_t new_base, uint32_t live,
                               hipStream_t st);

// Removes entries whose row is flagged in `dead` (bitmap over ring slots).
hipError_t launch_mark_dead(const int32_t* ids, int64_t count, int64_t head, int64_t cap, uint32_t* dead,
                            hipStream_t st);
hipError_t launch_bucket_purge(int nlists, int m, Entry* bkt, int32_t* size, const uint32_t* dead,
                               RowRing ring, hipStream_t st);

// ---------------- query ----------------
struct QueryBatch {
    const float* Q;          // nq x d
    int nq;
    const int32_t* k;        // per-query arrays, or null to use k0/s0/c0
    const int32_t* s;
    const int32_t* c;
    int k0, s0, c0;
    int mp;                  // n_probed_points
    const int64_t* out_off;  // per-query output offset (null: q * k0)
    int32_t* out_id;
    float* out_ip;
    // optional: the batch size read on the device (captured plans launch for a maximum
    // batch nq and take the actual count from here, nq_dev[0] <= nq)
    const int32_t* nq_dev;
    // CEOs-Hash search: rerank every distinct id of the probed lists, no score (c unused)
    int hash;
};

struct QueryConfig {
    int grid;          // persistent workgroups
    int table_log2;    // log2 of the hashed slots, rounded down (reporting; explicit sizes)
    int tile_buckets;  // hashed slots = 4 * tile_buckets (two-choice buckets of four keys)
    int lds_bytes;     // dynamic LDS per workgroup
    int s_max, c_max, k_max;
    int64_t cand_stride;  // scratch entries per workgroup
    int proj_waves = 4;   // query projection: waves per workgroup (2, 4 or 8)
    int prune = 1;        // rerank: bound every candidate from the int8 rows first (when present)
    // optional kernel timing: recorded before the query projection and before and
    // after the fused search kernel (all null: no events)
    hipEvent_t ev[3];
};

struct QueryScratch {
    float* qproj;      // nq * D*R query projections
    int32_t* probes;   // nq * s_max probed list ids, ascending |p|
    uint64_t* cand;    // grid * cand_stride candidate keys
    uint2* pairs;      // grid * cand_stride (slot, value) per probed bucket position
    uint4* stash;      // grid * kSearchStashList entries that missed both tile buckets
    uint32_t* keys;    // grid * search_tile_slots_nbk(tile_buckets) tile ids during the ordered pass
    int32_t* work;     // work counter (1 int)
    int32_t* status;   // error flags (1 int): bit0 = hash table overflow
    float* est;        // running distinct-ids-per-entry estimate carried across launches (1 float)
};

// Fused coCEOs search (projection, probe selection, score accumulation, candidate
// selection, rerank, top-k) over a persistent grid.
hipError_t launch_coceos_search(const IndexDims& dm, const uint32_t* signs, const Entry* bkt,
                                const int32_t* bsize, RowRing rows, const QueryBatch& qb,
                                const QueryConfig& cfg, const QueryScratch& scr, hipStream_t st);

int search_lds_bytes(int d, int tile_buckets, int s_max, int c_max, int k_max);
// Score tile: 4 * nbk hashed slots (nbk two-choice buckets of four keys) plus a
// linear-probing overflow region, a power of two near 1/16 of them, at most 2048.
constexpr int kSearchStashSlots = 2048;
__host__ __device__ inline int search_stash_slots(int nbk) {
    int s = 1;
    while (2 * s <= nbk / 4 && 2 * s <= kSearchStashSlots) s *= 2;
    return s;
}
__host__ __device__ inline int search_tile_slots_nbk(int nbk) { return 4 * nbk + search_stash_slots(nbk); }
__host__ __device__ inline int search_tile_slots(int table_log2) { return search_tile_slots_nbk(1 << (table_log2 - 2)); }
int search_ept(int mp);
constexpr int kSearchThreads = 512;
constexpr int kMaxProbedPoints = 4096;
constexpr int kSearchStashList = 4096;  // stash-list entries per workgroup and pass

// CEOs-Est (dense) search over column-major projections P [DR][n].
hipError_t launch_est_search(const IndexDims& dm, const uint32_t* signs, const float* P, int64_t n,
                             const float* X, const QueryBatch& qb, int grid, uint64_t* scratch,
                             int64_t scratch_stride, int32_t* work, hipStream_t st);

int est_search_lds_bytes(const IndexDims& dm, int s_max, int c_max, int k_max);

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);

// Up to four independent copies of 4-byte words in one launch (for example from mapped
// pinned host memory to device buffers, or back): a kernel on the stream instead of
// copy-engine transfers. No destination range may overlap a source range.
struct CopySegs {
    const uint32_t* src[4];
    uint32_t* dst[4];
    int64_t words[4];
    int n;
};
hipError_t launch_copy_segments(const CopySegs& c, hipStream_t st);

// Centering: center[j] = FP32 of the FP64 mean of column j over rows [0, n) of X, summed
// in a fixed order (partial: column_mean_blocks(n) x d doubles of workspace).
int column_mean_blocks(int64_t n);
hipError_t launch_column_mean(const float* X, int64_t n, int d, double* partial, float* center, hipStream_t st);
// X[(slot0 + i) mod cap][j] -= center[j] for i < count (count <= cap).
hipError_t launch_subtract_center(float* X, int64_t slot0, int64_t count, int64_t cap, int d, const float* center,
                                  hipStream_t st);

// Clears bit (slot0 + j) mod cap of a slot bitmap for j < count (count <= cap).
hipError_t launch_clear_slot_bits(uint32_t* bits, int64_t slot0, int64_t count, int64_t cap, hipStream_t st);

// Rerank bounds for ring slots (slot0 + j) mod cap, j < count: int8 row copies
// Xq [cap][quant_stride(d)]: codes and (s, r1, r2, 0) per slot; see k_quantize_rows.
hipError_t launch_quantize_rows(const float* X, int64_t slot0, int64_t count, int64_t cap, int d, int8_t* Xq,
                                hipStream_t st);

// Streaming read of n floats (n % 4 == 0, x 16-byte aligned) by grid x 256 threads, one
// partial sum per thread into out (grid * 256 floats).
hipError_t launch_stream_read(const float* x, int64_t n, float* out, int grid, hipStream_t st);

}  // namespace ceos
