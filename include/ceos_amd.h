/*
 * ceos_amd — C ABI of the MI355X-native coCEOs library (libceos_amd.so).
 *
 * This is the drop-in boundary for the reference's Python surface
 * (python/python_wrapper.cpp:13-75). The pybind11 module `CEOs` shipped beside it
 * binds these functions with the reference's names; any other FFI (ctypes, cgo,
 * JNI) can bind them directly. Conventions:
 *   - every function returns 0 on success or a CEOS_E* code, and then
 *     ceos_last_error() (per calling thread) describes the failure; the reference
 *     calls exit(1) in the same situations (src/CEOs.cpp:452-461,
 *     src/streamCEOs.cpp:179-189,380-390);
 *   - host-pointer functions take C-contiguous row-major FP32 arrays (rows x d) and
 *     fill caller-owned outputs; *_device functions take device pointers and a
 *     hipStream_t passed as void* (NULL = the index's own stream) and do not
 *     synchronise;
 *   - output rows are sorted by inner product descending and zero padded
 *     (id 0, 0.0) when fewer than k candidates exist (src/CEOs.cpp:659-666);
 *   - the index owns all device memory; builds copy X into HBM
 *     (src/CEOs.cpp:326).
 */
#ifndef CEOS_AMD_H
#define CEOS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CEOS_OK 0
#define CEOS_EINVAL 1   /* invalid argument or parameter combination */
#define CEOS_EDEVICE 2  /* HIP runtime / kernel failure */
#define CEOS_ESTATE 3   /* index not built / wrong call order */
#define CEOS_ENOMEM 4   /* device allocation failed or budget exceeded */
#define CEOS_EAGAIN 5   /* not finished yet (ceos_server_result without wait) */

typedef struct ceos_index ceos_index;
typedef struct ceos_stream ceos_stream;

/* Phase times of a build, update or search, in milliseconds of device time (HIP events),
 * and the index size in GB by the reference's formula (4 top_m x n_proj*n_repeats list
 * floats plus the stored rows). Builds and updates: copy = the points into HBM; project =
 * projections and candidate scan; select = exact top-m per list (or the stream merge).
 * Searches: project = query projection + probe selection; select = the fused kernel
 * (score accumulation, candidate selection, rerank, top-k). */
typedef struct {
    double copy_ms, project_ms, select_ms, total_ms;
    double index_gb;
} ceos_phase_times;

/* Description of the last failure on this thread ("" if none). */
const char* ceos_last_error(void);

/* ------------------------------ class CEOs ------------------------------ */

/* CEOs(n_points, n_features) — python_wrapper.cpp:15-21, CEOs.h:58-61 */
int ceos_create(int n_points, int n_features, ceos_index** out);
void ceos_destroy(ceos_index* h);

/* setIndexParam(n_proj, n_repeats, top_m, n_threads, random_seed) — python_wrapper.cpp:23-30,
 * CEOs.h:63-81. Also sets n_probed_points = top_m (CEOs.h:69). seed < 0 draws a
 * time-based seed (Utilities.cpp:23-25). n_threads only drives host-side work. */
int ceos_set_index_param(ceos_index* h, int n_proj, int n_repeats, int top_m, int n_threads, int seed);

/* Read/write fields n_probed_vectors, n_probed_points, n_cand, centering
 * (python_wrapper.cpp:32-35) and set_threads / clear (python_wrapper.cpp:36-37). */
int ceos_set_query_param(ceos_index* h, int n_probed_vectors, int n_probed_points, int n_cand);
int ceos_get_query_param(const ceos_index* h, int* n_probed_vectors, int* n_probed_points, int* n_cand);
int ceos_set_centering(ceos_index* h, int centering);
int ceos_get_centering(const ceos_index* h, int* centering);
int ceos_set_threads(ceos_index* h, int n_threads);
int ceos_get_threads(const ceos_index* h, int* n_threads);
int ceos_clear(ceos_index* h);

/* Index geometry: fhtDim (CEOs.h:76-79), n_proj, n_repeats, top_m and the seed used. */
int ceos_get_dims(const ceos_index* h, int* fht_dim, int* n_proj, int* n_repeats, int* top_m, int* seed);

/* build_coCEOs_Est(dataset) — python_wrapper.cpp:43, CEOs.cpp:311-441 */
int ceos_build_coceos(ceos_index* h, const float* X, int64_t n, int d);
int ceos_build_coceos_device(ceos_index* h, const float* dX, int64_t n, int d, void* stream);

/* search_coCEOs_Est(queries, n_neighbors) — python_wrapper.cpp:44-45, CEOs.cpp:450-694.
 * Uses the index's n_probed_vectors, n_probed_points and n_cand. ids/ips: nq x k. */
int ceos_search_coceos(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips);
int ceos_search_coceos_device(ceos_index* h, const float* dQ, int nq, int d, int k, int32_t* d_ids,
                              float* d_ips, void* stream);

/* Allocation-free form of ceos_search_coceos_device. No reference counterpart: it is for
 * callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results; it calls no
 * hipMalloc, hipFree or allocating library routine. Its search scratch is a static
 * device array of 64 MiB per GPU, held until the process ends and shared by every call
 * on that GPU: calls that overlap on one GPU race (calls ordered on one stream do not).
 * The persistent grid narrows and the queries go in chunks to fit it; a search whose
 * single workgroup share (about 16 B x n_probed_vectors x n_probed_points + 130 KB)
 * does not fit fails with CEOS_ENOMEM before any launch. Unlike the allocating form it
 * records no event: synchronise the stream before updating, rebuilding or freeing the
 * index. */
int ceos_search_coceos_device_noalloc(ceos_index* h, const float* dQ, int nq, int d, int k, int32_t* d_ids,
                                      float* d_ips, void* stream);

/* build_CEOs_Hash(dataset) — python_wrapper.cpp:46, CEOs.cpp:709-902: the coCEOs lists
 * (ids are what CEOs-Hash uses). With centering set, the stored points are X minus its
 * column mean. search_CEOs_Hash(queries, n_neighbors) — python_wrapper.cpp:47-48,
 * CEOs.cpp:904-1066: every distinct id among the first n_probed_points entries of the
 * n_probed_vectors probed lists is reranked (n_cand is not used). */
int ceos_build_hash(ceos_index* h, const float* X, int64_t n, int d);
int ceos_search_hash(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips);

/* Index files (no reference counterpart; the reference rebuilds on every start): the
 * parameters, the stored points, the buckets and the sign seed of a built coCEOs /
 * CEOs-Hash index, so that a restart skips the build. load replaces the handle's index;
 * the file must hold the handle's n_features. Searches after load return what the saved
 * index returned. */
int ceos_save(const ceos_index* h, const char* path);
int ceos_load(ceos_index* h, const char* path);

/* Batched coCEOs search with per-query k[q], n_probed_vectors s[q] and n_cand c[q]
 * (no reference counterpart: the batched form of search_coCEOs_Est). Query q's
 * results are written at off[q] .. off[q]+k[q]-1 of ids/ips, where
 * off[q] = k[0] + ... + k[q-1]. k_max/s_max/c_max bound the arrays' values. */
int ceos_search_batched(ceos_index* h, const float* Q, int nq, int d, const int32_t* k, const int32_t* s,
                        const int32_t* c, int32_t* ids, float* ips);
int ceos_search_batched_device(ceos_index* h, const float* dQ, int nq, int d, const int32_t* d_k,
                               const int32_t* d_s, const int32_t* d_c, const int64_t* d_off, int k_max,
                               int s_max, int c_max, int32_t* d_ids, float* d_ips, void* stream);

/* build(dataset) / search(queries, n_neighbors) = CEOs-Est — python_wrapper.cpp:39-41,
 * CEOs.cpp:18-98 and 107-297 (dense n x n_proj*n_repeats projections). */
int ceos_build_est(ceos_index* h, const float* X, int64_t n, int d);
int ceos_search_est(ceos_index* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips);

/* Introspection used by parity tests. Buckets are [n_proj*n_repeats][top_m] each:
 * close ids/values then far ids/values (far values hold -projection), the layout of
 * matrix_P rows (CEOs.cpp:417-428). Sign bits: fhtDim*3*n_repeats bytes (0/1) in the
 * order of bitHD3Generator (Utilities.cpp:21-41). Projections: n x (n_proj*n_repeats). */
int ceos_get_buckets(const ceos_index* h, int32_t* close_ids, float* close_vals, int32_t* far_ids,
                     float* far_vals);
int ceos_get_sign_bits(const ceos_index* h, uint8_t* bits);
int ceos_project(ceos_index* h, const float* rows, int64_t n, int d, float* out);

/* Tuning of the search launch (no reference counterpart): workgroups per CU and log2
 * of the LDS score-tile slots; 0 = automatic (for the tile: the most hashed slots, a
 * multiple of 256, that fit the workgroup's LDS share). get returns the last launch's
 * slots as floor(log2) and its LDS bytes. */
int ceos_set_search_config(ceos_index* h, int wg_per_cu, int table_log2);
int ceos_get_search_config(const ceos_index* h, int* wg_per_cu, int* table_log2, int* lds_bytes);

/* Memory bound of a build's list rescans (no reference counterpart). A build selects
 * each list's top_m from the points passing a threshold taken from a strided sample; a
 * list the sample misjudged (fewer than top_m passing, or more than the space set aside)
 * is scanned again. Rescans run over contiguous ranges of lists holding at most
 * `entries` candidate entries (8 bytes each; at least n_points is used), so a sample that
 * misjudges many lists costs more passes, not more memory. Default 2^27 (1 GiB). */
int ceos_set_build_budget(ceos_index* h, int64_t entries);
int ceos_stream_set_build_budget(ceos_stream* h, int64_t entries);
/* The last build's rescans: lists left under top_m, lists over their space, and the
 * rescan passes run (all 0 in the common case). */
int ceos_get_build_rescans(const ceos_index* h, int* under, int* over, int* passes);

/* Rerank bounds (no reference counterpart; on by default). While on, a coCEOs or
 * streamCEOs index keeps an int8 copy of every row with its scale and error bound, each
 * row padded to a multiple of 128 bytes (about a quarter of the FP32 rows), built with
 * the index and kept in step by
 * stream updates. A search whose k is at most 256 and at most a quarter of its
 * candidates then bounds every candidate's inner product from the int8 row and gives
 * the exact FP32 rerank only to candidates that can reach the top-k: same results,
 * about a quarter of the rerank's row reads. Off releases the copies and every
 * candidate gets the exact rerank. Needs n_features % 4 == 0 (otherwise always off).
 * Turning it on for a built index makes the copies at once. */
int ceos_set_rerank_bounds(ceos_index* h, int on);
int ceos_get_rerank_bounds(const ceos_index* h, int* on);
int ceos_stream_set_rerank_bounds(ceos_stream* h, int on);
int ceos_stream_get_rerank_bounds(const ceos_stream* h, int* on);

/* Kernel timing of coCEOs searches (no reference counterpart; for benchmarks). While
 * on, each search launch records HIP events on its stream (up to 4096 launches);
 * turning it on clears earlier records. get_kernel_timing waits for the recorded
 * launches and returns the summed milliseconds of the front kernels (query
 * projection + probe selection) and of the fused search kernel, and the count. */
int ceos_set_kernel_timing(ceos_index* h, int on);
int ceos_get_kernel_timing(ceos_index* h, double* front_ms, double* fused_ms, int* launches);

/* HBM read calibration (no reference counterpart; for benchmarks): times `iters`
 * launches of a plain streaming-read kernel over the device array x (n floats, n a
 * multiple of 4, 16-byte aligned) on `stream`, with n_out/256 workgroups of 256 threads
 * writing one partial sum each into out (n_out floats, a multiple of 256), and returns
 * the read rate in GB/s. Allocates nothing. */
int ceos_read_bandwidth(const float* x, int64_t n, float* out, int64_t n_out, int iters, void* stream,
                        double* gbs);

/* Phase times of the last build (or stream update) and of the last host-entry search.
 * The reference prints these instead of returning them: "Copying data time", "Size of
 * coCEOs-Est index in GB", "Projecting time", "Extracting top-m points time",
 * "Constructing time" (src/CEOs.cpp:326-328, 434-440; src/streamCEOs.cpp:36-37, 151-159,
 * 210-221, 355-368) and, verbose, the search's phase lines (src/CEOs.cpp:674-681). */
int ceos_get_build_times(const ceos_index* h, ceos_phase_times* out);
int ceos_get_search_times(const ceos_index* h, ceos_phase_times* out);
int ceos_stream_get_build_times(const ceos_stream* h, ceos_phase_times* out);
int ceos_stream_get_search_times(const ceos_stream* h, ceos_phase_times* out);

/* The centre subtracted from the stored points (n_features floats) when the last build
 * had centering on: CEOs-Hash (src/CEOs.cpp:721-728) and streamCEOs
 * (src/streamCEOs.cpp:22-25, vec_center). Computed on the device as the FP64 column mean
 * of the build rows, rounded once to FP32. CEOS_ESTATE when there is none. */
int ceos_get_center(const ceos_index* h, float* out);
int ceos_stream_get_center(const ceos_stream* h, float* out);

/* hipStream_t the index launches on (as void*). */
void* ceos_get_stream(const ceos_index* h);

/* ------------------------- captured search plans -------------------------
 * No reference counterpart: the batched form of search_coCEOs_Est / estimate_search
 * for callers that search repeatedly with batches of up to nq_max queries. A plan owns
 * device buffers for the queries (nq_max x d), the per-query k / n_probed_vectors /
 * n_cand (nq_max each, bounded by k_max / s_max / c_max) and the results (query q's
 * row at q*k_max, k[q] entries then zero padded to k_max), plus its own scratch. Its
 * launches (query projection, probe selection, fused search) are captured once in a
 * hipGraph and replayed by ceos_plan_run with the batch size read on the device.
 * With CEOS_PLAN_AUTOTUNE the first run times the candidate launch shapes
 * (workgroups per CU, score-tile size) on that batch and keeps the fastest; results
 * do not depend on the shape. A plan re-captures itself when the index changes (build,
 * update, delete, n_probed_points). Runs of one plan are ordered on the given stream;
 * the plan must not be destroyed while a run is in flight. */
typedef struct ceos_plan ceos_plan;
#define CEOS_PLAN_AUTOTUNE 1

int ceos_plan_create(ceos_index* h, int nq_max, int k_max, int s_max, int c_max, int flags, ceos_plan** out);
int ceos_stream_plan_create(ceos_stream* h, int nq_max, int k_max, int s_max, int c_max, int flags,
                            ceos_plan** out);
void ceos_plan_destroy(ceos_plan* p);
/* Device buffers of the plan: Q (nq_max x d FP32), k, s, c (nq_max int32 each), ids and
 * ips (nq_max x k_max). Any output pointer may be NULL. */
int ceos_plan_buffers(ceos_plan* p, float** dQ, int32_t** dk, int32_t** ds, int32_t** dc, int32_t** d_ids,
                      float** d_ips);
/* Replays the plan for the first nq queries of its buffers on `stream` (NULL: the
 * index's stream); does not synchronise. */
int ceos_plan_run(ceos_plan* p, int nq, void* stream);
/* Host convenience: copies Q, k, s, c in, runs, copies the results out (ids/ips as
 * nq x k_max) and synchronises. */
int ceos_plan_search(ceos_plan* p, const float* Q, int nq, int d, const int32_t* k, const int32_t* s,
                     const int32_t* c, int32_t* ids, float* ips);
/* Launch shape in use (after tuning; table_log2 0 = the automatic tile) and the tuned
 * kernel time in ms (0 if untuned). */
int ceos_plan_config(const ceos_plan* p, int* wg_per_cu, int* table_log2, double* tuned_ms);
/* The query projection's waves per workgroup in use (2, 4 or 8; tuned with the rest). */
int ceos_plan_proj_waves(const ceos_plan* p, int* waves_per_wg);
/* The autotuner's record: for each launch shape it timed, in order, (workgroups per CU,
 * table_log2 (0 = automatic tile), projection waves per workgroup) and the median kernel
 * time in ms of 3 timed launches after a warm-up. Writes at most `cap` entries and sets
 * *count to the number recorded (0 if untuned). */
int ceos_plan_tuning(const ceos_plan* p, int cap, int32_t* shapes, double* ms, int* count);

/* -------------------------------- serving --------------------------------
 * No reference counterpart: continuous batching over a resident index. Requests (one
 * query each, with its own k, n_probed_vectors and n_cand) are queued in arrival order;
 * a worker thread takes every request that has arrived, up to max_batch, into the next
 * free batch slot (a captured search plan with its own stream and pinned staging) as
 * soon as one is free, so batches overlap on the device and requests that arrive while
 * others run are admitted as slots free up. The number of slots is `inflight`, capped by
 * hbm_budget (bytes of device memory for the slots' buffers and scratch; 0 = no cap).
 * A request's results equal those of the same query searched alone. On a streamCEOs
 * server, updates and deletes are applied between batches in submission order. */
typedef struct ceos_server ceos_server;
typedef struct {
    int max_batch;       /* queries per batch slot */
    int k_max, s_max, c_max;  /* bounds of the per-request parameters */
    int inflight;        /* batch slots (>= 1), capped by hbm_budget */
    int64_t hbm_budget;  /* device bytes the slots may hold; 0 = no cap */
    int64_t max_pending; /* queued requests before submit fails with CEOS_ENOMEM; 0 = no cap */
} ceos_server_config;
typedef struct {
    int64_t submitted, completed, batches;
    double mean_batch;   /* queries per batch */
    double p50_ms, p99_ms, max_ms;  /* arrival -> results ready */
    double elapsed_s;    /* first arrival -> last completion */
    double qps;          /* completed / elapsed_s */
    int inflight_slots;
    int64_t device_bytes;
    int64_t inserts, deletes;  /* streamCEOs updates applied */
    double insert_mean_ms, insert_max_ms, delete_mean_ms, delete_max_ms;  /* time to apply one */
    /* where latency goes: queue wait (arrival -> its batch launched) per request, and batch
     * service (launch -> results in host memory) per batch */
    double wait_p50_ms, wait_p99_ms, wait_max_ms;
    double service_p50_ms, service_p99_ms, service_max_ms;
} ceos_server_stats;

int ceos_server_create(ceos_index* h, const ceos_server_config* cfg, ceos_server** out);
int ceos_stream_server_create(ceos_stream* h, const ceos_server_config* cfg, ceos_server** out);
/* Waits for every queued request, then stops the worker. */
void ceos_server_destroy(ceos_server* s);
/* Queues n requests (rows of Q, copied). arrival_s (optional) gives each request's
 * arrival time in seconds after the submission is queued, i.e. after its rows have been
 * copied (open-loop load); NULL = at once. Tickets are first_ticket .. first_ticket + n - 1. */
int ceos_server_submit(ceos_server* s, const float* Q, int n, int d, const int32_t* k, const int32_t* n_probed_vectors,
                       const int32_t* n_cand, const double* arrival_s, int64_t* first_ticket);
/* Results of one request: k entries into ids/ips (descending, zero padded), its k and
 * its latency (arrival to results ready). wait = 0 returns CEOS_EAGAIN while it runs.
 * A ticket can be collected once. */
int ceos_server_result(ceos_server* s, int64_t ticket, int wait, int32_t* ids, float* ips, int* k_out,
                       double* latency_ms);
/* streamCEOs servers: queued like requests, applied between batches. */
int ceos_server_update(ceos_server* s, const float* Xnew, int64_t n_new, int d, int n_del);
int ceos_server_delete(ceos_server* s, const int32_t* ids, int64_t count);
/* Waits until nothing is queued or running; reports a failed update. */
int ceos_server_drain(ceos_server* s);
int ceos_server_get_stats(ceos_server* s, ceos_server_stats* out);
/* Per-batch trace since the last reset (up to 2^20 batches, in the order they completed):
 * when each batch was launched,
 * how long the worker took to enqueue it, when its results were seen in host memory, the
 * wait of its oldest request, its size and its slot. count gets the number recorded;
 * at most max are copied. */
typedef struct {
    double launch_s;        /* seconds after the first arrival */
    double enqueue_ms;      /* launch -> rows copied and device work enqueued */
    double done_ms;         /* launch -> results in host memory, when the worker reaped them */
    double oldest_wait_ms;  /* arrival of the batch's oldest request -> launch */
    int32_t queries, slot;
    /* the enqueue's steps, ms: rows to the pinned staging, inputs to the device, the plan's
     * launch, the results' copy and the completion event, the server lock taken again */
    double steps_ms[5];
} ceos_server_batch;
int ceos_server_get_batches(ceos_server* s, ceos_server_batch* out, int64_t max, int64_t* count);
int ceos_server_reset_stats(ceos_server* s);

/* --------------------------- class streamCEOs --------------------------- */

/* streamCEOs(n_features) — python_wrapper.cpp:53, streamCEOs.h:47-51 */
int ceos_stream_create(int n_features, ceos_stream** out);
void ceos_stream_destroy(ceos_stream* h);

/* setIndexParam(n_proj, n_repeats, top_points, n_threads, random_seed) —
 * python_wrapper.cpp:55-57, streamCEOs.h:53-70 (sets n_probed_vectors = top_points). */
int ceos_stream_set_index_param(ceos_stream* h, int n_proj, int n_repeats, int top_points, int n_threads,
                                int seed);
int ceos_stream_set_query_param(ceos_stream* h, int n_probed_vectors, int n_probed_points, int n_cand);
int ceos_stream_get_query_param(const ceos_stream* h, int* n_probed_vectors, int* n_probed_points,
                                int* n_cand);
int ceos_stream_set_threads(ceos_stream* h, int n_threads);
int ceos_stream_get_threads(const ceos_stream* h, int* n_threads);
/* fhtDim, n_proj, n_repeats, top_m and the seed used (as ceos_get_dims). */
int ceos_stream_get_dims(const ceos_stream* h, int* fht_dim, int* n_proj, int* n_repeats, int* top_m, int* seed);
/* centering (streamCEOs.h:44, not bound in python_wrapper.cpp:52-71): stored rows are
 * minus the build rows' column mean, fixed for later updates (streamCEOs.cpp:22-34,
 * 198-203); queries are not centered. Takes effect at the next build. */
int ceos_stream_set_centering(ceos_stream* h, int centering);
int ceos_stream_get_centering(const ceos_stream* h, int* centering);
int ceos_stream_clear(ceos_stream* h);

/* build(dataset) — python_wrapper.cpp:65, streamCEOs.cpp:9-161 */
int ceos_stream_build(ceos_stream* h, const float* X, int64_t n, int d);
/* update(new_dataset, n_delPoints) — python_wrapper.cpp:66, streamCEOs.cpp:174-369:
 * appends the rows, drops the n_del oldest (ids shift down by n_del). */
int ceos_stream_update(ceos_stream* h, const float* Xnew, int64_t n_new, int d, int n_del);
/* estimate_search(queries, n_neighbors) — python_wrapper.cpp:67-68, streamCEOs.cpp:378-612 */
int ceos_stream_estimate_search(ceos_stream* h, const float* Q, int nq, int d, int k, int32_t* ids,
                                float* ips);
/* hash_search(queries, n_neighbors) — python_wrapper.cpp:69-70, streamCEOs.cpp:621-877 */
int ceos_stream_hash_search(ceos_stream* h, const float* Q, int nq, int d, int k, int32_t* ids, float* ips);
/* Random delete (no reference counterpart; the reference deletes FIFO only): the
 * points with these current ids leave every bucket at once and are never returned
 * again; ids of other points do not change. The rows keep their positions until the
 * FIFO delete of update() drops them. */
int ceos_stream_delete_ids(ceos_stream* h, const int32_t* ids, int64_t count);
int ceos_stream_size(const ceos_stream* h, int64_t* n_points);
/* streamCEOs index files: live points in id order, buckets, random-delete marks. */
int ceos_stream_save(const ceos_stream* h, const char* path);
int ceos_stream_load(ceos_stream* h, const char* path);
/* Buckets with per-list valid sizes ([n_proj*n_repeats] each). */
int ceos_stream_get_buckets(const ceos_stream* h, int32_t* close_ids, float* close_vals, int32_t* far_ids,
                            float* far_vals, int32_t* close_size, int32_t* far_size);

#ifdef __cplusplus
}
#endif

#endif /* CEOS_AMD_H */
