/* warpstore_amd: C ABI of the MI355X (gfx950) analytic operators.
 *
 * This is the drop-in boundary for the engine's GPU operator surface (reference src/gpu.h and
 * src/jit.h). Each entry point names the declaration it replaces. Conventions:
 *   - plain pointers and sizes only; plan structs are the byte-compatible ws_* types of ws_abi.h;
 *   - every fallible call returns a status: WS_OK, WS_ERR_OOM (1041: GPU out of memory),
 *     WS_ERR_INTERNAL (1105: HIP or internal error), WS_ERR_NEEDS_PLAN (1235: plan beyond the
 *     interpreter's limits without a compiled plan); ws_last_error() gives the message of the
 *     calling thread's last failure. The C++ shim (warpstore_amd/gpu.h) rethrows them as SqlError
 *     with the same codes, as the reference does (gpu.cu:23-30);
 *   - device buffers returned by the library ("tmp" buffers) come from its caching allocator and
 *     are released with ws_tmp_free (DevVec<T> semantics, gpu.h:59-79): a freed block is reused,
 *     in stream order, by later work of the thread that freed it; a block freed by another thread
 *     than the one that took it goes to a shared pool, and its next taker waits for the device
 *     first; ws_trim_pool returns the calling thread's cached blocks and the shared pool to the
 *     device;
 *   - all work of a host thread goes to that thread's default stream (hipStreamPerThread), the HIP
 *     form of the reference's per-thread default stream (CMakeLists.txt:50-52).
 */
#ifndef WARPSTORE_AMD_H
#define WARPSTORE_AMD_H

#include "warpstore_amd/ws_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define WS_OK 0
#define WS_ERR_OOM 1041
#define WS_ERR_INTERNAL 1105
#define WS_ERR_NEEDS_PLAN 1235

typedef struct ws_agg ws_agg;                 /* GpuAggregator (gpu.h:87-130) */
typedef struct ws_jit_program ws_jit_program; /* JitProgram (jit.h:14-28) */

/* ------------------------------------------------------------------ errors, device */
const char* ws_last_error(void);                    /* message of this thread's last failed call */
int ws_init(void);                                   /* gpuInit        gpu.h:12 */
int ws_sm_count(void);                               /* gpuSmCount     gpu.h:13 (compute units) */
int ws_arch(void);                                   /* gpuArch        gpu.h:14 (950 for gfx950) */
int ws_describe(char* buf, size_t cap);              /* gpuDescribe    gpu.h:15 */

/* ------------------------------------------------------------------ memory */
int ws_dev_alloc(size_t bytes, void** out);          /* devAlloc       gpu.h:16 */
void ws_dev_free(void* p);                           /* devFree        gpu.h:17 */
int ws_tmp_alloc(size_t bytes, void** out);          /* tmpAlloc       gpu.h:50 */
/* Page-locked host memory (the reference keeps a thread-local pinned stage, gpu.cu:226-240):
   result buffers the caller reads many groups into copy at link speed. No reference entry point. */
int ws_pinned_alloc(size_t bytes, void** out);
void ws_pinned_free(void* p);
void ws_trim_pool(void);                             /* gpuTrimPool    gpu.h:51 */
void ws_tmp_free(void* p);                           /* tmpFree        gpu.h:52 */
int ws_h2d(void* dst, const void* src, size_t bytes);/* h2d            gpu.h:53 */
int ws_d2h(void* dst, const void* src, size_t bytes);/* d2h            gpu.h:54 */
/* No reference entry point: a device-to-host copy queued on the library stream and not waited for
   (dst: page-locked memory from ws_pinned_alloc); the next synchronizing call completes it. */
int ws_d2h_async(void* dst, const void* src, size_t bytes);
int ws_d2d(void* dst, const void* src, size_t bytes);/* d2d            gpu.h:55 */
int ws_sync(void);                                   /* gpuSync        gpu.h:56 */
/* VRAM tiering (out of scope: every column is GPU resident here; these keep the link surface). */
void ws_set_vram_limit(int64_t bytes);               /* gpuSetVramLimit gpu.h:20 */
int64_t ws_vram_used(void);                          /* gpuVramUsed    gpu.h:21 */
double ws_vram_fill(void);                           /* gpuVramFill    gpu.h:22 */
int ws_place_column(void* host, size_t bytes, int host_only, int* mapped, void** out); /* placeColumn gpu.h:23 */
void ws_release_column(void* dev, void* host, size_t bytes, int mapped);              /* releaseColumn gpu.h:24 */
void* ws_vram_acquire(const void* host, size_t bytes);  /* vramAcquire gpu.h:44 (NULL: no room) */
void ws_vram_release(void* dev, size_t bytes);          /* vramRelease gpu.h:45 */
void* ws_host_map(void* host, size_t bytes);            /* hostMap     gpu.h:46 (NULL: unsupported) */
void ws_host_unmap(void* host);                         /* hostUnmap   gpu.h:47 */
int64_t ws_vram_limit(void);                            /* gpuVramLimit gpu.h:48 */
int64_t ws_vram_free(void);                             /* gpuVramFree  gpu.h:49 */
/* Column DML on the GPU (out of scope: NULL tells the caller to take its host path, which the
 * reference contract allows for both). */
void* ws_patch_column(const void* old_dev, size_t bytes, int width, const int32_t* rows, const void* vals,
                      int64_t m); /* gpuPatchColumn gpu.h:27 */
typedef struct ws_gpu_update {
  int kind;
  int64_t i0, add, nul, lo, hi;
  double d0, add_d;
  uint8_t not_null, bad, saw_null;
  double mn, mx;
} ws_gpu_update;                  /* GpuUpdate gpu.h:33-40 */
void* ws_update_column(const void* old_col, const void* src, int64_t n, int w, int f64, const int32_t* rows, int64_t m,
                       ws_gpu_update* u); /* gpuUpdateColumn gpu.h:41 */

/* ------------------------------------------------------------------ aggregation (GpuAggregator) */
int ws_agg_create(const ws_agg_plan* plan, int64_t capacity, const ws_dense_spec* dense, ws_agg** out); /* ctor gpu.h:89 */
void ws_agg_destroy(ws_agg* a);                                                              /* dtor gpu.h:90 */
int ws_agg_add_batch(ws_agg* a, const ws_expr_env* env, int ncols, int64_t n, const uint32_t* del); /* addBatch gpu.h:91 */
int ws_agg_run(ws_agg* a, const ws_expr_env* env, const ws_jit_program* jit);                /* run       gpu.h:92 */
int ws_agg_run_window(ws_agg* a, const ws_expr_env* env, const ws_jit_program* jit);         /* runWindow gpu.h:95 */
int ws_agg_batches(const ws_agg* a);                                                         /* batches   gpu.h:96 */
int ws_agg_windows(const ws_agg* a);                                                         /* windows   gpu.h:97 */
int64_t ws_agg_rows_in(const ws_agg* a);                                                     /* rowsIn    gpu.h:98 */
int ws_agg_dense(const ws_agg* a);                                                           /* dense     gpu.h:99 */
int ws_agg_overflowed(ws_agg* a, int* out);                                                  /* overflowed gpu.h:100 */
/* fetch (gpu.h:102) in two steps: compact the groups on the device and report their number, then
 * copy them row-major (keys groups x nkeys, accs groups x naggs) into caller buffers. Without GROUP BY
 * there is one "group"; accs holds the low words / values and ws_agg_sum_high the high words. */
int ws_agg_fetch_begin(ws_agg* a, int64_t* groups);
int ws_agg_fetch_copy(ws_agg* a, double* keys, double* accs);
int ws_agg_sum_bits_slots(const ws_agg* a, uint8_t* out); /* sumBitsSlots gpu.h:104 (naggs bytes) */
int ws_agg_sum_high(const ws_agg* a, int64_t* out);       /* sumHigh gpu.h:105 (naggs values, after fetch) */
/* compact (gpu.h:107): groups left in device memory, column-major; *keys / *accs are tmp buffers. */
int ws_agg_compact(ws_agg* a, double** keys, double** accs, int64_t* groups);
int64_t ws_agg_capacity(const ws_agg* a);                  /* capacity   gpu.h:108 */
int ws_agg_group_count(ws_agg* a, int64_t* out);           /* groupCount gpu.h:109 */
/* Batched form (no reference counterpart; the reference runs one aggregator at a time, with at
 * most 4 concurrent queries, engine.cpp:1262-1337): many dense aggregations of one plan shape,
 * each with its own batches, environment and accumulators, in one launch; and their groups
 * fetched in one pass (counts[q] groups of aggs[q], keys / accs row-major query after query,
 * ascending group index; fails with *total set when *total > cap). jit: a compiled dense plan of
 * the shape, or NULL for the interpreter. Results equal separate runs (double sums up to order). */
int ws_agg_run_batch(int n, ws_agg* const* aggs, const ws_expr_env* const* envs, const ws_jit_program* jit);
/* A batch prepared once and replayed (no reference counterpart: repeated dashboards, a serving
   loop's standing query sets). The aggregators (dense, one plan shape, batches added) must outlive
   the plan; every replay starts their tables from the initial values, runs the batched kernel and
   fetches like ws_agg_fetch_many. The launches are recorded in a HIP graph at prepare time. */
typedef struct ws_agg_batch_plan ws_agg_batch_plan;
int ws_agg_batch_prepare(int n, ws_agg* const* aggs, const ws_expr_env* const* envs, const ws_jit_program* jit,
                         ws_agg_batch_plan** out);
int ws_agg_batch_replay(ws_agg_batch_plan* p, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs);
/* A replay in two halves (no reference counterpart: a pipelined batch loop). ws_agg_batch_launch
   enqueues the plan's graph and the read-back of its group counts on the plan's own stream (after
   the calling thread's earlier work) and returns; ws_agg_batch_finish waits for the counts, copies
   the groups out as ws_agg_batch_replay does, and returns when they are in keys / accs. Between the
   two the host can launch another plan, whose kernels then run while this one's groups are copied
   (pinned keys / accs, e.g. from ws_pinned_alloc, copy without a staging pass). A plan holds one
   launch at a time: launching it again before its finish, or finishing it unlaunched, fails. */
int ws_agg_batch_launch(ws_agg_batch_plan* p);
int ws_agg_batch_finish(ws_agg_batch_plan* p, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs);
void ws_agg_batch_release(ws_agg_batch_plan* p);
int ws_agg_fetch_many(int n, ws_agg* const* aggs, int64_t* counts, int64_t* total, int64_t cap, double* keys, double* accs);
/* Aggregation strategy of a hash GROUP BY (no reference counterpart: the hook through which a
 * sizing model, or a test, chooses the physical plan; results do not depend on it).
 *   WS_STRAT_AUTO: partitioned when the first run covers >= 2^20 rows and >= 2^16 groups are
 *     expected; WS_STRAT_TABLE: the shared hash table (reference kernels.cuh:306-415);
 *   WS_STRAT_PARTITIONED: count / scatter / merge through hash partitions, no shared table unless a
 *     partition outgrows its LDS table.
 * expected_groups > 0 replaces the capacity-derived estimate (capacity / 2) for sizing the
 * partitions and the spill table. Call before the first run. */
enum { WS_STRAT_AUTO = 0, WS_STRAT_TABLE = 1, WS_STRAT_PARTITIONED = 2 };
int ws_agg_set_strategy(ws_agg* a, int strategy, int64_t expected_groups);
int ws_agg_partitioned(const ws_agg* a); /* 1 when the last run produced a partitioned result */
/* Process-wide tuning knobs of the physical plans (no reference counterpart: the autotuner's
 * interface; results do not depend on them). Keys: "part_merge_lds_kib" (8..152, default 60),
 * "part_scatter_lds_kib" (16..152, default 120), "part_chunks_per_cu" (1..64, default 1),
 * "part_auto_rows" (default 2^20), "part_auto_groups" (default 2^16), "batch_tile_rows" (256..2^20,
 * default 4096), "batch_lds_slots" (64..8192, default 1024), "batch_blocks_per_cu" (1..32,
 * default 4), "part_pack" (0 or 1, default 1: packed partial-group records for integer keys),
 * "select_chunks_per_cu" (1..64, default 8: chunked selection of long inputs),
 * "select_mask_min_rows" (default 2^16: compiled selections of at least this many rows take the
 * ballot pass with set bitmaps staged in LDS, then a count scan and an expansion),
 * "dense_staged_min_set_words" (default 1024: compiled dense GROUP BYs whose WHERE reads set
 * bitmaps of at least this many words stage them in LDS), "batch_key_ranges" (0..64, default 0:
 * batched queries with more groups than the LDS table, in at most this many table-sized key
 * ranges, read their rows once per range), "topk_steps" (1..64, default 8) and
 * "topk_blocks_per_cu" (1..16, default 1): the first pass of top-k selection.
 * Unknown keys and values out of range return WS_ERR_INTERNAL. */
int ws_tune_set(const char* key, int64_t value);
int ws_tune_get(const char* key, int64_t* value);

/* ------------------------------------------------------------------ column statistics (sizing)
 * No reference counterpart: the segment statistics the sizing model reads (the reference engine
 * keeps min / max per segment, engine_core.h:3869-3882; the HyperLogLog registers are new).
 * Distinct identity: integer types by exact value, doubles by the bits of (v + 0.0); NULLs are
 * counted in `nulls` and nowhere else. HyperLogLog with p = 11: register x >> 53 of
 * x = mix64(value bits) (the splitmix64 finalizer), rank = leading zeros of (x << 11 | 1 << 10) + 1. */
typedef struct ws_col_stats {
  int64_t rows;       /* live (not deleted) rows */
  int64_t nulls;      /* live NULL rows */
  int64_t imin, imax; /* integer types: exact min / max (INT64_MAX / INT64_MIN when none) */
  double dmin, dmax;  /* min / max as doubles (NaN when none) */
  uint8_t hll[2048];  /* HyperLogLog registers */
} ws_col_stats;
int ws_column_stats(const ws_colview* col, int64_t n, const uint32_t* del, ws_col_stats* out); /* one batch */
void ws_col_stats_merge(ws_col_stats* dst, const ws_col_stats* src);
double ws_hll_estimate(const uint8_t* regs); /* distinct-count estimate of 2048 registers */

/* ------------------------------------------------------------------ learned sizing
 * No reference counterpart (the reference sizes from key ranges, engine_select.cpp:1154-1235): an
 * MLP 16 -> 32 -> 32 -> 1 (ReLU, fp32) predicting log2(groups + 1) / 32 from ws_sizing_features,
 * trained on the GPU (forward, backward and Adam kernels) on exact counts the aggregations log.
 * Parameter layout: W1[32][16], b1[32], W2[32][32], b2[32], w3[32], b3. */
#define WS_SIZER_INPUTS 16
#define WS_SIZER_PARAMS 1633
typedef struct ws_sizer ws_sizer;
int ws_sizer_create(uint64_t seed, ws_sizer** out); /* He-uniform weights from the seed */
void ws_sizer_destroy(ws_sizer* s);
int ws_sizer_get_params(const ws_sizer* s, float* out);
int ws_sizer_set_params(ws_sizer* s, const float* in);
int ws_sizer_predict(ws_sizer* s, const float* x, int64_t n, float* y);     /* x: n x 16, host */
int ws_sizer_loss_grad(ws_sizer* s, const float* x, const float* t, int64_t n, float* loss, float* grad); /* MSE, dL/dparams */
int ws_sizer_train(ws_sizer* s, const float* x, const float* t, int64_t n, int epochs, int batch, float lr, uint64_t seed,
                   float* final_loss);
int ws_sizer_groups(ws_sizer* s, const float* features, int64_t* groups);  /* predicted group count */
int ws_sizing_features(const ws_col_stats* keys, int nkeys, const int64_t* dict_sizes, int64_t rows, double selectivity,
                       float* out);                                           /* 16 features of a GROUP BY */
int ws_agg_set_features(ws_agg* a, const float* features); /* log (features, exact groups) after the run */
int64_t ws_sizing_log_size(void);
int64_t ws_sizing_log_copy(float* x, float* t, int64_t cap); /* x: cap x 16; t: log2(groups + 1) / 32 */
void ws_sizing_log_clear(void);
/* Join output size model (no reference counterpart; the reference counts, scans and reads the size
   back, gpu.cu:1129-1193): features of an equi-join from its key columns' statistics, and a log of
   exact output sizes for training a sizer model, whose ws_sizer_groups then predicts pairs. */
int ws_join_features(const ws_col_stats* build_key, int64_t build_rows, int64_t build_table_rows, const ws_col_stats* probe_key,
                     int64_t probe_rows, float* out);
void ws_join_log_add(const float* features, int64_t pairs);
int64_t ws_join_log_size(void);
int64_t ws_join_log_copy(float* x, float* t, int64_t cap); /* x: cap x 16; t: log2(pairs + 1) / 32 */
void ws_join_log_clear(void);

/* ------------------------------------------------------------------ selection / gathers */
int ws_select(const ws_expr_env* env, int has_where, ws_prog where, int64_t n, const uint32_t* del, int32_t* out_idx,
              const ws_jit_program* jit, int64_t* count); /* gpuSelect gpu.h:133-134 */
/* gpuSelectMulti gpu.h:138-140: *out_idx is a tmp buffer of the selected global row indices. */
int ws_select_multi(const ws_expr_env* env, int has_where, ws_prog where, int nb, const int64_t* bn,
                    const uint32_t* const* bdel, const ws_colview* bcols, int ncols, int32_t** out_idx,
                    int64_t* per_batch, const ws_jit_program* jit, int64_t* count);
/* No counterpart in the reference: ws_select for callers that cannot allocate device memory
   (code under graph capture, serving loops that forbid allocation). Interpreted plans only, one
   batch of at most 2^28 rows (268,435,456; larger inputs fail with 1105). Its descriptor and
   look-back state live in a 2 MiB device array per GPU held until the process ends; no device
   allocation or free happens in the call. Calls that overlap on one GPU share that array and race:
   serialize them. Same results as ws_select (ascending row indices of rows passing the WHERE and
   not deleted). */
int ws_select_noalloc(const ws_expr_env* env, int has_where, ws_prog where, int64_t n, const uint32_t* del,
                      int32_t* out_idx, int64_t* count);
int ws_gather_multi(const ws_colview* src, const int64_t* boff, int nb, int ptype, int esize, const int32_t* idx,
                    int64_t m, void* dst);                                                   /* gpuGatherMulti gpu.h:142-143 */
int ws_remap_codes(const int32_t* in, int64_t n, const int32_t* map, int64_t map_n, int32_t* out); /* gpuRemapCodes gpu.h:145 */
int ws_bits_to_double(const double* src, int64_t n, double* dst);                          /* gpuBitsToDouble gpu.h:147 */
int ws_eval_at_int(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t* out); /* gpuEvalAtInt gpu.h:149 */
/* gpuDistinctInts gpu.h:152: sorted distinct values into out (capacity out_cap), count in *n_out. */
int ws_distinct_ints(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, int64_t base, int64_t span,
                     int64_t* out, int64_t out_cap, int64_t* n_out);
typedef struct ws_temp_col_spec {
  int8_t src;
  int32_t idx, hi, cnt;
  uint8_t kind, key_int, sum_bits, out;
  double div, mul;
} ws_temp_col_spec; /* TempColSpec gpu.h:155-163 */
int ws_bits_from_column(const void* col, int out, int64_t n, int64_t base, int64_t nbits, int* saw_null,
                        uint32_t** bits); /* gpuBitsFromColumn gpu.h:165 (free with ws_dev_free) */
int ws_temp_column(const double* keys, int nk, const double* accs, int na, int64_t g, const ws_temp_col_spec* s,
                   double* mn, double* mx, void** out); /* gpuTempColumn gpu.h:166 (free with ws_dev_free) */
int ws_eval_at(const ws_expr_env* env, ws_prog prog, const int32_t* idx, int64_t m, double* out); /* gpuEvalAt gpu.h:168 */
int ws_gather(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst);             /* gpuGather gpu.h:170 */
int ws_gather_enc(const ws_colview* v, int ptype, const int32_t* idx, int64_t m, void* dst);     /* gpuGatherEnc gpu.h:172 */
int ws_key_bitmap(ws_colview key, int64_t n, int64_t lo, int64_t nbits, uint32_t* bits);         /* gpuKeyBitmap gpu.h:174 */
int ws_set_bits(uint32_t* bits, const int32_t* idx, int64_t m);                                  /* gpuSetBits gpu.h:176 */
/* No reference entry point: the number of set bits among `words` words (synchronous). The star-join
   driver compares it with the filtered build rows: equal means the build keys are distinct. */
int ws_bitmap_count(const uint32_t* bits, int64_t words, int64_t* count);
/* No reference entry point: materialize(build, WHERE, {key}) followed by gpuKeyBitmap (gpu.h:174)
   in one scan, for a join filter whose build rows are not needed otherwise. Zeroes the
   ceil(nbits / 32) words of `bits`, then sets bit key - lo for every live row of the nb batches
   that passes the WHERE and whose integer key (a program over the batch columns) lies in
   [lo, lo + nbits); *count (device memory, may be null) gets the number of passing rows.
   Asynchronous on the library stream. 0 < nbits <= 2^31. jit (may be null): a compiled selection
   plan (WS_JF_FLAG / WS_JF_FLAGM) whose keys[0] is `key`; without one the interpreter runs. */
int ws_key_bitmap_where(const ws_expr_env* env, int has_where, ws_prog where, ws_prog key, int nb, const int64_t* bn,
                        const uint32_t* const* bdel, const ws_colview* bcols, int ncols, int64_t lo, int64_t nbits,
                        uint32_t* bits, uint64_t* count, const ws_jit_program* jit);
int ws_gather_null(const void* src, int ptype, const int32_t* idx, int64_t m, void* dst);        /* gpuGatherNull gpu.h:193 */
int ws_gather_bytes(const void* src, int esize, const int32_t* idx, int64_t m, void* dst, int nulls); /* gpuGatherBytes gpu.h:195 */

/* ------------------------------------------------------------------ segment encodings (host side)
 * Table::encodeColumn storage.cu:1097-1194: a sealed segment column in the compressed format the
 * kernels decode inline (ws_colview.enc): 1 / 2 / 3 integer frame of reference (u8 / u16 / u32
 * codes, top code = NULL), 4 double dictionary (<= 255 distinct values), 5 / 6 decimal doubles
 * (u16 / u32 codes of round(x * 10^s), s <= 4). *enc = 0 when no encoding fits (keep the column
 * raw); otherwise *bytes bytes were written to out (cap >= ws_encoded_bytes_max(n)). The
 * reference encodes only from 1024 rows on and under its compression setting; that policy stays
 * with the caller. */
size_t ws_encoded_bytes_max(int64_t n);
int ws_encode_column(const void* col, int ptype, int64_t n, uint8_t* out, size_t cap, int* enc, size_t* bytes);

/* ------------------------------------------------------------------ joins (outputs are tmp buffers) */
int ws_hash_join(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int32_t** out_build,
                 int32_t** out_probe, int64_t* m); /* gpuHashJoin gpu.h:178-179 */
/* gpuHashJoinParts gpu.h:184-186: b_off has pb+1 entries (pb == 1 or pb == pp), p_off pp+1; out_off
 * (optional, pp+1 entries) receives the output offset of each probe partition. */
int ws_hash_join_parts(ws_colview build_key, const int64_t* b_off, int pb, ws_colview probe_key, const int64_t* p_off,
                       int pp, int32_t** out_build, int32_t** out_probe, int64_t* out_off, int64_t* m);
int ws_hash_join_outer(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int32_t** out_build,
                       int32_t** out_probe, int full, int64_t* m); /* gpuHashJoinOuter gpu.h:189-190 */
/* gpuHashJoin with a predicted output size (expected_pairs > 0): the single-probe form writes into
   buffers of that size without reading the count back first; a smaller real size is exact, a larger
   one costs a second write pass. Same pairs and order as ws_hash_join. */
int ws_hash_join_sized(ws_colview build_key, int64_t nb, ws_colview probe_key, int64_t np, int64_t expected_pairs,
                       int32_t** out_build, int32_t** out_probe, int64_t* m);
int ws_cross_join(int64_t na, int64_t nb, int32_t** out_a, int32_t** out_b, int64_t* m); /* gpuCrossJoin gpu.h:191 */

/* ------------------------------------------------------------------ sorting */
/* gpuSortPerm gpu.h:197: stable multi-key sort (keys[0] most significant, desc[i] != 0 descending),
 * *perm is a tmp buffer of m int32. */
int ws_sort_perm(const double* const* keys, const uint8_t* desc, int nkeys, int64_t m, int32_t** perm);
/* ORDER BY ... LIMIT k (engine_select.cpp:1320-1342 sorts all m rows and keeps perm[0, k)):
 * *perm (a tmp buffer of min(k, m) int32) is exactly the first min(k, m) entries of ws_sort_perm's
 * permutation, found by selection for up to 4 keys and k <= 256 (else by the full sort). */
int ws_topk_perm(const double* const* keys, const uint8_t* desc, int nkeys, int64_t m, int64_t k, int32_t** perm);
int ws_map_rank(double* keys, const int32_t* rank, int64_t m);                      /* gpuMapRank gpu.h:199 */
int ws_sort_perm_host(const void* host_col, int ptype, int64_t n, int32_t* out);    /* gpuSortPermHost gpu.h:201 */

/* ------------------------------------------------------------------ per-query compiled plans (jit.h) */
int ws_jit_enabled(void);                       /* jitEnabled       jit.h:37 */
int ws_jit_mode(void);                          /* jitMode          jit.h:38 */
void ws_jit_set_mode(int mode);                 /* jitSetMode       jit.h:39 (this thread) */
void ws_jit_set_default_mode(int mode);         /* jitSetDefaultMode jit.h:40 */
void ws_jit_set_cache_dir(const char* dir);     /* jitSetCacheDir   jit.h:42 ("" = memory only) */
/* jitGet jit.h:47: *out is NULL when codegen is off, still compiling (interpret first) or failed. */
int ws_jit_get(const ws_expr_env* env, int ncols, const ws_agg_plan* plan, int fn, const char* kind,
               const ws_jit_program** out);
int ws_jit_program_id(const ws_jit_program* p);                 /* JitProgram::id */
int ws_jit_launch(const ws_jit_program* p, int fn, unsigned grid, unsigned block, size_t shmem, void** args); /* jitLaunch jit.h:48 */
int ws_prog_is_int(const ws_expr_env* env, ws_prog p, int* out); /* progIsInt jit.h:51 */
void ws_jit_begin_query(const char* sql);                         /* jitBeginQuery jit.h:60 */
typedef struct ws_jit_query_info {
  int compiled, hits, from_disk, background;
  double compile_ms;
  int nids;
  int ids[16];
} ws_jit_query_info;
void ws_jit_end_query(double query_ms, ws_jit_query_info* info);  /* jitEndQuery jit.h:61 */
typedef struct ws_plan_cache_row {
  int id;
  char kind[64];
  char origin[16];
  double compile_ms;
  int64_t execs;
  double avg_ms, last_ms;
  size_t source_bytes;
  char sql[256];
} ws_plan_cache_row;
int ws_jit_plan_cache(ws_plan_cache_row* rows, int cap);          /* jitPlanCache jit.h:73 (returns count) */
int ws_jit_source_of(int id, char* buf, size_t cap);              /* jitSourceOf jit.h:74 (returns length) */
void ws_jit_clear_cache(void);                                    /* jitClearCache jit.h:75 */
/* Tooling, no counterpart in the reference: compiles a plan's generated kernel for `arch`
 * (e.g. "gfx950") with hipRTC without loading it (no device needed); buf gets the generated
 * evaluator, or the compiler log on failure. Used to check code generation on build hosts. */
int ws_jit_compile_check(const ws_expr_env* env, int ncols, const ws_agg_plan* plan, int fn, const char* arch, char* buf,
                         size_t cap, size_t* code_bytes);

/* ------------------------------------------------------------------ profiling (PROFILE stages) */
/* Per-kernel timing with HIP events recorded on the launching stream, the GPU-side counterpart of
 * the engine's PROFILE stages (engine_core.h:4016-4028). While enabled, every kernel launch of the
 * calling thread is bracketed by events; ws_prof_collect synchronizes and sums them per kernel. */
void ws_prof_enable(int on);
void ws_prof_only(const char* kernel); /* restrict recording to one kernel name (NULL or "" = all) */
void ws_prof_reset(void);
/* Totals per kernel name: returns the number of kernels (at most cap). */
typedef struct ws_prof_row {
  char name[96];
  int64_t launches;
  double total_ms;
} ws_prof_row;
int ws_prof_collect(ws_prof_row* rows, int cap);
/* Tooling, no counterpart in the reference: a plain device-to-device streaming copy kernel
 * (bytes a multiple of 16), the known-size reference for calibrating hardware byte counters. */
int ws_copy_kernel(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* WARPSTORE_AMD_H */
