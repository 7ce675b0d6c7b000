/* ArrowMetal C ABI, MI355X (gfx950) build: the argsort / take / group-by subset.
 *
 * The declarations keep the reference's signatures, argument meanings, return codes and ownership
 * rules, so that its callers (the DuckDB extension, the Spark FFM downcalls, the Polars plugin, the
 * Python/Go/Node/R/Rust bindings) load libArrowMetalC.so in place of libArrowMetalC.dylib unchanged.
 * Each declaration cites the reference declaration it replaces (include/arrowmetal.h:LINE in the
 * reference checkout) and the Swift entry point that implemented it there.
 *
 * Conventions (reference include/arrowmetal.h:1-3, Sources/ArrowMetalC/ArrowMetalC.swift:9-34):
 *   return 0 on success; 1 on an operation error, with the message in am_last_error() (thread-local,
 *   valid until the next call on the thread); 2 for a NULL or bad handle argument (no message).
 *   Every am_array* / am_groupby* written to an out-parameter is a new reference the caller releases.
 *   Calls are synchronous: results are complete on return, except between am_batch_begin and
 *   am_batch_end on the calling thread, where the GPU work is ordered on the thread's stream and the
 *   next sync point (am_batch_end, am_export, a count read back) completes it.
 *
 * Device arrays are HIP device memory on the current ROCm device. am_export_device reports
 * ARROW_DEVICE_ROCM (10) and a NULL sync_event (the data is ready when the call returns).
 */
#ifndef ARROWMETAL_H
#define ARROWMETAL_H
#include <stdint.h>
#include "arrow_abi.h"
#ifdef __cplusplus
extern "C" {
#endif

/* reference include/arrowmetal.h:12 */
typedef struct am_array am_array;
/* reference include/arrowmetal.h:900 */
typedef struct am_groupby am_groupby;
/* Extension of this build: a plan over many record batches (see am_batched_create). */
typedef struct am_batched am_batched;
/* Extension of this build: a request server (see am_server_create). */
typedef struct am_server am_server;

/* reference include/arrowmetal.h:15-17 (ArrowMetalC.swift am_version / am_device_name / am_last_error) */
const char* am_version(void);
const char* am_device_name(void);
const char* am_last_error(void);

/* reference include/arrowmetal.h:20-21 (CInterop.swift:119-139): the producer's array is moved (its
 * release is set to NULL). A CPU array is copied into HBM once; am_import_device with
 * device_type ARROW_DEVICE_ROCM keeps the producer's device buffers (zero copy) and releases the
 * producer when the last handle dies. ARROW_DEVICE_CPU goes through am_import. */
int  am_import(const struct ArrowSchema* schema, struct ArrowArray* array, am_array** out);
int  am_import_device(const struct ArrowSchema* schema, struct ArrowDeviceArray* array, am_array** out);
/* reference include/arrowmetal.h:22-34 (ArrowMetalC.swift:52-67, ChunkedImport.swift): `n` arrays of the
 * one type `schema` describes (`arrays` points at n contiguous structs), imported as one array of their
 * total length; each chunk's offset, length, validity bitmap (or its absence) and null count (-1
 * included) are honoured. On success every chunk is moved (its release is NULL on return); on failure
 * a chunk whose release is still set remains the caller's. n == 1 is am_import. Taken here: the
 * integer, float and temporal types. Returns 3, reading and moving nothing, with "Unsupported Arrow
 * type: am_import_chunks does not take F arrays; import their concatenation" for any other type.
 * am_import_chunks_supported returns 1 when the type is taken, 0 otherwise (and for NULL). */
int  am_import_chunks(const struct ArrowSchema* schema, struct ArrowArray* arrays, int64_t n, am_array** out);
int  am_import_chunks_supported(const struct ArrowSchema* schema);
/* reference include/arrowmetal.h:48-58 (CInterop.swift:338-412): am_export fills host buffers
 * (copied from HBM once per handle); am_export_device hands out the device buffers themselves. */
int  am_export(am_array* a, struct ArrowSchema* schema, struct ArrowArray* array);
int  am_export_device(am_array* a, struct ArrowSchema* schema, struct ArrowDeviceArray* array);
void am_release(am_array* a);
int64_t     am_length(am_array* a);
int64_t     am_null_count(am_array* a);
const char* am_format(am_array* a);
const char* am_compute_format(am_array* a);

/* reference include/arrowmetal.h:93 (ArrowMetalC.swift:372-383, Take.swift:41-80). indices: int32,
 * int64 or uint32; a null index gives a null; an index outside [0, length) is the error
 * "take: index out of range (array length N)". */
int  am_take(am_array* a, am_array* indices, am_array** out);

/* reference include/arrowmetal.h:94 (ArrowMetalC.swift:384-389): a zero-copy window; offset < 0,
 * length < 0 or offset + length > am_length(a) is the error "slice out of range". */
int  am_slice(am_array* a, int64_t offset, int64_t length, am_array** out);

/* reference include/arrowmetal.h:108 (ArrowMetalC.swift:394-399): stable uint32 ("I") indices,
 * ascending or descending, nulls last, float order ieee. */
int  am_argsort(am_array* a, int descending, am_array** out);

/* reference include/arrowmetal.h:109-110 (ArrowMetalC.swift:400-415): am_sort is the values in
 * am_argsort's order (same type, nulls last); am_top_k the first k indices of am_argsort with
 * descending = largest (k < 0 is the error "top_k needs k >= 0"; k above the length gives them all). */
int  am_sort(am_array* a, int descending, am_array** out);
int  am_top_k(am_array* a, int64_t k, int largest, am_array** out);

/* reference include/arrowmetal.h:458 (ArrowMetalC_Window.swift:106-118): stable multi-key argsort,
 * descending per key (NULL: all ascending), nulls last, float order ieee. */
int  am_lexsort(am_array** columns, const int* descending, int64_t count, am_array** out);

/* reference include/arrowmetal.h:113 (ArrowMetalC.swift:416-469): dense keys int32/int64/uint32 in
 * [0, key_count); null and out-of-range keys are skipped. agg: 0 sum, 1 count(rows), 2 min, 3 max,
 * 4 mean, 5 count(values). sum and mean take integer values (int64 / uint64 / float64 out); sum also
 * takes float32 values (the float32 total per key, returned as float64); min/max take 32-bit or
 * narrower values; any other pairing is the error "Unsupported Arrow type: group-by agg N on <fmt>". */
int  am_group_by(am_array* keys, int64_t key_count, int agg, am_array* values, am_array** out);

/* reference include/arrowmetal.h:184-185 (MetalContext.swift:245-312): per-thread batching. */
int  am_batch_begin(void);
int  am_batch_end(void);

/* reference include/arrowmetal.h:911-918 (ArrowMetalC_GroupByExtra.swift:42-89). Group order is
 * ascending by key (nulls last) for numeric columns, lexicographic in column order for several
 * columns; ids are int32 and never null; am_group_by_group_count returns -1 for a NULL handle. */
int  am_group_by_keys(am_array** columns, int64_t count, am_groupby** out);
int64_t am_group_by_group_count(am_groupby* gb);
int  am_group_by_keys_result(am_groupby* gb, int64_t i, am_array** out);
int  am_group_by_ids(am_groupby* gb, am_array** out);
void am_group_by_release(am_groupby* gb);

/* reference include/arrowmetal.h:920-954 (ArrowMetalC_GroupByExtra.swift:92-181). Ops taken here:
 *   0 hash_sum       integers wrap in 64 bits (int64 out; uint64 out for uint64 values);
 *                    float64 and float32 values: the exact sum rounded once (float64 out)
 *   1 hash_count_all rows per group (values may be NULL)
 *   2 hash_count     non-null values per group
 *   3 hash_mean      float64 out; for float values the exact sum over the count, rounded once
 *   4 hash_min       values' type; float64 compares -0.0 < +0.0 and skips NaN
 *   5 hash_max
 *   6 hash_min_max   struct<min, max> of the two above, the struct itself never null
 *                    (Kernels/AggregatesExtra.swift:90-96)
 * A group with no value to answer with is null. hash_count takes a column of any type. Ops 7-25
 * return 1 with "am_group_agg_ex op N is not supported by this build"; any other op with
 * "unknown am_group_agg_ex op N". */
int  am_group_agg_ex(am_groupby* gb, am_array* values, int op, double p1, am_array** out);

/* reference include/arrowmetal.h:1331, 1348 (ArrowMetalC_Options.swift:110-119, 166-175).
 * null_placement 1 at_start, any other value at_end (ArrowMetalC_Options.swift:50); float_order 0
 * ieee, 1 total, 2 nan_largest, any other value the error "unknown float_order N" (:52-59). */
int  am_argsort_ex(am_array* a, int descending, int null_placement, am_array** out);
int  am_argsort_ex2(am_array* a, int descending, int null_placement, int float_order, am_array** out);

/* reference include/arrowmetal.h:1332-1333, 1349-1354 (ArrowMetalC_Options.swift:126-228):
 *   am_sort_ex / am_sort_ex2   the values in am_argsort_ex / am_argsort_ex2's order, the input's type,
 *                              the null block where null_placement puts it; -0.0 and NaN payload bits
 *                              come out unchanged.
 *   am_top_k_ex                the first k indices am_argsort_ex2 gives with the same options
 *                              (largest = descending); k < 0 is the error "top_k needs k >= 0".
 *   am_lexsort_ex              stable multi-key argsort, one null_placement for every key.
 *   am_lexsort_ex2             each of descending / null_placement / float_order holds `count` entries,
 *                              or is NULL for the defaults (ascending, at_end, ieee).
 * Index results are uint32 ("I"); more than 2^32 - 1 rows is the error "<op>: row R does not fit the
 * UInt32 index type (...)" with op "top_k" or "lexsort". */
int  am_sort_ex(am_array* a, int descending, int null_placement, am_array** out);
int  am_sort_ex2(am_array* a, int descending, int null_placement, int float_order, am_array** out);
int  am_top_k_ex(am_array* a, int64_t k, int largest, int null_placement, int float_order, am_array** out);
int  am_lexsort_ex(am_array** columns, const int* descending /* or NULL */, int64_t count, int null_placement,
                   am_array** out);
int  am_lexsort_ex2(am_array** columns, const int* descending /* or NULL */, const int* null_placement /* or NULL */,
                    const int* float_order /* or NULL */, int64_t count, am_array** out);

/* reference include/arrowmetal.h:1359 (ArrowMetalC_Options.swift:284-297,
 * Kernels/UniqueOrder.swift:117-125, Kernels/Unique.swift:46-51): Arrow `unique` of a numeric
 * column. order 0 (and any value other than 1): first appearance, with the null kept once at the
 * position of the first null row; order 1: ascending, nulls dropped. */
int  am_unique_ex(am_array* a, int order, am_array** out);

/* reference include/arrowmetal.h:1515-1600 (Sources/ArrowMetalC/ArrowMetalC_Stream.swift): streaming
 * group-by over a producer of record batches (any Arrow C Stream: pyarrow's RecordBatchReader, Polars,
 * DuckDB). am_stream_from_c_stream moves `in` (NULL plus am_stream_last_error() on failure); a stream is
 * consumed by one terminal. am_stream_group_by: `n_keys` key columns; per aggregate an op (0 sum,
 * 1 count, 2 min, 3 max, 4 mean; 5 variance, 6 stddev and 7 count_distinct_approx return 1 with
 * "... is not supported by this build"), a value column (NULL only for count, which then counts rows)
 * and an output name. dense_key_count > 0: one integer key column inside [0, dense_key_count), returned
 * as int32, keys no row landed on (and keys outside the range) dropped. Result: the key columns, then
 * one column per aggregate, groups in ascending key order with the null key last; sum null for a group
 * without a value (int64 for signed, uint64 for unsigned columns, float64 for float values: the
 * correctly rounded sum of the batches' correctly rounded sums), count int64, min / max the values'
 * type, mean float64. Memory is one batch plus the batches' partial group states, folded back to the
 * running answer as they grow. ddof and prefetch_depth are accepted and unused here. */
typedef struct am_stream am_stream;
typedef struct am_stream_result am_stream_result;
const char* am_stream_last_error(void);
am_stream*  am_stream_from_c_stream(struct ArrowArrayStream* in, int prefetch_depth);
void        am_stream_release(am_stream* s);
int am_stream_group_by(am_stream* s, const char** keys, int64_t n_keys, const int* ops, const char** columns,
                       const char** names, int64_t n_aggs, int64_t dense_key_count, int ddof,
                       am_stream_result** out);
int64_t     am_stream_result_column_count(am_stream_result* r);
const char* am_stream_result_column_name(am_stream_result* r, int64_t i);
int         am_stream_result_column(am_stream_result* r, int64_t i, am_array** out);
int64_t     am_stream_result_scalar_count(am_stream_result* r);
const char* am_stream_result_scalar_name(am_stream_result* r, int64_t i);
int         am_stream_result_scalar(am_stream_result* r, int64_t i, int64_t* out_i64, double* out_f64,
                                    int* out_kind, int* is_null);
int64_t     am_stream_result_rows_out(am_stream_result* r);
int         am_stream_result_stats(am_stream_result* r, int64_t* batches, int64_t* rows, int64_t* bytes_read,
                                   double* wall_s, double* read_s, double* gpu_s, double* merge_s,
                                   double* overlap);
void        am_stream_result_release(am_stream_result* r);

/* reference include/arrowmetal.h:1178-1236 (Sources/ArrowMetalC/ArrowMetalC_Expr.swift:36-123): am_query
 * over equal-length named columns. This build runs its fused dense group-by, the form the DuckDB
 * extension sends for small key spans (duckdb-extension/src/arrowmetal_rewrite.cpp:1319-1350):
 *   (query (group_by K "name" KEY) (aggregate AGG+))
 *   KEY := (col "c") | (cast KEY INT-TYPE) | (sub KEY LIT) | (add KEY LIT)   -- keys outside [0,K) drop
 *   AGG := (count "n") | (count "n" V) | (sum "n" V) | (mean "n" V) | (min "n" V) | (max "n" V)
 *   V   := (col "c") | (shr (col "c") LIT) | (bit_and (col "c") LIT);  LIT := (int N) | (i64 N) | ...
 * Result: column 0 "name" = int32 0..K-1, then one column per aggregate: counts int64 (never null);
 * sums int64 (uint64 for unsigned columns, float64 for float32 values, rounded once from the exact
 * sum); means float64; min / max the values' type; null where no value contributed. As in the
 * reference: a float64 sum or mean, and min / max wider than 32 bits, are refused. Filters, projections,
 * whole-input aggregates and other expressions return 1 with "Unsupported Arrow type: am_query: ... is
 * not supported by this build". No scalars are produced (am_query_scalar_count 0). */
typedef struct am_query_result_t am_query_result;
int  am_query(am_array** columns, const char** names, int64_t n_columns, const char* expr_text,
              am_query_result** out);
int64_t     am_query_column_count(am_query_result* r);
const char* am_query_column_name(am_query_result* r, int64_t i);
int         am_query_column(am_query_result* r, int64_t i, am_array** out);
int64_t     am_query_scalar_count(am_query_result* r);
const char* am_query_scalar_name(am_query_result* r, int64_t i);
int         am_query_scalar(am_query_result* r, int64_t i, int64_t* out_i64, double* out_f64,
                            int* out_kind, int* is_null);
void        am_query_result_release(am_query_result* r);
const char* am_query_canonical(const char* expr_text);

/* reference include/arrowmetal.h:1362-1433 (Sources/ArrowMetalC/ArrowMetalC_Plan.swift,
 * Sources/ArrowMetal/Engine/PlanJSON.swift): the plan engine, one JSON plan per call over tables
 * registered with am_plan_source_create (which retains the column handles it is given; columns of
 * unequal length are the error "Array length mismatch: A vs B").
 * This build runs the subset on the ORDER BY / group-by path:
 *   plan  := {"op": "scan",     "source": NAME, "columns": [NAME, ...]?}
 *          | {"op": "sort",     "input": plan, "by": [KEY, ...], "nulls": NULLS?, "float_order": FORDER?}
 *          | {"op": "limit",    "input": plan, "count": INT, "offset": INT?}
 *          | {"op": "unique",   "input": plan, "subset": [NAME, ...]?}
 *          | {"op": "group_by", "input": plan, "keys": [[NAME, "(col \"c\")"], ...],
 *                               "aggs": [[AGGOP, NAME, "(col \"c\")"?], ...]}
 *   KEY, NULLS, FORDER and AGGOP as in the reference grammar; a sort under a limit of 1..1024 rows with
 *   no offset and one numeric key is a top-k selection. Expressions are column references only. The
 *   reference's other operators (filter, select, with_columns, aggregate, join, join_asof, concat,
 *   window, explode) and expressions return 1 with "Unsupported Arrow type: plan: ... is not
 *   supported by this build"; an unknown operator with "Invalid ArrowArray: plan: unknown operator".
 * Results: group_by keys in ascending key order with the null group last, then one column per
 * aggregate (sum: int64 for signed integers, uint64 for unsigned, float64 for float values rounded
 * once from the exact sum; mean float64; count int64; min/max the values' type). `optimize` is
 * accepted; it changes no result on this subset. am_plan_explain's text is valid until the next call
 * on the thread (NULL plus am_last_error() when the plan does not parse). am_plan_column hands out a
 * new reference the caller releases with am_release. */
typedef struct am_plan_source_t am_plan_source;
typedef struct am_plan_result_t am_plan_result;
int  am_plan_source_create(const char* name, am_array** columns, const char** names,
                           int64_t n_columns, am_plan_source** out);
void am_plan_source_release(am_plan_source* s);
int  am_plan_run(const char* plan_json, am_plan_source** sources, int64_t n_sources, int optimize,
                 am_plan_result** out);
const char* am_plan_explain(const char* plan_json, am_plan_source** sources, int64_t n_sources,
                            int optimize);
int64_t     am_plan_column_count(am_plan_result* r);
int64_t     am_plan_row_count(am_plan_result* r);
const char* am_plan_column_name(am_plan_result* r, int64_t i);
int         am_plan_column(am_plan_result* r, int64_t i, am_array** out);
void        am_plan_result_release(am_plan_result* r);

/* ---------------------------------------------------------------------------------------------
 * Extensions of this build (no counterpart in the reference header). They serve the batched and
 * served forms of the same kernels, the keep-first unique of the plan engine's "unique" operator
 * (reference Sources/ArrowMetal/Engine/Executor.swift:134-148), and measurement.
 * ------------------------------------------------------------------------------------------- */

/* No counterpart in the reference: am_group_by for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The integer
 * aggregates of am_group_by (agg 0 sum, 1 count rows, 2 min, 3 max of 32-bit or narrower values,
 * 4 mean, 5 count values; not the float32 sum) over key_count <= 65536 groups, written into caller
 * device buffers: out_values holds key_count elements (int64 for sums and counts, uint64 for sums of
 * uint64, float64 for means, the values' type for min / max), out_validity ceil(key_count / 64)
 * little-endian bitmap words (bit k: group k has a value), out_null_count one int64. The only other
 * memory it touches is a fixed workspace in each GPU's global memory (about 9 MiB, held until the
 * process ends); a larger key_count returns 1 with "Invalid ArrowArray: am_group_by_noalloc: key_count
 * N exceeds the 65536 groups of its fixed workspace". Calls that overlap on one GPU share that
 * workspace and race. Results are ready when the call returns. */
int  am_group_by_noalloc(am_array* keys, int64_t key_count, int agg, am_array* values, void* out_values,
                         uint64_t* out_validity, int64_t* out_null_count);

/* Keep-first unique rows over `count` key columns: the lowest row of every distinct key combination,
 * ascending, as uint32 row numbers (take them with am_take). */
int  am_unique_rows(am_array** columns, int64_t count, am_array** out);

/* Batched form of am_argsort_ex2 and of am_group_by_keys + am_group_agg_ex sum / count_all, over
 * `nbatches` record batches stored back to back: batch b is rows [offsets[b], offsets[b+1]) of each
 * column (offsets has nbatches + 1 entries, offsets[0] == 0, at most 2^32 rows in total). The
 * reference runs such work one call per batch (its per-thread command-buffer batching,
 * Sources/ArrowMetal/MetalContext.swift:245-312); here all batches go through one fixed kernel
 * sequence, captured in a hipGraph on the first run (use_graph != 0) and replayed.
 *
 *   am_batched_create   plan for the batch layout; hash_capacity (rounded up to a power of two, at
 *                       most what one workgroup's LDS holds) bounds the distinct group keys per batch
 *                       handled by the batched tables; tile_items 8, 16 or 20 keys per thread.
 *   am_batched_tune     times the candidate tile sizes, hash capacities and group-by strategies
 *                       (LDS table per chunk or one global table per batch) on these columns and keeps
 *                       the fastest in the plan.
 *   am_batched_run      sort_keys: any numeric column, or NULL; group_keys + values: integer columns,
 *                       or both NULL. No nulls. Each batch's sort is stable, nulls-last semantics of
 *                       am_argsort_ex2 with `descending` and `float_order` (NaN after every value under
 *                       ieee). Each batch's groups are ascending by key; sums wrap in 64 bits. A batch
 *                       with more distinct keys than the capacity is grouped again by the general
 *                       hash group-by, so every result is complete.
 *   am_batched_result   of the last run: 0 sort indices (uint32, row numbers local to their batch,
 *                       batches back to back); 1 group keys (the key column's type); 2 sums (int64);
 *                       3 counts (int64); 4 group offsets (int64, nbatches + 1: batch b's groups are
 *                       rows [o[b], o[b+1]) of results 1-3).
 *   am_batched_config   up to `count` of: tile items, hash capacity, group-by strategy, digit bits,
 *                       batches regrouped by the general path in the last run, sort tiles, and 1 when
 *                       the last run sorted by the MSD radix sort (64-bit keys), 0 by the LSD passes. */
int  am_batched_create(const int64_t* offsets, int64_t nbatches, int hash_capacity, int tile_items,
                       am_batched** out);
int  am_batched_tune(am_batched* b, am_array* sort_keys, am_array* group_keys, am_array* values,
                     int descending, int float_order);
int  am_batched_run(am_batched* b, am_array* sort_keys, am_array* group_keys, am_array* values,
                    int descending, int float_order, int use_graph);
int  am_batched_result(am_batched* b, int which, am_array** out);
int  am_batched_config(am_batched* b, int64_t* out, int count);
void am_batched_release(am_batched* b);

/* Request server with continuous batching over HBM-resident columns. The reference serves one
 * synchronous call at a time on the caller's thread; here `slots` worker threads (1..16), each with its
 * own HIP stream, take requests from one FIFO queue. A request is admitted when a worker is free and
 * its estimated device footprint fits in what `budget_bytes` has left (one larger than the whole budget
 * runs alone); requests that arrive while others run join them and their kernels overlap. Admission
 * keeps arrival order. Each request runs the code of the plain calls, so its results are those of the
 * plain calls, alone or in the mix.
 *
 *   am_server_submit   op 0: group-by of `columns` with `values`: results keys (one array per key
 *                      column, ascending groups), int64 sum (hash_sum), int64 count (hash_count_all);
 *                      op 1: argsort of columns[0] with descending / null_placement / float_order as
 *                      am_argsort_ex2: result uint32 indices; op 2: keep-first unique rows over
 *                      `columns`: results the uint32 rows (am_unique_rows) and each column gathered at
 *                      them. The server holds a reference to every array until the request is done.
 *   am_server_wait     blocks until the request is done, hands over up to `cap` results (new
 *                      references; *count gets how many there are) and writes times_ms[3]: queued,
 *                      running, submit to done. A failed request returns 1 with its error.
 *   am_server_stats    up to `count` of: submitted, completed, failed, peak admitted footprint (bytes),
 *                      peak requests running together, budget, queued now, slots.
 *   am_server_release  finishes the queued requests, then stops the workers. */
int  am_server_create(int slots, int64_t budget_bytes, am_server** out);
int  am_server_submit(am_server* s, int op, am_array** columns, int64_t count, am_array* values,
                      int descending, int null_placement, int float_order, int64_t* ticket);
int  am_server_wait(am_server* s, int64_t ticket, am_array** out, int64_t cap, double* times_ms,
                    int64_t* count);
int  am_server_stats(am_server* s, int64_t* out, int count);
void am_server_release(am_server* s);

/* Test hook, the counterpart of the reference's Dispatch.maxRowIndex (Sources/ArrowMetal/Kernels/
 * Dispatch.swift:190-196, lowered in Tests/ArrowMetalTests/IndexTypeTests.swift:142-144): the largest
 * row number an index-returning call (argsort, unique rows, unique) may produce, 2^32 - 1 by default.
 * Past it such calls fail with "Invalid ArrowArray: <op>: row R does not fit the UInt32 index type (row
 * numbers go up to M)". Returns the previous limit; a value outside [0, 2^32 - 1] restores the default.
 * Process-wide. */
int64_t am_set_max_row_index(int64_t max_row);

/* Kernel timing of the library's own launches, measured with HIP events on the stream each kernel
 * is launched on. am_prof_report writes "name count total_ms\n" lines into buf (NUL-terminated) and
 * returns the number of bytes needed. */
int  am_prof_enable(int on);
int  am_prof_reset(void);
int64_t am_prof_report(char* buf, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* ARROWMETAL_H */
