/*
 * gpjson.h -- C ABI of the MI355X (gfx950) GPJSON index + query library
 * (libgpjson_hip.so).
 *
 * The reference reaches the GPU through Truffle NFI bindings of libcudart /
 * libcuda / libnvrtc made by CUDARuntime
 * (gpjson/src/main/java/com/koenv/gpjson/gpu/CUDARuntime.java:76-94, symbol
 * table CUDARuntimeFunction.java:39-194, CUDADriverFunction.java:34-46,
 * NVRuntimeCompiler.java:297-307) and drives seven kernels from Java
 * (functions/QueryFunction.java:48-265).  This library replaces that whole
 * layer: a HIPRuntime on the Java side binds the entry points below instead
 * of cudaMalloc / cuLaunchKernel / nvrtc*.  Paths cited below are relative to
 * gpjson/src/main/java/com/koenv/gpjson/ in the reference.
 *
 * Conventions
 *  - Every entry point returns an int status: GPJSON_OK (0) or a GPJSON_ERR_*
 *    code, mirroring CUDARuntime.checkCUDAReturnCode (gpu/CUDARuntime.java:506-516);
 *    gpjson_last_error() gives the message of the calling thread's last failure.
 *  - The library owns all device memory behind opaque handles; the caller owns
 *    host buffers and passes their capacity.
 *  - A context owns one HIP stream; calls on one context are serialized by the
 *    caller (the reference is single-threaded per context).
 *  - Offsets are 64-bit (the reference is limited to files < 2 GiB).
 */
#ifndef GPJSON_H
#define GPJSON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GPJSON_OK = 0,
  GPJSON_ERR_HIP = 1,          /* HIP runtime error (CUDA error in the reference) */
  GPJSON_ERR_INVALID_ARG = 2,  /* bad argument (interop UnsupportedTypeException / ArityException) */
  GPJSON_ERR_PARSE = 3,        /* JSONPathException: "Failed to parse JSON path" */
  GPJSON_ERR_UNSUPPORTED = 4,  /* UnsupportedJSONPathException (reference falls back to jayway) */
  GPJSON_ERR_IO = 5,           /* file missing / changed while reading */
  GPJSON_ERR_CAPACITY = 6,     /* caller buffer too small or a stated limit exceeded */
  GPJSON_ERR_NO_MEMORY = 7,    /* device or pinned allocation failed / HBM budget exceeded */
  /* gpjson_query_file: a path is outside the supported subset and the options
     allow the reference's fallback (functions/QueryFunction.java:122-128,
     267-314): nothing ran on the GPU, and the caller evaluates the query with
     its own JSONPath library (the reference uses jayway json-path) */
  GPJSON_FALLBACK_REQUIRED = 8
};

/* Index export selectors for gpjson_index_export. */
enum {
  GPJSON_INDEX_NEWLINE = 0, /* int64[lines] */
  GPJSON_INDEX_STRING = 1,  /* uint64[ceil(n/64)] */
  GPJSON_INDEX_LEVELED = 2, /* uint64[levels][ceil(n/64)], level-major */
  GPJSON_INDEX_ESCAPE = 3,  /* uint64[ceil(n/64)], only with GPJSON_BUILD_VERIFY */
  GPJSON_INDEX_QUOTE = 4,   /* uint64[ceil(n/64)], only with GPJSON_BUILD_VERIFY */
  GPJSON_INDEX_COLON = 5,   /* uint64[ceil(n/64)]: every ':' byte (no reference counterpart) */
  GPJSON_INDEX_RAW_QUOTE = 6, /* uint64[ceil(n/64)]: every '"' byte, escaped or not (no reference counterpart) */
  /* uint64[L][ceil(n/4096)]: bit k of word s of level l is set iff word 64*s+k
     of leveled bitmap l is not zero (no reference counterpart; lets the
     interpreter skip 4 KB without structure at a level with one load) */
  GPJSON_INDEX_LEVEL_SUMMARY = 7
};

/* Flags of gpjson_index_build. */
enum { GPJSON_BUILD_VERIFY = 1 /* also materialize escape and quote bitmaps */ };

/* Flags of gpjson_query_file (QueryOptions, functions/QueryOptions.java:8-14). */
enum {
  GPJSON_QUERY_DISABLE_FALLBACK = 1,
  /* (no reference counterpart) gather the result bytes on the device and copy
     them back once, instead of decoding each element from a second map of
     the file (QueryFunction.java:247-259): see gpjson_query_gather */
  GPJSON_QUERY_GATHER = 2
};

#define GPJSON_MAX_LEVELS 64   /* reference: 16, stages/LeveledBitmapsIndex.java:16 */
#define GPJSON_MAX_QUERIES 255 /* functions/QueryFunction.java:74-77 */

typedef struct gpjson_ctx gpjson_ctx;
typedef struct gpjson_buf gpjson_buf;
typedef struct gpjson_index gpjson_index;
typedef struct gpjson_result gpjson_result;
typedef struct gpjson_batch gpjson_batch;
typedef struct gpjson_server gpjson_server;

/* ---- errors -------------------------------------------------------------- */

/* Message of the calling thread's last failed call (replaces GPJSONException
 * text, GPJSONException.java:33-39).  Never NULL. */
const char *gpjson_last_error(void);

/* ---- context ------------------------------------------------------------- */

/* Opens device `device` with one stream (replaces `new CUDARuntime`,
 * gpu/CUDARuntime.java:76-94, and the lazy context of :480-488). */
int gpjson_ctx_create(int device, gpjson_ctx **out);
/* Frees every device allocation of the context (CUDARuntime has no close). */
void gpjson_ctx_destroy(gpjson_ctx *ctx);
/* Waits for the context's stream (cudaDeviceSynchronize, gpu/CUDARuntime.java:388). */
int gpjson_ctx_synchronize(gpjson_ctx *ctx);
/* Compute units of the device (cudaDeviceGetAttribute, CUDARuntimeFunction.java:39-194). */
int gpjson_ctx_num_cus(gpjson_ctx *ctx, int *out);

/* ---- JSONPath compiler ---------------------------------------------------- */

/* Compiles a JSONPath of the supported subset to IR bytes
 * (jsonpath/JSONPathParser.java:20-261 + IRBuilder.java:19-81).  Writes at
 * most cap bytes; *ir_len gets the full size.  GPJSON_ERR_PARSE or
 * GPJSON_ERR_UNSUPPORTED on failure. */
int gpjson_compile(const char *path, uint8_t *ir, size_t cap, size_t *ir_len, int *max_depth,
                   int *num_results);

/* ---- staged input -------------------------------------------------------- */

/* Copies n host bytes to device memory through pinned staging buffers with
 * hipMemcpyAsync (replaces cudaMalloc + synchronous pageable cudaMemcpy of
 * QueryFunction.readFile, functions/QueryFunction.java:149,316-335). */
int gpjson_buf_from_host(gpjson_ctx *ctx, const void *data, int64_t n, gpjson_buf **out);
/* Reads a file straight into pinned staging and on to the device; fails with
 * GPJSON_ERR_IO if its size changes while reading (QueryFunction.java:320-322). */
int gpjson_buf_from_file(gpjson_ctx *ctx, const char *path, gpjson_buf **out);
/* Wraps n bytes already resident on the context's device (no copy; the caller
 * keeps ownership and must keep them alive while the handle is used). */
int gpjson_buf_wrap_device(gpjson_ctx *ctx, const void *device_ptr, int64_t n, gpjson_buf **out);
int64_t gpjson_buf_size(const gpjson_buf *buf);
void gpjson_buf_free(gpjson_buf *buf);

/* ---- structural index ---------------------------------------------------- */

/* Builds newline, string and leveled indexes for `levels` levels in one fused
 * launch (replaces CombinedIndex.create + LeveledBitmapsIndex.create,
 * stages/CombinedIndex.java:28-56, stages/LeveledBitmapsIndex.java:40-54, and
 * their kernels).  The index keeps a reference to buf. */
int gpjson_index_build(gpjson_ctx *ctx, gpjson_buf *buf, int levels, int flags, gpjson_index **out);
/* Reads a file and indexes it (replaces QueryFunction.readFile followed by the
 * index build, functions/QueryFunction.java:149,316-335, stages/CombinedIndex.
 * java:28-56, stages/LeveledBitmapsIndex.java:40-54).  The result equals
 * gpjson_buf_from_file + gpjson_index_build; the caller frees *idx_out, then
 * *buf_out.  Errors as those two (GPJSON_ERR_IO when the file size changes).
 * A pipelined form that indexes each staged chunk while the next is read and
 * copied exists in the library but is switched off. */
int gpjson_index_file(gpjson_ctx *ctx, const char *path, int levels, int flags, gpjson_buf **buf_out,
                      gpjson_index **idx_out);
/* Rebuilds an index in place over the same buffers (no allocation). */
int gpjson_index_rebuild(gpjson_index *idx);
int64_t gpjson_index_lines(const gpjson_index *idx);
int gpjson_index_levels(const gpjson_index *idx);
/* Copies one index array to host memory (GPUUtils.readLongs,
 * debug/GPUUtils.java:51-72); *size_bytes gets the full size. */
int gpjson_index_export(const gpjson_index *idx, int which, void *dst, int64_t cap_bytes, int64_t *size_bytes);
void gpjson_index_free(gpjson_index *idx);

/* ---- per-line query ------------------------------------------------------ */

/* Runs nq compiled queries over an index in one interpreter launch
 * (replaces the per-query find_value loop, QueryFunction.java:201-241).
 * spans receives int64[nq][lines][R][2] (start, end) with -1 for "no result";
 * cap_elems is its capacity in int64 elements.  R = results per line. */
int gpjson_query_run(gpjson_ctx *ctx, const gpjson_index *idx, const uint8_t *const *irs,
                     const int64_t *ir_lens, int nq, int R, int64_t *spans, int64_t cap_elems);

/* ---- the whole query() member -------------------------------------------- */

/* query(file, paths[, options]) of the polyglot library object
 * (GPJSONLibrary.java:89-98 -> QueryFunction.call, QueryFunction.java:48-265):
 * compile, stage, index, query, copy spans back.  1 <= npaths <= 255. */
int gpjson_query_file(gpjson_ctx *ctx, const char *file, const char *const *paths, int npaths, int flags,
                      gpjson_result **out);
/* Result shape [query][line][k] (result/Result.java:23-72, ResultArray.java:50-69). */
int64_t gpjson_result_queries(const gpjson_result *res);
int64_t gpjson_result_lines(const gpjson_result *res);
int gpjson_result_per_line(const gpjson_result *res);
/* int64[lines][R][2] spans of query q (the long[] of ResultArray); NULL if q is out of range. */
const int64_t *gpjson_result_spans(const gpjson_result *res, int64_t q);
/* UTF-8 bytes of element [q][line][k] (ResultItem.readArrayElement,
 * result/ResultItem.java:32-52): returns 1 and sets ptr and len, or 0 for null. */
int gpjson_result_value(const gpjson_result *res, int64_t q, int64_t line, int k, const char **ptr, int64_t *len);
void gpjson_result_free(gpjson_result *res);

/* ---- device-side result gather ------------------------------------------- */

typedef struct gpjson_gathered gpjson_gathered;

/* (no reference counterpart: replaces the span copy of QueryFunction.java:240
 * and the per-element decoding from a second map of the file,
 * QueryFunction.java:247-259 / result/ResultItem.java:32-52.)  Runs nq queries
 * over idx, then gathers the bytes of every result element [q][line][k] into
 * one buffer on the device (in element order) and copies it and the spans back
 * once.  *out: per query the file spans int64[lines][R][2] (as gpjson_query_run
 * returns them) and the gathered spans int64[lines][R][2], offsets into the
 * gathered bytes (-1, -1 where there is no result). */
int gpjson_query_gather(gpjson_ctx *ctx, const gpjson_index *idx, const uint8_t *const *irs,
                        const int64_t *ir_lens, int nq, int R, gpjson_gathered **out);
int64_t gpjson_gathered_lines(const gpjson_gathered *g);
int gpjson_gathered_per_line(const gpjson_gathered *g);
/* total gathered bytes; *bytes (optional) points at them */
int64_t gpjson_gathered_bytes(const gpjson_gathered *g, const uint8_t **bytes);
const int64_t *gpjson_gathered_spans(const gpjson_gathered *g, int64_t q);
const int64_t *gpjson_gathered_file_spans(const gpjson_gathered *g, int64_t q);
/* milliseconds on the device: the interpreter launch, the gather kernels, and
 * the copy of spans and bytes to the host */
int gpjson_gathered_timing(const gpjson_gathered *g, double *query_ms, double *gather_ms, double *copy_ms);
void gpjson_gathered_free(gpjson_gathered *g);

/* ---- prepared plain path -------------------------------------------------- */

typedef struct gpjson_plan gpjson_plan;

/* Prepares the plain path "index then nq queries" over a resident buffer:
 * allocates the index and the device span buffer once and builds the index
 * once to size the newline array (QueryFunction.call without the file I/O,
 * functions/QueryFunction.java:165-241).  Later runs allocate nothing. */
int gpjson_plan_create(gpjson_ctx *ctx, gpjson_buf *buf, const uint8_t *const *irs, const int64_t *ir_lens, int nq,
                       gpjson_plan **out);
/* Enqueues one pass (index kernel + one interpreter launch) on the context's
 * stream without waiting; with timed != 0, HIP events bracket both launches. */
int gpjson_plan_run(gpjson_plan *plan, int timed);
/* Waits, then returns index / query time in ms of the last timed run. */
int gpjson_plan_timing(gpjson_plan *plan, double *index_ms, double *query_ms);
/* Whether the plan evaluates lines inside index tiles (gpjson_set_fusion), and
 * (waits) how many lines the last run left to the interpreter (-1 if not). */
int gpjson_plan_fused(gpjson_plan *plan, int *fused, int64_t *interpreted_lines);
int64_t gpjson_plan_lines(const gpjson_plan *plan);
int gpjson_plan_results_per_line(const gpjson_plan *plan);
/* Waits, then copies int64[nq][lines][R][2] spans of the last run to host memory. */
int gpjson_plan_spans(gpjson_plan *plan, int64_t *spans, int64_t cap_elems);
/* The plan's index (valid until gpjson_plan_free). */
const gpjson_index *gpjson_plan_index(const gpjson_plan *plan);
void gpjson_plan_free(gpjson_plan *plan);

/* ---- plain path without allocation (no counterpart in the reference) ------ */

#define GPJSON_NOALLOC_SCRATCH_BYTES (32LL << 20) /* fixed scratch array per GPU */
#define GPJSON_NOALLOC_MAX_QUERIES 8
#define GPJSON_NOALLOC_MAX_IR_BYTES 1024          /* all queries' IR bytes together */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the allocation-free variant of gpjson_plan_create +
 * gpjson_plan_run, running the same index and interpreter kernels.  Indexes
 * the n bytes at `input` (device memory of ctx's device) for the queries'
 * depth and runs the nq (<= GPJSON_NOALLOC_MAX_QUERIES) compiled queries over
 * every line in one interpreter launch.
 *  - Scratch (tile states, newline scratch rows, records, leveled bitmaps and
 *    summaries, the job table and IR bytes) lives in one fixed device array of
 *    GPJSON_NOALLOC_SCRATCH_BYTES per GPU, held from the first context created
 *    on that GPU until the process ends.  An input whose index needs more
 *    (about 42 MiB at 2 levels, 21 MiB at 8) fails with GPJSON_ERR_CAPACITY
 *    before anything is launched.
 *  - Outputs are caller-owned device buffers: newlines int64[newline_cap]
 *    gets the newline index; spans int64[nq][lines][R][2] (R = the largest
 *    STORE count of the queries; capacity span_cap elements); info int64[2]
 *    gets the line count and 1 if it exceeds min(newline_cap,
 *    span_cap / (nq * R * 2)), in which case no query runs.
 *  - Enqueued on the context's stream without a host wait; calls no
 *    allocator.  Every call on one GPU uses that GPU's one array: calls that
 *    overlap on one GPU (two contexts or two threads) race. */
int gpjson_query_noalloc(gpjson_ctx *ctx, const uint8_t *input, int64_t n, const uint8_t *const *irs,
                         const int64_t *ir_lens, int nq, int64_t *newlines, int64_t newline_cap, int64_t *spans,
                         int64_t span_cap, int64_t *info);

/* ---- batched execution (no counterpart in the reference) ------------------ */

/* Prepares a batch of nfiles resident buffers.  File f runs nq_per_file[f]
 * queries; their IRs are irs[k..k+nq_per_file[f]) in file order.  Each file's
 * index has the depth of its deepest query.  Launch shapes come from the
 * context's tuning cache for (bucket of the batch's total bytes, its deepest
 * level, batch), else from the context.  The first run happens here and sizes
 * every array; later runs allocate nothing. */
int gpjson_batch_create(gpjson_ctx *ctx, gpjson_buf *const *bufs, int nfiles, const uint8_t *const *irs,
                        const int64_t *ir_lens, const int *nq_per_file, gpjson_batch **out);
/* Enqueues one run: one persistent index launch over the tiles of every file
 * (atomic tile queue), a device-side line scan, and one interpreter launch over
 * every (file, query, line).  With use_graph, the sequence is captured into a
 * hipGraph on first use and replayed.  No host wait. */
int gpjson_batch_run(gpjson_batch *batch, int use_graph, int timed);
/* Waits; index_ms (-1 for graph runs) and total_ms of the last timed run. */
int gpjson_batch_timing(gpjson_batch *batch, double *index_ms, double *total_ms);
int gpjson_batch_lines(gpjson_batch *batch, int file, int64_t *lines);
int gpjson_batch_results_per_line(const gpjson_batch *batch, int file, int q);
/* Waits, then copies int64[lines][R][2] spans of (file, query q) to host memory. */
int gpjson_batch_spans(gpjson_batch *batch, int file, int q, int64_t *spans, int64_t cap_elems);
/* Batch autotuner: times the batch's own runs under every index tile shape
 * (the chunk of input a tile covers = block x wpt x 64 bytes) and every
 * interpreter shape (mode, block, lanes per line, grid, lines per claim),
 * installs the fastest pair in the batch (its arrays are laid out again, a
 * captured graph is dropped and captured anew on the next graph run; the
 * spans are those of a run with the new shapes) and records it in the tuning
 * cache.  Allocates; waits.  report (optional) gets one line per candidate. */
int gpjson_batch_autotune(gpjson_batch *batch, int iterations, char *report, size_t report_cap);
/* Index launches of a batch: one over every file's tiles (enable 0), or one per
 * level class (enable 1): the files that need a smaller kernel variant than
 * the deepest file's take a launch of their own, so they do not run the
 * deeper variant.  Lays the batch out again and runs it once (allocates;
 * waits).  The batch autotuner picks this too.  get: 1 when two launches run,
 * 0 when one does (one level class), -1 for a NULL batch. */
int gpjson_batch_set_level_split(gpjson_batch *batch, int enable);
int gpjson_batch_get_level_split(const gpjson_batch *batch);
/* The batch's work order (gpjson_set_query_order), -1 for a NULL batch. */
int gpjson_batch_query_order(const gpjson_batch *batch);
/* The batch's launch shapes (the fields of gpjson_get_config). */
int gpjson_batch_get_config(const gpjson_batch *batch, int *index_block, int *index_wpt, int *query_mode,
                            int *query_block, int *query_group, int *query_waves_per_cu, int *lines_per_wave);
void gpjson_batch_free(gpjson_batch *batch);

/* ---- serving (no counterpart in the reference) ---------------------------- */

/* A query server on `device` that keeps files and their indexes (built for
 * `levels` levels) resident in HBM within hbm_budget_bytes (input + index
 * bytes), evicting least-recently-used files (the host copy stays), and
 * coalesces concurrent queries on resident files into one interpreter launch.
 * Owns two streams (serving, admission) and two threads. */
int gpjson_server_create(int device, int64_t hbm_budget_bytes, int levels, gpjson_server **out);
/* Registers n host bytes under `key` (copied; not resident until needed). */
int gpjson_server_add(gpjson_server *server, const char *key, const void *data, int64_t n);
/* Makes `key` resident now (stage + index), evicting as needed; waits. */
int gpjson_server_admit(gpjson_server *server, const char *key);
/* Queues query `path` on file `key`; returns at once with a ticket. */
int gpjson_server_submit(gpjson_server *server, const char *key, const char *path, uint64_t *ticket);
/* Waits for a ticket (timeout_ms < 0: no timeout); gives lines, results per
 * line, latency from submit to completion in microseconds, and (spans not
 * NULL) copies int64[lines][R][2] spans, after which the ticket is forgotten. */
int gpjson_server_wait(gpjson_server *server, uint64_t ticket, int64_t timeout_ms, int64_t *lines, int *R,
                       double *latency_us, int64_t *spans, int64_t cap_elems);
/* Zero-copy variant of gpjson_server_wait: *spans points at the request's
 * int64[lines][R][2] spans where the server keeps them (pinned host memory).
 * The pointer stays valid, and the batch buffer holding it stays reserved,
 * until gpjson_server_release(ticket); release promptly: while every batch
 * buffer holds unreleased results the server launches nothing new. */
int gpjson_server_result(gpjson_server *server, uint64_t ticket, int64_t timeout_ms, const int64_t **spans,
                         int64_t *lines, int *R, double *latency_us);
int gpjson_server_release(gpjson_server *server, uint64_t ticket);
/* Fills min(n, 8) counters: resident bytes, resident files, interpreter
 * launches, requests served by them, largest batch, admissions, evictions,
 * spills (batches whose uncollected results were moved to host memory). */
int gpjson_server_stats(gpjson_server *server, int64_t *out, int n);
void gpjson_server_destroy(gpjson_server *server);

/* ---- launch shapes and the autotuner -------------------------------------- */

/* Current launch shapes of the context: index tile = index_block threads x
 * index_wpt 64-byte words; interpreter mode (query_mode: 0 = one lane group
 * per line; 3 = mode 0, and lines of 8 KB or more are handed to a second
 * launch that gives each a wavefront; modes 1 and 2 no longer exist and are
 * rejected), block size, lanes per line (query_group: 1, 4, 16 or 64),
 * persistent grid in wavefronts per CU and lines a lane group claims at a
 * time.  gpjson_set_config clears the tuning cache.  The reference
 * hard-codes 8x1024 (stages/CombinedIndex.java:11-14) and 512x1024 threads with
 * one thread per line (functions/QueryFunction.java:229). */
int gpjson_get_config(gpjson_ctx *ctx, int *index_block, int *index_wpt, int *query_mode, int *query_block,
                      int *query_group, int *query_waves_per_cu, int *lines_per_wave);
int gpjson_set_config(gpjson_ctx *ctx, int index_block, int index_wpt, int query_mode, int query_block,
                      int query_group, int query_waves_per_cu, int lines_per_wave);
/* In-tile evaluation (no reference counterpart; off by default): a prepared
 * plain pass (gpjson_plan_create) of ONE member-chain query $.k0.k1[.k2] (at
 * most 3 keys, 120 key bytes) evaluates every line that starts and ends inside
 * an index tile in the index kernel itself, from the tile's masks in LDS, and
 * writes the tile's records and leveled words only around the lines that cross
 * into its neighbours; the interpreter then runs on those lines alone.  Spans
 * are the same either way.  Plans prepared while it is off interpret every
 * line from the full index. */
/* Work order of a launch with several queries over one file's lines (no
 * reference counterpart; the reference runs one find_value launch per query,
 * functions/QueryFunction.java:201-241): qblock 0 runs query after query over
 * all lines; qblock B > 0 lets the queries take turns over blocks of B lines,
 * so the queries of a line run close together and share its cached index
 * words.  Results are the same either way.  Setting it clears the tuning
 * cache, as gpjson_set_config does. */
int gpjson_set_query_order(gpjson_ctx *ctx, int qblock);
int gpjson_get_query_order(gpjson_ctx *ctx, int *qblock);
int gpjson_set_fusion(gpjson_ctx *ctx, int enable);
int gpjson_get_fusion(gpjson_ctx *ctx, int *enabled);
/* Times every supported shape on buf with the given queries and records the
 * fastest in the context's tuning cache under (floor(log2 of buf's size), the
 * queries' depth, plain pass): every later plan, one-shot query and index
 * build of an input in that size bucket at that depth uses it; the context's
 * own shapes (gpjson_set_config) serve every other input.  report (optional)
 * gets one text line per candidate.  No reference counterpart: the reference
 * hard-codes its launch shapes (stages/CombinedIndex.java:11-14,
 * functions/QueryFunction.java:229). */
int gpjson_autotune(gpjson_ctx *ctx, gpjson_buf *buf, const uint8_t *const *irs, const int64_t *ir_lens, int nq,
                    int iterations, char *report, size_t report_cap);
/* The tuning cache (no reference counterpart).  gpjson_set_config clears it.
 * A saved cache is a text file with one line per entry, tagged with the
 * device architecture; loading takes the entries of this context's
 * architecture and rejects shapes this build does not support. */
int gpjson_tune_cache_clear(gpjson_ctx *ctx);
int64_t gpjson_tune_cache_size(const gpjson_ctx *ctx);
int gpjson_tune_cache_save(const gpjson_ctx *ctx, const char *path);
int gpjson_tune_cache_load(gpjson_ctx *ctx, const char *path, int *loaded);

/* ---- measurement --------------------------------------------------------- */

/* Times `iterations` passes of index + nq queries on a resident buffer with
 * HIP events on the context's stream (no allocation inside the timed region).
 * ms_per_pass: mean wall time of one pass; index_ms / query_ms: kernel time
 * split of the last pass. */
int gpjson_time_plain(gpjson_ctx *ctx, gpjson_buf *buf, const uint8_t *const *irs, const int64_t *ir_lens,
                      int nq, int iterations, double *ms_per_pass, double *index_ms, double *query_ms);
/* Device-to-device streaming copy of n bytes (16-byte vector loads and
 * stores), timed with HIP events over `iterations` launches: the achievable
 * HBM bandwidth is 2n / ms_per_copy, and the kernel is the known-traffic
 * reference for calibrating FETCH_SIZE / WRITE_SIZE counters.  No counterpart
 * in the reference. */
int gpjson_time_copy(gpjson_ctx *ctx, int64_t n, int iterations, double *ms_per_copy);
/* Device streaming read of n bytes (16-byte vector loads, nothing written),
 * timed with HIP events over `iterations` launches: the achievable HBM read
 * bandwidth is n / ms_per_read.  No counterpart in the reference. */
int gpjson_time_read(gpjson_ctx *ctx, int64_t n, int iterations, double *ms_per_read);

#ifdef __cplusplus
}
#endif

#endif /* GPJSON_H */
