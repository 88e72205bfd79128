// runtime.h -- internal host-side structures of libgpjson_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gpjson.h"
#include "index_common.h"
#include "query_common.h"

namespace gj {

// Kernel entry points (index_kernel.hip, query_kernel.hip).
hipError_t index_launch(const IndexLaunch& L, int block, int wpt, int nlv, bool verify, int grid,
                        hipStream_t stream);
// Zeroes the tile states and the tickets (and *also, when set) with a kernel.
hipError_t index_reset(TileState* states, int64_t ntiles, unsigned int* ticket, hipStream_t stream,
                       const int* run_if = nullptr, unsigned long long* also = nullptr);
hipError_t index_lines(const IndexLaunch& L, int64_t* lines_out, hipStream_t stream);
// Deferred newline bookkeeping of a one-file launch (IndexLaunch.nl_scratch):
// tile_lines, the line count, the newline index and the speculation check
// (*mismatch = 1 when a speculated carry-in was wrong).  block_sums: one int64
// per 256 tiles; *overflow must start at 0.
hipError_t index_finish(const IndexLaunch& L, int64_t* block_sums, int64_t* lines_out, int* mismatch,
                        int* overflow, hipStream_t stream);
// The same for a launch over many files (IndexLaunch.jobs): file_base and
// lines_out hold one int64 per file.
hipError_t index_finish_batch(const IndexLaunch& L, int64_t* block_sums, int64_t* file_base, int64_t* lines_out,
                              int* mismatch, int* overflow, hipStream_t stream);
bool index_config_supported(int block, int wpt);
int64_t index_tile_bytes(int block, int wpt);
hipError_t query_launch(const QueryLaunch& L, int block, int group, int levels, int grid, hipStream_t stream,
                        int long_grid = 0);
// D for a member-chain program (KEY k DOWN)^D STORE END, else 0.
int chain_depth(const uint8_t* ir, int64_t len);
bool query_config_supported(int block, int group);
// gather.hip: result bytes gathered into one buffer (phase 1: sizes and the
// total; phase 2: offsets and bytes).  block_sums holds gather_blocks(n) int64.
int64_t gather_blocks(int64_t n);
hipError_t gather_phase1(const int64_t* spans, int64_t n, int64_t N, int64_t* block_sums, int64_t* total_dev,
                         hipStream_t stream);
hipError_t gather_phase2(const uint8_t* F, int64_t N, const int64_t* spans, int64_t n, const int64_t* block_sums,
                         int64_t* out_spans, uint8_t* out_bytes, int copy_grid, hipStream_t stream);
hipError_t zero_u64(unsigned long long* p, int64_t n, hipStream_t stream);
// Per-file line counts (from index_finish_batch) to query jobs; *overflow = 1
// when a file has more lines than its newline array holds.
hipError_t batch_plan(const FileJob* files, int nfiles, QueryJob* qjobs, const int* job_file, int njobs,
                      int64_t* total_out, const int64_t* lines, int* overflow, hipStream_t stream);

// Error reporting: every C entry point funnels failures through these.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define GJ_HIP(expr)                                  \
  do {                                                \
    hipError_t gj_e_ = (expr);                        \
    if (gj_e_ != hipSuccess) return hip_fail(gj_e_, #expr); \
  } while (0)

// Default index tile: 256 threads x 2 rows (32 KiB), 4 blocks per CU.  On 4 GiB
// of tweets it indexed in 1.72 / 1.68 / 1.73 / 1.78 ms for L = 2, 1, 2, 3
// against 1.79 / 1.73 / 1.79 / 1.88 ms with 3 rows (3 blocks per CU by LDS).
struct IndexTune {
  int block = 256;
  int wpt = 2;
};

// Interpreter modes: one lane group per line (the general interpreter);
// line-long (line mode, and lines of kLongLineBytes or more go to a second
// launch with a wavefront per line).  Modes 1 and 2 (a tiled interpreter and a
// member-chain evaluator) lost every autotuner run and were removed.
enum QueryMode { kLineMode = 0, kLineLongMode = 3 };

// Defaults: the shape the autotuner measures fastest on 4 GiB of tweets for
// $.user.lang: line mode, one lane per line, 16 waves per CU, 4 lines per claim
// (1.11 ms; 1.25 ms with long lines to a wavefront, 1.92 ms with 1 line per
// claim, 1.90 ms with 4 lanes per line).
struct QueryTune {
  int mode = kLineMode;
  int block = 256;
  int group = 1;          // lanes per line
  int waves_per_cu = 16;  // persistent grid size factor
  int lines_per_wave = 4; // lines each group takes per wavefront claim (at most)
  int qblock = 0;         // several queries: they take turns over blocks of this many lines (0: query after query)
  // A prepared plain pass of one member-chain query evaluates the lines inside
  // each index tile in the index kernel and interprets only the others.  Off
  // by default: on 4 GiB of tweets the fused index kernel takes 2.6-3.3 ms
  // against 1.9 ms for the plain one plus 0.4-1.2 ms of interpreter.
  int fuse = 0;
};

int nlv_for_levels(int L);

// Tuning cache: the launch shapes the autotuner measured fastest, keyed by the
// input size's power-of-two bucket (floor(log2 n)), the level count and the
// kind of run (0: a plain pass over one file, 1: a batch, n = its total bytes);
// one context serves one device, so its architecture completes the key.
struct TuneKey {
  int bucket = 0, levels = 0, batch = 0;
  bool operator<(const TuneKey& o) const {
    return std::tie(bucket, levels, batch) < std::tie(o.bucket, o.levels, o.batch);
  }
};
struct TuneEntry {
  IndexTune itune;
  QueryTune qtune;
  double ms = 0;  // the winning shapes' measured time per run
};
int size_bucket(int64_t n);
// The autotuner's candidates (autotune.cpp).
const std::vector<IndexTune>& index_candidates();
std::vector<QueryTune> query_candidates(const QueryTune& base);

// Grow-only device allocation owned by a context.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
};

}  // namespace gj

struct gpjson_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;  // host-to-device staging overlapped with indexing
  int num_cus = 0;
  gj::DevBuf states;      // lookback tile states
  gj::DevBuf ticket;      // index tickets (kTicketBytes), then the bookkeeping flags
  gj::DevBuf jobs;        // device job arrays
  gj::DevBuf irbuf;       // IR bytes of a query launch
  gj::DevBuf nl_scratch;  // an index build's newline offsets per tile, before compaction
  gj::DevBuf gather[4];   // result gather: file spans, block sums + total, gathered spans, bytes
  void* pinned[2] = {nullptr, nullptr};
  size_t pinned_size = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  void* host_scratch = nullptr;  // pinned scratch for small uploads
  size_t host_scratch_size = 0;
  gj::IndexTune itune;
  gj::QueryTune qtune;
  std::string arch;                               // gcnArchName of the device
  std::map<gj::TuneKey, gj::TuneEntry> tune_cache;  // filled by the autotuners
  bool tuning = false;                            // an autotuner is timing candidates: use itune/qtune
  void* noalloc_scratch = nullptr;  // this device's fixed scratch array (gpjson_query_noalloc)
};

struct gpjson_buf {
  gpjson_ctx* ctx = nullptr;
  uint8_t* d = nullptr;
  int64_t n = 0;
  bool owned = false;
};

struct gpjson_index {
  gpjson_ctx* ctx = nullptr;
  const gpjson_buf* buf = nullptr;
  int64_t n = 0, W = 0, Wp = 0;
  int L = 0;
  int flags = 0;
  uint64_t* rec = nullptr;  // records: string, colon, raw-quote and level-0 words (kRecWords per word)
  uint64_t* LB = nullptr;  // leveled bitmaps of levels 1 .. L-1 (level 0 is in the records)
  uint64_t* LS = nullptr;  // level summaries, inside the LB allocation
  int64_t Ws = 0;
  int64_t* NL = nullptr;
  int64_t nl_cap = 0;
  uint64_t* E = nullptr;
  uint64_t* Q = nullptr;
  int64_t lines = 0;
  int64_t* lines_dev = nullptr;  // device copy of the line count
  int64_t* tile_lines = nullptr; // first line id of each index tile (+ total)
  int64_t tile_lines_cap = 0;
  int64_t ntiles = 0;            // tiles of the last build
  int64_t tile_bytes = 0;
};

namespace gj {
// The shapes for a run over n bytes at L levels (batch: 1 for a batch): the
// tuning cache's entry when there is one, else the context's shapes; the
// in-tile evaluation switch always comes from the context.
void shapes_for(const gpjson_ctx* ctx, int64_t n, int L, int batch, IndexTune* it, QueryTune* qt);
// Device job table + launch descriptor for nq queries over one index.
struct PreparedQueries {
  QueryLaunch launch{};
  int long_grid = 0;   // blocks of the long-line launch
  int grid = 0;
  int block = 256;
  int group = 4;
  int levels = 0;
};
// The job table (work counter, jobs, IR bytes) lives in *table, which the
// caller owns for as long as the prepared launch may run.
int prepare_queries(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs, const int64_t* ir_lens,
                    int nq, int R, int64_t* spans_dev, DevBuf* table, PreparedQueries* out);
int launch_prepared(gpjson_ctx* ctx, const PreparedQueries& pq, hipEvent_t before_kernel = nullptr,
                    hipEvent_t after_kernel = nullptr);
// Enqueues reset + index + line count for one index (no host sync).
// tune: tile shape to use (NULL: the context's current shape).
// In-tile evaluation of one member-chain query (IndexLaunch.fuse_*): its keys,
// and the device arrays the bookkeeping fills.
struct FuseSpec {
  int depth = 0;
  int klen[kFuseMaxDepth] = {0, 0, 0};
  uint8_t keys[kFuseKeyBytes] = {};
  int64_t* spans = nullptr;             // int64[lines][2]
  int64_t* list = nullptr;              // lines left to the interpreter
  int64_t list_cap = 0;
  unsigned long long* count = nullptr;  // their number
};
// Fills spec's keys from a member chain of at most kFuseMaxDepth keys whose
// key bytes fit kFuseKeyBytes; false for any other IR.
bool fuse_spec_from_ir(const uint8_t* ir, int64_t len, FuseSpec* spec);
int enqueue_index(gpjson_index* idx, hipEvent_t before_kernel = nullptr, hipEvent_t after_kernel = nullptr,
                  const IndexTune* tune = nullptr, const FuseSpec* fuse = nullptr);
// Enqueues one interpreter launch of nq queries over idx with the line count
// read on the device; spans_dev receives int64[nq][lines][R][2].
int enqueue_queries_dev(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs,
                        const int64_t* ir_lens, int nq, int R, int64_t* spans_dev);
// The launches of one index build over prepared arrays (no allocation): reset,
// index kernel, newline bookkeeping, and the second pass that runs only after
// a wrong speculation.  flags: 16 ints of device memory (zeroed here); the line
// count lands in *lines_dev.
int enqueue_index_launches(const IndexLaunch& L, const IndexTune& t, int levels, bool verify, int* flags,
                           int64_t* lines_dev, hipStream_t stream, hipEvent_t before_kernel = nullptr,
                           hipEvent_t after_kernel = nullptr);
// Device bytes an index holds (every array, after any newline-array growth).
int64_t index_device_bytes(const gpjson_index* idx);
// Context scratch an index build of n bytes needs (tile states, newline
// scratch rows and block sums); a context keeps the largest it has built.
int64_t index_scratch_bytes(int64_t n, const IndexTune& t);
// Levels (deepest DOWN nesting) and result slots (STOREs) a set of IRs needs.
void ir_shape(const uint8_t* const* irs, const int64_t* ir_lens, int nq, int* L_out, int* R_out);
// noalloc.hip: the device address of this device's fixed scratch array (the
// current device must be `device`); looked up once per device.
int noalloc_scratch_setup(int device, void** out);
int upload_irs(gpjson_ctx* ctx, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
               std::vector<int64_t>* offsets, const uint8_t** dev_base);
}  // namespace gj
