// runtime.cpp -- context, staging, index and query entry points of the C ABI.
//
// Host orchestration that replaces the reference's Java driver
// (J/functions/QueryFunction.java:48-265, J/stages/CombinedIndex.java:28-151,
// J/stages/LeveledBitmapsIndex.java:40-103) and its CUDA runtime binding
// (J/gpu/CUDARuntime.java:76-566).  J/ = gpjson/src/main/java/com/koenv/gpjson/.
//
// Differences by design: one stream per context and no host synchronization
// between the index and query launches (the reference synchronizes after every
// kernel, J/gpu/CUDARuntime.java:388); the three host-side carry scans of the
// reference are done on the device (decoupled lookback); no per-call
// cudaMalloc of scratch (grow-only buffers owned by the context).
#include "runtime.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <memory>

#include "jsonpath.h"

namespace gj {

static thread_local std::string tls_error;

int fail(int code, const std::string& msg) {
  tls_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  tls_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  return GPJSON_ERR_HIP;
}

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return GPJSON_OK;
  if (p) {
    hipError_t e = hipFree(p);
    p = nullptr;
    cap = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree");
  }
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return fail(GPJSON_ERR_NO_MEMORY, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  cap = want;
  return GPJSON_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

int nlv_for_levels(int L) { return L <= 2 ? 2 : (L <= 4 ? 4 : 8); }

static int dev_alloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(GPJSON_ERR_NO_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  return GPJSON_OK;
}

static int ensure_host_scratch(gpjson_ctx* ctx, size_t bytes) {
  if (ctx->host_scratch_size >= bytes) {
    // the previous upload from this scratch must have been consumed
    GJ_HIP(hipStreamSynchronize(ctx->stream));
    return GPJSON_OK;
  }
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->host_scratch) (void)hipHostFree(ctx->host_scratch);
  ctx->host_scratch = nullptr;
  size_t want = std::max<size_t>(bytes, 1 << 16);
  hipError_t e = hipHostMalloc(&ctx->host_scratch, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    ctx->host_scratch = nullptr;
    ctx->host_scratch_size = 0;
    return fail(GPJSON_ERR_NO_MEMORY, "pinned allocation failed");
  }
  ctx->host_scratch_size = want;
  return GPJSON_OK;
}

// Staging chunk: 64 MiB, or GPJSON_STAGE_CHUNK_BYTES (at least 4 KiB) when set
// (small chunks let tests cover many chunk boundaries with small files).
static size_t stage_chunk_bytes() {
  const char* v = std::getenv("GPJSON_STAGE_CHUNK_BYTES");
  if (v && *v) {
    const long long b = std::atoll(v);
    if (b >= 4096) return (size_t)b;
  }
  return (size_t)64 << 20;
}

static int ensure_pinned(gpjson_ctx* ctx) {
  if (ctx->pinned[0]) return GPJSON_OK;
  const size_t sz = stage_chunk_bytes();
  for (int b = 0; b < 2; b++) {
    hipError_t e = hipHostMalloc(&ctx->pinned[b], sz, hipHostMallocDefault);
    if (e != hipSuccess) {
      ctx->pinned[b] = nullptr;
      return fail(GPJSON_ERR_NO_MEMORY, "pinned staging allocation failed");
    }
    GJ_HIP(hipEventCreateWithFlags(&ctx->pin_ev[b], hipEventDisableTiming));
  }
  ctx->pinned_size = sz;
  return GPJSON_OK;
}

// Staged copy of n bytes: `fill(dst, off, len)` writes host bytes into a pinned
// chunk, which is then copied asynchronously while the next chunk is filled.
template <class Fill>
static int staged_upload(gpjson_ctx* ctx, uint8_t* d, int64_t n, Fill fill) {
  int rc = ensure_pinned(ctx);
  if (rc) return rc;
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  int64_t chunk = (int64_t)ctx->pinned_size;
  int k = 0;
  for (int64_t off = 0; off < n; off += chunk, k++) {
    int b = k & 1;
    int64_t len = std::min<int64_t>(chunk, n - off);
    GJ_HIP(hipEventSynchronize(ctx->pin_ev[b]));
    rc = fill(ctx->pinned[b], off, len);
    if (rc) {
      (void)hipStreamSynchronize(ctx->stream);
      return rc;
    }
    GJ_HIP(hipMemcpyAsync(d + off, ctx->pinned[b], (size_t)len, hipMemcpyHostToDevice, ctx->stream));
    GJ_HIP(hipEventRecord(ctx->pin_ev[b], ctx->stream));
  }
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  return GPJSON_OK;
}

int upload_irs(gpjson_ctx* ctx, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
               std::vector<int64_t>* offsets, const uint8_t** dev_base) {
  int64_t total = 0;
  offsets->assign(nq, 0);
  for (int q = 0; q < nq; q++) {
    (*offsets)[q] = total;
    total += ir_lens[q];
  }
  int rc = ensure_host_scratch(ctx, (size_t)total + 1);
  if (rc) return rc;
  rc = ctx->irbuf.ensure((size_t)total + 1);
  if (rc) return rc;
  for (int q = 0; q < nq; q++)
    if (ir_lens[q]) std::memcpy((uint8_t*)ctx->host_scratch + (*offsets)[q], irs[q], (size_t)ir_lens[q]);
  GJ_HIP(hipMemcpyAsync(ctx->irbuf.p, ctx->host_scratch, (size_t)total + 1, hipMemcpyHostToDevice, ctx->stream));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  *dev_base = (const uint8_t*)ctx->irbuf.p;
  return GPJSON_OK;
}

// Tile geometry, lookback state and launch descriptor of one index build.
static int index_setup(gpjson_index* idx, const IndexTune& t, IndexLaunch* out) {
  gpjson_ctx* ctx = idx->ctx;
  const int64_t tile_bytes = index_tile_bytes(t.block, t.wpt);
  const int64_t ntiles = (idx->n + tile_bytes - 1) / tile_bytes;
  if (ntiles + 1 > idx->tile_lines_cap) return fail(GPJSON_ERR_CAPACITY, "tile table too small");
  idx->ntiles = ntiles;
  idx->tile_bytes = tile_bytes;
  const size_t states_bytes = (size_t)std::max<int64_t>(ntiles, 1) * sizeof(TileState);
  // growing the tile-state array frees the old one: earlier launches must be done
  if (states_bytes > ctx->states.cap) GJ_HIP(hipStreamSynchronize(ctx->stream));
  int rc = ctx->states.ensure(states_bytes);
  if (rc) return rc;
  rc = ctx->ticket.ensure(4096);
  if (rc) return rc;
  FileJob j{};
  j.F = idx->buf->d;
  j.N = idx->n;
  j.W = idx->W;
  j.Wp = idx->Wp;
  j.rec = idx->rec;
  j.LB = idx->LB;
  j.LS = idx->LS;
  j.Ws = idx->Ws;
  j.NL = idx->NL;
  j.nl_cap = idx->nl_cap;
  j.E = idx->E;
  j.Q = idx->Q;
  j.tile_lines = idx->tile_lines;
  j.tile0 = 0;
  j.ntiles = ntiles;
  j.L = idx->L;
  IndexLaunch L{};
  L.jobs = nullptr;
  L.one = j;
  L.njobs = 1;
  L.total_tiles = ntiles;
  L.states = (TileState*)ctx->states.p;
  L.ticket = (unsigned int*)ctx->ticket.p;
  L.speculate = 1;
  // newline offsets per tile before compaction: a slot per word of every tile
  L.nl_scratch_cap = index_tile_bytes(t.block, t.wpt) / 64;
  // (then one int64 per 256 tiles for the bookkeeping's block sums)
  const size_t scratch_bytes =
      ((size_t)std::max<int64_t>(ntiles, 1) * (size_t)L.nl_scratch_cap + (size_t)(ntiles / 256 + 2)) * 8;
  if (scratch_bytes > ctx->nl_scratch.cap) GJ_HIP(hipStreamSynchronize(ctx->stream));
  rc = ctx->nl_scratch.ensure(scratch_bytes);
  if (rc) return rc;
  L.nl_scratch = (int64_t*)ctx->nl_scratch.p;
  *out = L;
  return GPJSON_OK;
}

int enqueue_index_launches(const IndexLaunch& L, const IndexTune& t, int levels, bool verify, int* flags,
                           int64_t* lines_dev, hipStream_t stream, hipEvent_t before_kernel, hipEvent_t after_kernel) {
  const int64_t ntiles = L.total_tiles;
  const int nlv = nlv_for_levels(levels);
  int64_t* block_sums = L.nl_scratch + ntiles * L.nl_scratch_cap;
  GJ_HIP(index_reset(L.states, ntiles, L.ticket, stream, nullptr, L.fuse_count));
  GJ_HIP(zero_u64((unsigned long long*)flags, 8, stream));
  if (before_kernel) GJ_HIP(hipEventRecord(before_kernel, stream));
  if (ntiles > 0) GJ_HIP(index_launch(L, t.block, t.wpt, nlv, verify, (int)ntiles, stream));
  if (after_kernel) GJ_HIP(hipEventRecord(after_kernel, stream));
  GJ_HIP(index_finish(L, block_sums, lines_dev, flags, flags + 8, stream));
  // A wrong speculation (input that is not NDJSON somewhere): the same build
  // again with every tile taking the lookback.  These launches return at once
  // unless the check above raised its flag, so no host round trip is needed.
  IndexLaunch L2 = L;
  L2.speculate = 0;
  L2.run_if = flags;
  GJ_HIP(index_reset(L.states, ntiles, L.ticket, stream, flags, L.fuse_count));
  // a quarter of the resident grid at most: the launch nearly always returns at once
  if (ntiles > 0) GJ_HIP(index_launch(L2, t.block, t.wpt, nlv, verify, (int)std::min<int64_t>(ntiles, 256), stream));
  GJ_HIP(index_finish(L2, block_sums, lines_dev, flags + 16, flags + 8, stream));
  return GPJSON_OK;
}

int chain_depth(const uint8_t* ir, int64_t len) {
  int64_t pc = 0;
  int D = 0;
  while (pc < len && ir[pc] == kMoveToKey) {
    pc++;
    int v = 0, shift = 0;
    for (;;) {
      if (pc >= len) return 0;
      const uint8_t b = ir[pc++];
      v |= (int)(b & 0x7F) << shift;
      if (!(b & 0x80)) break;
      shift += 7;
      if (shift > 28) return 0;
    }
    pc += v;
    if (pc >= len || ir[pc] != kMoveDown) return 0;
    pc++;
    D++;
  }
  if (D == 0 || pc + 2 != len || ir[pc] != kStoreResult || ir[pc + 1] != kEnd) return 0;
  return D;
}

bool fuse_spec_from_ir(const uint8_t* ir, int64_t len, FuseSpec* spec) {
  const int D = chain_depth(ir, len);
  if (D < 1 || D > kFuseMaxDepth) return false;
  int64_t pc = 0;
  int off = 0;
  for (int l = 0; l < D; l++) {
    pc++;  // KEY
    uint32_t v = 0;
    int sh = 0;
    for (;;) {
      const uint8_t b = ir[pc++];
      v |= (uint32_t)(b & 0x7F) << sh;
      if (!(b & 0x80)) break;
      sh += 7;
    }
    if (off + (int64_t)v > kFuseKeyBytes) return false;
    // the in-tile key test relies on keys without a newline byte
    if (std::memchr(ir + pc, '\n', v)) return false;
    std::memcpy(spec->keys + off, ir + pc, v);
    spec->klen[l] = (int)v;
    off += (int)v;
    pc += v + 1;  // the key, then DOWN
  }
  spec->depth = D;
  return true;
}

int enqueue_index(gpjson_index* idx, hipEvent_t before_kernel, hipEvent_t after_kernel, const IndexTune* tune,
                  const FuseSpec* fuse) {
  gpjson_ctx* ctx = idx->ctx;
  IndexTune t;
  if (tune) t = *tune;
  else shapes_for(ctx, idx->n, idx->L, 0, &t, nullptr);
  IndexLaunch L{};
  int rc = index_setup(idx, t, &L);
  if (rc) return rc;
  if (fuse && fuse->depth > 0) {
    L.fuse_depth = fuse->depth;
    for (int l = 0; l < kFuseMaxDepth; l++) L.fuse_klen[l] = fuse->klen[l];
    std::memcpy(L.fuse_keys, fuse->keys, kFuseKeyBytes);
    L.fuse_spans = fuse->spans;
    L.fuse_list = fuse->list;
    L.fuse_list_cap = fuse->list_cap;
    L.fuse_count = fuse->count;
  }
  // bookkeeping flags, after the tickets: [0] a speculated carry-in was wrong
  // (it gates the second pass), [8] a tile overflowed its scratch row, [16]
  // the second pass's check (not read)
  static_assert(kTicketBytes + 128 <= 4096, "flags past the ticket buffer");
  int* flags = (int*)((uint8_t*)ctx->ticket.p + kTicketBytes);
  return enqueue_index_launches(L, t, idx->L, (idx->flags & GPJSON_BUILD_VERIFY) != 0, flags, idx->lines_dev,
                                ctx->stream, before_kernel, after_kernel);
}

int prepare_queries(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs,
                           const int64_t* ir_lens, int nq, int R, int64_t* spans_dev, DevBuf* table,
                           PreparedQueries* out) {
  int64_t ir_total = 0;
  for (int q = 0; q < nq; q++) ir_total += ir_lens[q];
  // table layout: [work counters][QueryJob x nq][IR bytes][long-line
  // counters, 64 B][long-line items]
  const size_t jobs_off = kQueryCounterWords * 8;
  const size_t jobs_bytes = sizeof(QueryJob) * (size_t)nq;
  const size_t long_off = (jobs_off + jobs_bytes + (size_t)ir_total + 16 + 63) / 64 * 64;
  const int64_t long_cap = (int64_t)nq * (idx->n / kLongLineBytes + 1);
  const size_t bytes = long_off + 64 + (size_t)long_cap * 8;
  // (also waits for the stream, so a table being replaced is no longer read;
  // only the part before the long-line area is uploaded)
  int rc = ensure_host_scratch(ctx, long_off);
  if (rc) return rc;
  rc = table->ensure(bytes);
  if (rc) return rc;
  uint8_t* hbase = (uint8_t*)ctx->host_scratch;
  uint8_t* dbase = (uint8_t*)table->p;
  std::memset(hbase, 0, jobs_off);
  QueryJob* hj = (QueryJob*)(hbase + jobs_off);
  int64_t off = (int64_t)(jobs_off + jobs_bytes);
  QueryTune qt;
  shapes_for(ctx, idx->n, idx->L, 0, nullptr, &qt);
  // the line count is read on the device
  for (int q = 0; q < nq; q++) {
    QueryJob j{};
    j.F = idx->buf->d;
    j.N = idx->n;
    j.NL = idx->NL;
    j.lines = 0;
    j.rec = idx->rec;
    j.LB = idx->LB;
    j.LS = idx->LS;
    j.Ws = idx->Ws;
    j.Wp = idx->Wp;
    j.L = idx->L;
    j.R = R;
    j.ir = dbase + off;
    j.ir_len = ir_lens[q];
    j.out = spans_dev;
    j.item0 = 0;
    std::memcpy(hbase + off, irs[q], (size_t)ir_lens[q]);
    off += ir_lens[q];
    hj[q] = j;
  }
  GJ_HIP(hipMemcpyAsync(dbase, hbase, long_off, hipMemcpyHostToDevice, ctx->stream));
  QueryLaunch L{};
  L.jobs = (const QueryJob*)(dbase + jobs_off);
  L.njobs = nq;
  L.total_items = 0;
  L.lines_dev = idx->lines_dev;
  L.total_dev = nullptr;
  L.counter = (unsigned long long*)dbase;
  L.lines_per_wave = qt.lines_per_wave;
  L.ir_base = dbase + jobs_off + jobs_bytes;
  L.ir_bytes = ir_total;
  L.qblock = nq > 1 ? qt.qblock : 0;
  if (qt.mode == kLineLongMode && qt.group != 64) {
    L.long_count = (unsigned long long*)(dbase + long_off);
    L.long_items = (int64_t*)(dbase + long_off + 64);
    L.long_cap = long_cap;
    L.long_min = kLongLineBytes;
  }
  out->launch = L;
  out->block = qt.block;
  out->group = qt.group;
  out->levels = idx->L;
  const int waves_per_block = qt.block / 64;
  out->grid = std::max(1, ctx->num_cus * qt.waves_per_cu / waves_per_block);
  out->long_grid = ctx->num_cus * 2;
  return GPJSON_OK;
}

int launch_prepared(gpjson_ctx* ctx, const PreparedQueries& pq, hipEvent_t before_kernel, hipEvent_t after_kernel) {
  GJ_HIP(zero_u64(pq.launch.counter, kQueryCounterWords, ctx->stream));
  if (pq.launch.long_count) GJ_HIP(zero_u64(pq.launch.long_count, 2, ctx->stream));
  if (before_kernel) GJ_HIP(hipEventRecord(before_kernel, ctx->stream));
  GJ_HIP(query_launch(pq.launch, pq.block, pq.group, pq.levels, pq.grid, ctx->stream, pq.long_grid));
  if (after_kernel) GJ_HIP(hipEventRecord(after_kernel, ctx->stream));
  return GPJSON_OK;
}

int enqueue_queries_dev(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs,
                        const int64_t* ir_lens, int nq, int R, int64_t* spans_dev) {
  PreparedQueries pq;
  int rc = prepare_queries(ctx, idx, irs, ir_lens, nq, R, spans_dev, &ctx->jobs, &pq);
  if (rc) return rc;
  return launch_prepared(ctx, pq);
}

static int read_lines(gpjson_index* idx) {
  gpjson_ctx* ctx = idx->ctx;
  int rc = ensure_host_scratch(ctx, 64);
  if (rc) return rc;
  GJ_HIP(hipMemcpyAsync(ctx->host_scratch, idx->lines_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(&idx->lines, ctx->host_scratch, 8);
  return GPJSON_OK;
}

// The index object and its device arrays for buf (nothing is launched).
static int alloc_index(gpjson_ctx* ctx, gpjson_buf* buf, int levels, int flags,
                       std::unique_ptr<gpjson_index>* out);
static int index_file(gpjson_ctx* ctx, const char* path, int levels, int flags, gpjson_buf** buf_out,
                      gpjson_index** idx_out);

static void free_index_arrays(gpjson_index* idx) {
  void* ps[] = {idx->rec, idx->LB, idx->NL, idx->E, idx->Q, idx->lines_dev, idx->tile_lines};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  idx->LS = nullptr;
  idx->rec = idx->LB = idx->E = idx->Q = nullptr;
  idx->NL = idx->lines_dev = idx->tile_lines = nullptr;
}

int64_t index_device_bytes(const gpjson_index* idx) {
  int64_t b = (int64_t)idx->W * kRecWords * 8 + leveled_alloc_words(idx->Wp, idx->L) * 8 + idx->nl_cap * 8 + 64 +
              idx->tile_lines_cap * 8;
  if (idx->E) b += idx->W * 8;
  if (idx->Q) b += idx->W * 8;
  return b;
}

int64_t index_scratch_bytes(int64_t n, const IndexTune& t) {
  const int64_t tile_bytes = index_tile_bytes(t.block, t.wpt);
  const int64_t ntiles = std::max<int64_t>((n + tile_bytes - 1) / tile_bytes, 1);
  return ntiles * (int64_t)sizeof(TileState) + (ntiles * (tile_bytes / 64) + ntiles / 256 + 2) * 8;
}

// Builds (or rebuilds) an index; grows the newline array if the first pass
// finds more lines than it holds.
static int build_index(gpjson_index* idx) {
  int rc = enqueue_index(idx, nullptr, nullptr);
  if (rc) return rc;
  rc = read_lines(idx);
  if (rc) return rc;
  if (idx->lines > idx->nl_cap) {
    (void)hipFree(idx->NL);
    idx->NL = nullptr;
    idx->nl_cap = idx->lines;
    rc = dev_alloc((void**)&idx->NL, (size_t)idx->nl_cap * 8);
    if (rc) return rc;
    rc = enqueue_index(idx, nullptr, nullptr);
    if (rc) return rc;
    rc = read_lines(idx);
    if (rc) return rc;
  }
  return GPJSON_OK;
}

}  // namespace gj

using namespace gj;

extern "C" {

const char* gpjson_last_error(void) { return tls_error.c_str(); }

int gpjson_ctx_create(int device, gpjson_ctx** out) {
  if (!out) return fail(GPJSON_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device < 0 || device >= count)
    return fail(GPJSON_ERR_INVALID_ARG, "device " + std::to_string(device) + " out of range (" +
                                            std::to_string(count) + " devices)");
  GJ_HIP(hipSetDevice(device));
  std::unique_ptr<gpjson_ctx> ctx(new gpjson_ctx());
  ctx->device = device;
  GJ_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  GJ_HIP(hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device));
  hipDeviceProp_t prop;
  GJ_HIP(hipGetDeviceProperties(&prop, device));
  ctx->arch = prop.gcnArchName;
  const int rc = noalloc_scratch_setup(device, &ctx->noalloc_scratch);
  if (rc) {
    (void)hipStreamDestroy(ctx->stream);
    return rc;
  }
  *out = ctx.release();
  return GPJSON_OK;
}

void gpjson_ctx_destroy(gpjson_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->states.release();
  ctx->ticket.release();
  ctx->jobs.release();
  ctx->irbuf.release();
  ctx->nl_scratch.release();
  for (auto& b : ctx->gather) b.release();
  for (int b = 0; b < 2; b++) {
    if (ctx->pinned[b]) (void)hipHostFree(ctx->pinned[b]);
    if (ctx->pin_ev[b]) (void)hipEventDestroy(ctx->pin_ev[b]);
  }
  if (ctx->host_scratch) (void)hipHostFree(ctx->host_scratch);
  if (ctx->copy_stream) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    (void)hipStreamDestroy(ctx->copy_stream);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int gpjson_ctx_synchronize(gpjson_ctx* ctx) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  GJ_HIP(hipSetDevice(ctx->device));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  return GPJSON_OK;
}

int gpjson_ctx_num_cus(gpjson_ctx* ctx, int* out) {
  if (!ctx || !out) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  *out = ctx->num_cus;
  return GPJSON_OK;
}

int gpjson_compile(const char* path, uint8_t* ir, size_t cap, size_t* ir_len, int* max_depth, int* num_results) {
  if (!path) return fail(GPJSON_ERR_INVALID_ARG, "path is NULL");
  CompiledPath cp;
  std::string err;
  PathStatus st = compile_path(path, &cp, &err);
  if (st == PathStatus::kUnsupported) return fail(GPJSON_ERR_UNSUPPORTED, err);
  if (st == PathStatus::kParseError) return fail(GPJSON_ERR_PARSE, err);
  if (ir_len) *ir_len = cp.ir.size();
  if (max_depth) *max_depth = cp.max_depth;
  if (num_results) *num_results = cp.num_results;
  if (ir && cap) std::memcpy(ir, cp.ir.data(), std::min(cap, cp.ir.size()));
  if (ir && cap < cp.ir.size()) return fail(GPJSON_ERR_CAPACITY, "IR buffer too small");
  return GPJSON_OK;
}

int gpjson_buf_from_host(gpjson_ctx* ctx, const void* data, int64_t n, gpjson_buf** out) {
  if (!ctx || !out || (n > 0 && !data) || n < 0) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<gpjson_buf> b(new gpjson_buf());
  b->ctx = ctx;
  b->n = n;
  b->owned = true;
  int rc = dev_alloc((void**)&b->d, (size_t)n + 64);
  if (rc) return rc;
  const uint8_t* src = (const uint8_t*)data;
  rc = staged_upload(ctx, b->d, n, [&](void* dst, int64_t off, int64_t len) -> int {
    std::memcpy(dst, src + off, (size_t)len);
    return GPJSON_OK;
  });
  if (rc) {
    (void)hipFree(b->d);
    return rc;
  }
  *out = b.release();
  return GPJSON_OK;
}

int gpjson_buf_from_file(gpjson_ctx* ctx, const char* path, gpjson_buf** out) {
  if (!ctx || !out || !path) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(GPJSON_ERR_IO, std::string("Failed to open file ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return fail(GPJSON_ERR_IO, "Failed to get size of file");
  }
  const int64_t n = (int64_t)st.st_size;
  std::unique_ptr<gpjson_buf> b(new gpjson_buf());
  b->ctx = ctx;
  b->n = n;
  b->owned = true;
  int rc = dev_alloc((void**)&b->d, (size_t)n + 64);
  if (rc) {
    close(fd);
    return rc;
  }
  rc = staged_upload(ctx, b->d, n, [&](void* dst, int64_t off, int64_t len) -> int {
    int64_t got = 0;
    while (got < len) {
      ssize_t r = pread(fd, (uint8_t*)dst + got, (size_t)(len - got), (off_t)(off + got));
      if (r <= 0) return fail(GPJSON_ERR_IO, "Size of file has changed while reading");
      got += r;
    }
    return GPJSON_OK;
  });
  struct stat st2;
  if (rc == GPJSON_OK && (fstat(fd, &st2) != 0 || st2.st_size != st.st_size))
    rc = fail(GPJSON_ERR_IO, "Size of file has changed while reading");
  close(fd);
  if (rc) {
    (void)hipFree(b->d);
    return rc;
  }
  *out = b.release();
  return GPJSON_OK;
}

int gpjson_buf_wrap_device(gpjson_ctx* ctx, const void* device_ptr, int64_t n, gpjson_buf** out) {
  if (!ctx || !out || n < 0 || (n > 0 && !device_ptr)) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  gpjson_buf* b = new gpjson_buf();
  b->ctx = ctx;
  b->d = (uint8_t*)device_ptr;
  b->n = n;
  b->owned = false;
  *out = b;
  return GPJSON_OK;
}

int64_t gpjson_buf_size(const gpjson_buf* buf) { return buf ? buf->n : -1; }

void gpjson_buf_free(gpjson_buf* buf) {
  if (!buf) return;
  if (buf->owned && buf->d) {
    (void)hipSetDevice(buf->ctx->device);
    (void)hipStreamSynchronize(buf->ctx->stream);
    (void)hipFree(buf->d);
  }
  delete buf;
}

int gpjson_index_build(gpjson_ctx* ctx, gpjson_buf* buf, int levels, int flags, gpjson_index** out) {
  if (!ctx || !buf || !out) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  if (levels < 0 || levels > GPJSON_MAX_LEVELS)
    return fail(GPJSON_ERR_CAPACITY, "levels must be in [0, " + std::to_string(GPJSON_MAX_LEVELS) + "], got " +
                                         std::to_string(levels));
  *out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<gpjson_index> idx;
  int rc = alloc_index(ctx, buf, levels, flags, &idx);
  if (rc == GPJSON_OK) rc = build_index(idx.get());
  if (rc) {
    if (idx) free_index_arrays(idx.get());
    return rc;
  }
  *out = idx.release();
  return GPJSON_OK;
}

int gpjson_index_file(gpjson_ctx* ctx, const char* path, int levels, int flags, gpjson_buf** buf_out,
                      gpjson_index** idx_out) {
  if (!ctx || !path || !buf_out || !idx_out) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  if (levels < 0 || levels > GPJSON_MAX_LEVELS)
    return fail(GPJSON_ERR_CAPACITY, "levels must be in [0, " + std::to_string(GPJSON_MAX_LEVELS) + "], got " +
                                         std::to_string(levels));
  *buf_out = nullptr;
  *idx_out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  return gj::index_file(ctx, path, levels, flags, buf_out, idx_out);
}

}  // extern "C"

namespace gj {

static int alloc_index(gpjson_ctx* ctx, gpjson_buf* buf, int levels, int flags,
                       std::unique_ptr<gpjson_index>* out) {
  std::unique_ptr<gpjson_index>& idx = *out;
  idx.reset(new gpjson_index());
  idx->ctx = ctx;
  idx->buf = buf;
  idx->n = buf->n;
  idx->W = (buf->n + 63) / 64;
  idx->Wp = (idx->W + 7) / 8 * 8;
  idx->L = levels;
  idx->flags = flags;
  idx->nl_cap = buf->n / 256 + 4096;
  int rc = GPJSON_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc == GPJSON_OK) rc = dev_alloc(p, bytes);
  };
  alloc((void**)&idx->rec, (size_t)idx->W * kRecWords * 8);
  alloc((void**)&idx->LB, (size_t)leveled_alloc_words(idx->Wp, levels) * 8);
  idx->Ws = summary_stride(idx->Wp);
  if (idx->LB) idx->LS = summaries_of(idx->LB, idx->Wp, levels);
  alloc((void**)&idx->NL, (size_t)idx->nl_cap * 8);
  alloc((void**)&idx->lines_dev, 64);
  // one entry per tile of the smallest supported tile size, plus the total
  idx->tile_lines_cap = buf->n / index_tile_bytes(256, 1) + 2;
  alloc((void**)&idx->tile_lines, (size_t)idx->tile_lines_cap * 8);
  if (flags & GPJSON_BUILD_VERIFY) {
    alloc((void**)&idx->E, (size_t)idx->W * 8);
    alloc((void**)&idx->Q, (size_t)idx->W * 8);
  }
  return rc;
}

// Staging overlapped with indexing.  The file is read into the two pinned
// buffers in turn and each chunk is copied to the device on the copy stream;
// after a chunk's copy, one index launch on the context stream takes the tiles
// that chunk completes.  The tile states persist across those launches (reset
// once), so a launch's lookbacks read the prefixes earlier launches published,
// and the result is the one-launch index.  Reading chunk k+2, copying chunk k+1
// and indexing chunk k run at once.
//
// Switched off (kOverlapStaging): with 64 MiB chunks a 1 GiB file never
// finished indexing on the GPU, while files of a few MB in 1 MiB and 20 KB
// chunks matched the one-launch index; the cause is not known.  Until it is,
// the file is read and copied whole, then indexed in one launch.
constexpr bool kOverlapStaging = false;

static int index_file(gpjson_ctx* ctx, const char* path, int levels, int flags, gpjson_buf** buf_out,
                      gpjson_index** idx_out) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(GPJSON_ERR_IO, std::string("Failed to open file ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return fail(GPJSON_ERR_IO, "Failed to get size of file");
  }
  const int64_t n = (int64_t)st.st_size;
  std::unique_ptr<gpjson_buf> b(new gpjson_buf());
  b->ctx = ctx;
  b->n = n;
  b->owned = true;
  std::unique_ptr<gpjson_index> idx;
  bool launched = false;
  auto run = [&]() -> int {
    int rc = dev_alloc((void**)&b->d, (size_t)n + 64);
    if (rc) return rc;
    rc = alloc_index(ctx, b.get(), levels, flags, &idx);
    if (rc) return rc;
    if (!kOverlapStaging) {
      rc = staged_upload(ctx, b->d, n, [&](void* dst, int64_t off, int64_t len) -> int {
        int64_t got = 0;
        while (got < len) {
          const ssize_t r = pread(fd, (uint8_t*)dst + got, (size_t)(len - got), (off_t)(off + got));
          if (r <= 0) return fail(GPJSON_ERR_IO, "Size of file has changed while reading");
          got += r;
        }
        return GPJSON_OK;
      });
      if (rc) return rc;
      struct stat st2;
      if (fstat(fd, &st2) != 0 || st2.st_size != st.st_size)
        return fail(GPJSON_ERR_IO, "Size of file has changed while reading");
      return build_index(idx.get());
    }
    rc = ensure_pinned(ctx);
    if (rc) return rc;
    if (!ctx->copy_stream) GJ_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    GJ_HIP(hipStreamSynchronize(ctx->stream));
    GJ_HIP(hipStreamSynchronize(ctx->copy_stream));
    IndexTune t;
    shapes_for(ctx, n, levels, 0, &t, nullptr);
    IndexLaunch L{};
    rc = index_setup(idx.get(), t, &L);
    if (rc) return rc;
    L.nl_scratch = nullptr;  // the tiles' lookback gives the newline offsets
    const int64_t ntiles = L.total_tiles, tile_bytes = idx->tile_bytes;
    const int nlv = nlv_for_levels(levels);
    const bool verify = (flags & GPJSON_BUILD_VERIFY) != 0;
    launched = true;
    GJ_HIP(index_reset(L.states, ntiles, L.ticket, ctx->stream));
    const int64_t chunk = (int64_t)ctx->pinned_size;
    int64_t done_tiles = 0;
    int k = 0;
    for (int64_t off = 0; off < n; off += chunk, k++) {
      const int pb = k & 1;
      const int64_t len = std::min<int64_t>(chunk, n - off);
      GJ_HIP(hipEventSynchronize(ctx->pin_ev[pb]));  // its previous copy is done
      int64_t got = 0;
      while (got < len) {
        const ssize_t r = pread(fd, (uint8_t*)ctx->pinned[pb] + got, (size_t)(len - got), (off_t)(off + got));
        if (r <= 0) return fail(GPJSON_ERR_IO, "Size of file has changed while reading");
        got += r;
      }
      GJ_HIP(hipMemcpyAsync(b->d + off, ctx->pinned[pb], (size_t)len, hipMemcpyHostToDevice, ctx->copy_stream));
      GJ_HIP(hipEventRecord(ctx->pin_ev[pb], ctx->copy_stream));
      const int64_t end_tiles = off + len >= n ? ntiles : (off + len) / tile_bytes;
      if (end_tiles > done_tiles) {
        GJ_HIP(hipStreamWaitEvent(ctx->stream, ctx->pin_ev[pb], 0));
        L.tile_begin = done_tiles;
        L.total_tiles = end_tiles;
        GJ_HIP(zero_u64((unsigned long long*)L.ticket, kTicketBytes / 8, ctx->stream));
        GJ_HIP(index_launch(L, t.block, t.wpt, nlv, verify, (int)(end_tiles - done_tiles), ctx->stream));
        done_tiles = end_tiles;
      }
    }
    L.tile_begin = 0;
    L.total_tiles = ntiles;
    GJ_HIP(index_lines(L, idx->lines_dev, ctx->stream));
    rc = read_lines(idx.get());
    if (rc) return rc;
    struct stat st2;
    if (fstat(fd, &st2) != 0 || st2.st_size != st.st_size)
      return fail(GPJSON_ERR_IO, "Size of file has changed while reading");
    if (idx->lines > idx->nl_cap) rc = build_index(idx.get());  // a larger newline index, from the device copy
    return rc;
  };
  const int rc = run();
  close(fd);
  if (rc) {
    if (launched) {
      (void)hipStreamSynchronize(ctx->copy_stream);
      (void)hipStreamSynchronize(ctx->stream);
    }
    if (idx) free_index_arrays(idx.get());
    if (b->d) (void)hipFree(b->d);
    return rc;
  }
  *buf_out = b.release();
  idx->buf = *buf_out;
  *idx_out = idx.release();
  return GPJSON_OK;
}

}  // namespace gj

extern "C" {

int gpjson_index_rebuild(gpjson_index* idx) {
  if (!idx) return fail(GPJSON_ERR_INVALID_ARG, "idx is NULL");
  GJ_HIP(hipSetDevice(idx->ctx->device));
  return build_index(idx);
}

int64_t gpjson_index_lines(const gpjson_index* idx) { return idx ? idx->lines : -1; }
int gpjson_index_levels(const gpjson_index* idx) { return idx ? idx->L : -1; }

int gpjson_index_export(const gpjson_index* idx, int which, void* dst, int64_t cap_bytes, int64_t* size_bytes) {
  if (!idx) return fail(GPJSON_ERR_INVALID_ARG, "idx is NULL");
  GJ_HIP(hipSetDevice(idx->ctx->device));
  GJ_HIP(hipStreamSynchronize(idx->ctx->stream));
  int64_t size = 0;
  const void* src = nullptr;
  switch (which) {
    case GPJSON_INDEX_NEWLINE: size = idx->lines * 8; src = idx->NL; break;
    case GPJSON_INDEX_STRING: size = idx->W * 8; src = idx->rec; break;
    case GPJSON_INDEX_LEVELED: size = idx->W * 8 * idx->L; src = idx->LB; break;
    case GPJSON_INDEX_ESCAPE: size = idx->W * 8; src = idx->E; break;
    case GPJSON_INDEX_QUOTE: size = idx->W * 8; src = idx->Q; break;
    case GPJSON_INDEX_COLON: size = idx->W * 8; src = idx->rec ? idx->rec + 1 : nullptr; break;
    case GPJSON_INDEX_RAW_QUOTE: size = idx->W * 8; src = idx->rec ? idx->rec + 2 : nullptr; break;
    case GPJSON_INDEX_LEVEL_SUMMARY: size = (idx->W + 63) / 64 * 8 * idx->L; src = idx->LS; break;
    default: return fail(GPJSON_ERR_INVALID_ARG, "unknown index array " + std::to_string(which));
  }
  if (size_bytes) *size_bytes = size;
  if (!dst) return GPJSON_OK;
  if (!src && size > 0) return fail(GPJSON_ERR_INVALID_ARG, "array not built (needs GPJSON_BUILD_VERIFY)");
  if (cap_bytes < size) return fail(GPJSON_ERR_CAPACITY, "export buffer too small");
  if (size == 0) return GPJSON_OK;
  if (which == GPJSON_INDEX_LEVELED) {
    // level 0 from the records, levels 1 .. L-1 from their level-major rows
    GJ_HIP(hipMemcpy2D(dst, 8, idx->rec + kRecLevel0, (size_t)kRecWords * 8, 8, (size_t)idx->W,
                       hipMemcpyDeviceToHost));
    if (idx->L > 1)
      GJ_HIP(hipMemcpy2D((uint8_t*)dst + idx->W * 8, (size_t)idx->W * 8, src, (size_t)idx->Wp * 8,
                         (size_t)idx->W * 8, (size_t)(idx->L - 1), hipMemcpyDeviceToHost));
  } else if (which == GPJSON_INDEX_STRING || which == GPJSON_INDEX_COLON || which == GPJSON_INDEX_RAW_QUOTE) {
    // one word of each record
    GJ_HIP(hipMemcpy2D(dst, 8, src, (size_t)kRecWords * 8, 8, (size_t)idx->W, hipMemcpyDeviceToHost));
  } else if (which == GPJSON_INDEX_LEVEL_SUMMARY) {
    const size_t row = (size_t)(idx->W + 63) / 64 * 8;
    GJ_HIP(hipMemcpy2D(dst, row, src, (size_t)idx->Ws * 8, row, (size_t)idx->L, hipMemcpyDeviceToHost));
  } else {
    GJ_HIP(hipMemcpy(dst, src, (size_t)size, hipMemcpyDeviceToHost));
  }
  return GPJSON_OK;
}

void gpjson_index_free(gpjson_index* idx) {
  if (!idx) return;
  (void)hipSetDevice(idx->ctx->device);
  (void)hipStreamSynchronize(idx->ctx->stream);
  free_index_arrays(idx);
  delete idx;
}

int gpjson_query_run(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs, const int64_t* ir_lens,
                     int nq, int R, int64_t* spans, int64_t cap_elems) {
  if (!ctx || !idx || !irs || !ir_lens || !spans) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  if (nq < 1 || nq > GPJSON_MAX_QUERIES) return fail(GPJSON_ERR_INVALID_ARG, "nq must be in [1, 255]");
  if (R < 1) return fail(GPJSON_ERR_INVALID_ARG, "R must be >= 1");
  const int64_t need = (int64_t)nq * idx->lines * R * 2;
  if (cap_elems < need) return fail(GPJSON_ERR_CAPACITY, "span buffer too small");
  GJ_HIP(hipSetDevice(ctx->device));
  int64_t* d = nullptr;
  int rc = dev_alloc((void**)&d, (size_t)need * 8);
  if (rc) return rc;
  rc = enqueue_queries_dev(ctx, idx, irs, ir_lens, nq, R, d);
  if (rc == GPJSON_OK) {
    hipError_t e = hipMemcpyAsync(spans, d, (size_t)need * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(e, "query span copy");
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// query(file, paths, options) and its Result
// ---------------------------------------------------------------------------

struct gpjson_gathered {
  int64_t nq = 0, lines = 0;
  int R = 0;
  std::vector<int64_t> file_spans;  // [nq][lines][R][2] into the file
  std::vector<int64_t> spans;       // [nq][lines][R][2] into bytes
  std::vector<uint8_t> bytes;
  double query_ms = 0, gather_ms = 0, copy_ms = 0;
};

struct gpjson_result {
  int64_t nq = 0, lines = 0;
  int R = 0;
  std::vector<std::vector<int64_t>> spans;
  const uint8_t* map = nullptr;
  int64_t map_len = 0;
  gpjson_gathered* gathered = nullptr;  // values from the gathered bytes (no map)
};

namespace {

// Interpreter launch + device gather + one copy back (gpjson_query_gather).
int run_gather(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
               int R, gpjson_gathered* g) {
  using namespace gj;
  const int64_t n = (int64_t)nq * idx->lines * R;  // elements
  g->nq = nq;
  g->lines = idx->lines;
  g->R = R;
  int rc = ctx->gather[0].ensure((size_t)std::max<int64_t>(n, 1) * 16);
  if (!rc) rc = ctx->gather[1].ensure((size_t)(gather_blocks(n) + 8) * 8);
  if (rc) return rc;
  int64_t* spans_dev = (int64_t*)ctx->gather[0].p;
  int64_t* sums = (int64_t*)ctx->gather[1].p + 8;
  int64_t* total_dev = (int64_t*)ctx->gather[1].p;
  hipEvent_t ev[4];
  for (auto& e : ev) GJ_HIP(hipEventCreate(&e));
  auto done = [&](int r) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    return r;
  };
  if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) return done(hip_fail(hipErrorUnknown, "event"));
  rc = enqueue_queries_dev(ctx, idx, irs, ir_lens, nq, R, spans_dev);
  if (rc) return done(rc);
  hipError_t e = hipEventRecord(ev[1], ctx->stream);
  if (e == hipSuccess) e = gather_phase1(spans_dev, n, idx->n, sums, total_dev, ctx->stream);
  int64_t total = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&total, total_dev, 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return done(hip_fail(e, "gather sizes"));
  rc = ctx->gather[2].ensure((size_t)std::max<int64_t>(n, 1) * 16);
  if (!rc) rc = ctx->gather[3].ensure((size_t)std::max<int64_t>(total, 1));
  if (rc) return done(rc);
  int64_t* out_spans = (int64_t*)ctx->gather[2].p;
  uint8_t* out_bytes = (uint8_t*)ctx->gather[3].p;
  e = gather_phase2(idx->buf->d, idx->n, spans_dev, n, sums, out_spans, out_bytes, ctx->num_cus * 8, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev[2], ctx->stream);
  g->file_spans.resize((size_t)n * 2);
  g->spans.resize((size_t)n * 2);
  g->bytes.resize((size_t)total);
  if (e == hipSuccess && n > 0)
    e = hipMemcpyAsync(g->file_spans.data(), spans_dev, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && n > 0)
    e = hipMemcpyAsync(g->spans.data(), out_spans, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && total > 0)
    e = hipMemcpyAsync(g->bytes.data(), out_bytes, (size_t)total, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev[3], ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return done(hip_fail(e, "gather"));
  float a = 0, b = 0, c = 0;
  (void)hipEventElapsedTime(&a, ev[0], ev[1]);
  (void)hipEventElapsedTime(&b, ev[1], ev[2]);
  (void)hipEventElapsedTime(&c, ev[2], ev[3]);
  g->query_ms = a;
  g->gather_ms = b;  // includes the one 8-byte read of the total
  g->copy_ms = c;
  return done(GPJSON_OK);
}

}  // namespace

extern "C" {

int gpjson_query_file(gpjson_ctx* ctx, const char* file, const char* const* paths, int npaths, int flags,
                      gpjson_result** out) {
  if (!ctx || !file || !out) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (npaths > GPJSON_MAX_QUERIES) return fail(GPJSON_ERR_INVALID_ARG, "Maximum size of queries is 255");
  if (npaths < 1 || !paths) return fail(GPJSON_ERR_INVALID_ARG, "expected queries to contain at least one query");
  std::vector<CompiledPath> cps(npaths);
  int L = 0, R = 0;
  for (int i = 0; i < npaths; i++) {
    std::string err;
    PathStatus st = compile_path(paths[i] ? paths[i] : "", &cps[i], &err);
    if (st == PathStatus::kUnsupported) {
      if (flags & GPJSON_QUERY_DISABLE_FALLBACK)
        return fail(GPJSON_ERR_UNSUPPORTED, "Unsupported JSON path (" + std::to_string(i) + ", fallback disabled): " + err);
      return fail(GPJSON_FALLBACK_REQUIRED,
                  "Unsupported JSON path (" + std::to_string(i) + "): " + err +
                      "; fallback allowed: evaluate the query with the caller's JSONPath library");
    }
    if (st == PathStatus::kParseError)
      return fail(GPJSON_ERR_PARSE, "Failed to parse JSON path (" + std::to_string(i) + "): " + err);
    L = std::max(L, cps[i].max_depth);
    R = std::max(R, cps[i].num_results);
  }
  if (L > GPJSON_MAX_LEVELS) return fail(GPJSON_ERR_CAPACITY, "JSON path deeper than 64 levels");
  if (R < 1) R = 1;
  GJ_HIP(hipSetDevice(ctx->device));
  gpjson_buf* buf = nullptr;
  gpjson_index* idx = nullptr;
  int rc = gpjson_index_file(ctx, file, L, 0, &buf, &idx);
  if (rc) return rc;
  std::unique_ptr<gpjson_result> res(new gpjson_result());
  res->nq = npaths;
  res->lines = idx->lines;
  res->R = R;
  res->spans.resize(npaths);
  std::vector<const uint8_t*> irs(npaths);
  std::vector<int64_t> lens(npaths);
  for (int i = 0; i < npaths; i++) {
    irs[i] = cps[i].ir.data();
    lens[i] = (int64_t)cps[i].ir.size();
  }
  const int64_t per = idx->lines * R * 2;
  if (flags & GPJSON_QUERY_GATHER) {
    // values come from the bytes gathered on the device: no map of the file
    std::unique_ptr<gpjson_gathered> g(new gpjson_gathered());
    rc = run_gather(ctx, idx, irs.data(), lens.data(), npaths, R, g.get());
    gpjson_index_free(idx);
    gpjson_buf_free(buf);
    if (rc) return rc;
    for (int i = 0; i < npaths; i++)
      res->spans[i].assign(g->file_spans.begin() + per * i, g->file_spans.begin() + per * (i + 1));
    res->gathered = g.release();
    *out = res.release();
    return GPJSON_OK;
  }
  std::vector<int64_t> all((size_t)(per * npaths));
  rc = gpjson_query_run(ctx, idx, irs.data(), lens.data(), npaths, R, all.data(), (int64_t)all.size());
  gpjson_index_free(idx);
  gpjson_buf_free(buf);
  if (rc) return rc;
  for (int i = 0; i < npaths; i++) res->spans[i].assign(all.begin() + per * i, all.begin() + per * (i + 1));
  // map the file for value decoding (QueryFunction.java:247-259)
  int fd = open(file, O_RDONLY);
  if (fd < 0) return fail(GPJSON_ERR_IO, "Failed to open file");
  struct stat st;
  if (fstat(fd, &st) == 0 && st.st_size > 0) {
    void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m != MAP_FAILED) {
      res->map = (const uint8_t*)m;
      res->map_len = (int64_t)st.st_size;
    }
  }
  close(fd);
  *out = res.release();
  return GPJSON_OK;
}

int64_t gpjson_result_queries(const gpjson_result* res) { return res ? res->nq : -1; }
int64_t gpjson_result_lines(const gpjson_result* res) { return res ? res->lines : -1; }
int gpjson_result_per_line(const gpjson_result* res) { return res ? res->R : -1; }

const int64_t* gpjson_result_spans(const gpjson_result* res, int64_t q) {
  if (!res || q < 0 || q >= res->nq) return nullptr;
  return res->spans[q].data();
}

int gpjson_result_value(const gpjson_result* res, int64_t q, int64_t line, int k, const char** ptr, int64_t* len) {
  if (!res || q < 0 || q >= res->nq || line < 0 || line >= res->lines || k < 0 || k >= res->R) {
    fail(GPJSON_ERR_INVALID_ARG, "index out of range");
    return -1;
  }
  if (res->gathered) {
    const gpjson_gathered* g = res->gathered;
    const int64_t* gs = g->spans.data() + ((q * res->lines + line) * res->R + k) * 2;
    if (gs[0] == -1) return 0;
    if (ptr) *ptr = (const char*)(g->bytes.data() + gs[0]);
    if (len) *len = gs[1] - gs[0];
    return 1;
  }
  const int64_t* s = res->spans[q].data() + (line * res->R + k) * 2;
  if (s[0] == -1) return 0;
  int64_t a = std::min(s[0], res->map_len);
  int64_t b = std::min(std::max(s[1], a), res->map_len);
  if (ptr) *ptr = (const char*)(res->map + a);
  if (len) *len = b - a;
  return 1;
}

void gpjson_result_free(gpjson_result* res) {
  if (!res) return;
  if (res->map) munmap((void*)res->map, (size_t)res->map_len);
  delete res->gathered;
  delete res;
}

int gpjson_query_gather(gpjson_ctx* ctx, const gpjson_index* idx, const uint8_t* const* irs, const int64_t* ir_lens,
                        int nq, int R, gpjson_gathered** out) {
  if (!ctx || !idx || !irs || !ir_lens || !out) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (nq < 1 || nq > GPJSON_MAX_QUERIES) return fail(GPJSON_ERR_INVALID_ARG, "nq must be in [1, 255]");
  if (R < 1) return fail(GPJSON_ERR_INVALID_ARG, "R must be >= 1");
  GJ_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<gpjson_gathered> g(new gpjson_gathered());
  const int rc = run_gather(ctx, idx, irs, ir_lens, nq, R, g.get());
  if (rc) return rc;
  *out = g.release();
  return GPJSON_OK;
}

int64_t gpjson_gathered_lines(const gpjson_gathered* g) { return g ? g->lines : -1; }
int gpjson_gathered_per_line(const gpjson_gathered* g) { return g ? g->R : -1; }

int64_t gpjson_gathered_bytes(const gpjson_gathered* g, const uint8_t** bytes) {
  if (!g) return -1;
  if (bytes) *bytes = g->bytes.data();
  return (int64_t)g->bytes.size();
}

const int64_t* gpjson_gathered_spans(const gpjson_gathered* g, int64_t q) {
  if (!g || q < 0 || q >= g->nq) return nullptr;
  return g->spans.data() + q * g->lines * g->R * 2;
}

const int64_t* gpjson_gathered_file_spans(const gpjson_gathered* g, int64_t q) {
  if (!g || q < 0 || q >= g->nq) return nullptr;
  return g->file_spans.data() + q * g->lines * g->R * 2;
}

int gpjson_gathered_timing(const gpjson_gathered* g, double* query_ms, double* gather_ms, double* copy_ms) {
  if (!g) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  if (query_ms) *query_ms = g->query_ms;
  if (gather_ms) *gather_ms = g->gather_ms;
  if (copy_ms) *copy_ms = g->copy_ms;
  return GPJSON_OK;
}

void gpjson_gathered_free(gpjson_gathered* g) { delete g; }

}  // extern "C"

namespace gj {
// Levels and result slots a set of IRs needs: DOWN nesting depth and STORE count.
void ir_shape(const uint8_t* const* irs, const int64_t* ir_lens, int nq, int* L_out, int* R_out) {
  int L = 0, R = 1;
  for (int q = 0; q < nq; q++) {
    // levels / results are recovered from the IR: count DOWN depth and STOREs
    int depth = 0, maxd = 0, stores = 0;
    for (int64_t pc = 0; pc < ir_lens[q];) {
      uint8_t op = irs[q][pc++];
      if (op == kMoveDown) { depth++; maxd = std::max(maxd, depth); }
      else if (op == kMoveUp) depth--;
      else if (op == kStoreResult) stores++;
      else if (op == kMoveToKey || op == kExpressionStringEquals || op == kMoveToIndex) {
        uint32_t v = 0;
        int sh = 0;
        while (pc < ir_lens[q]) {
          uint8_t b = irs[q][pc++];
          v |= (uint32_t)(b & 0x7F) << sh;
          if (!(b & 0x80)) break;
          sh += 7;
        }
        if (op != kMoveToIndex) pc += v;
      }
    }
    L = std::max(L, maxd);
    R = std::max(R, stores);
  }
  *L_out = std::min(L, GPJSON_MAX_LEVELS);
  *R_out = R;
}
}  // namespace gj

struct gpjson_plan {
  gpjson_ctx* ctx = nullptr;
  gpjson_index* idx = nullptr;
  int nq = 0, R = 1;
  int64_t* spans = nullptr;
  gj::PreparedQueries pq;
  gj::DevBuf table;    // this plan's counter + job table + IR bytes
  gj::IndexTune itune;  // index tile shape the prepared launch was sized for
  std::vector<uint8_t> irs_flat;
  std::vector<int64_t> ir_lens;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // index kernel, query kernel
  bool timed_pending = false;
  // one member-chain query: its lines inside index tiles are evaluated by the
  // index kernel; list holds [count (64 B)][lines left to the interpreter]
  gj::FuseSpec fuse;
  int64_t* list = nullptr;
};

extern "C" {

int gpjson_plan_create(gpjson_ctx* ctx, gpjson_buf* buf, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
                       gpjson_plan** out) {
  if (!ctx || !buf || !irs || !ir_lens || !out || nq < 1 || nq > GPJSON_MAX_QUERIES)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  int L = 0, R = 1;
  ir_shape(irs, ir_lens, nq, &L, &R);
  std::unique_ptr<gpjson_plan> p(new gpjson_plan());
  p->ctx = ctx;
  p->nq = nq;
  p->R = R;
  int rc = gpjson_index_build(ctx, buf, L, 0, &p->idx);
  if (rc) return rc;
  rc = dev_alloc((void**)&p->spans, (size_t)nq * (size_t)std::max<int64_t>(p->idx->nl_cap, 1) * R * 2 * 8);
  shapes_for(ctx, buf->n, L, 0, &p->itune, nullptr);  // the shape the index was just built with
  if (rc == GPJSON_OK) rc = prepare_queries(ctx, p->idx, irs, ir_lens, nq, R, p->spans, &p->table, &p->pq);
  if (rc == GPJSON_OK && nq == 1 && ctx->qtune.fuse &&
      fuse_spec_from_ir(irs[0], ir_lens[0], &p->fuse)) {
    // lines the interpreter may get: at most every line (and the empty file's one)
    p->fuse.list_cap = p->idx->nl_cap + 1;
    rc = dev_alloc((void**)&p->list, 64 + (size_t)p->fuse.list_cap * 8);
    if (rc == GPJSON_OK) {
      p->fuse.count = (unsigned long long*)p->list;
      p->fuse.list = p->list + 8;
      p->fuse.spans = p->spans;
      QueryLaunch& ql = p->pq.launch;
      ql.line_list = p->fuse.list;
      ql.list_count = p->fuse.count;
      ql.list_cap = p->fuse.list_cap;
      ql.long_items = nullptr;  // the listed lines are few: no second long-line launch
      ql.long_count = nullptr;
    }
  } else {
    p->fuse.depth = 0;
  }
  for (int k = 0; k < 4 && rc == GPJSON_OK; k++)
    if (hipEventCreate(&p->ev[k]) != hipSuccess) rc = fail(GPJSON_ERR_HIP, "hipEventCreate");
  if (rc) {
    gpjson_plan_free(p.release());
    return rc;
  }
  *out = p.release();
  return GPJSON_OK;
}

int gpjson_plan_run(gpjson_plan* p, int timed) {
  if (!p) return fail(GPJSON_ERR_INVALID_ARG, "plan is NULL");
  gpjson_ctx* ctx = p->ctx;
  GJ_HIP(hipSetDevice(ctx->device));
  // the plan keeps the tile shape it was prepared with, whatever the context's
  // shapes or tuning cache say now: a prepared plan runs the same launches
  // every time
  int rc = enqueue_index(p->idx, timed ? p->ev[0] : nullptr, timed ? p->ev[1] : nullptr, &p->itune,
                         p->fuse.depth > 0 ? &p->fuse : nullptr);
  if (rc) return rc;
  rc = launch_prepared(ctx, p->pq, timed ? p->ev[2] : nullptr, timed ? p->ev[3] : nullptr);
  if (rc) return rc;
  if (timed) p->timed_pending = true;
  return GPJSON_OK;
}

int gpjson_plan_timing(gpjson_plan* p, double* index_ms, double* query_ms) {
  if (!p) return fail(GPJSON_ERR_INVALID_ARG, "plan is NULL");
  if (!p->timed_pending) return fail(GPJSON_ERR_INVALID_ARG, "no timed run");
  GJ_HIP(hipEventSynchronize(p->ev[3]));
  float a = 0, b = 0;
  GJ_HIP(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
  GJ_HIP(hipEventElapsedTime(&b, p->ev[2], p->ev[3]));
  if (index_ms) *index_ms = a;
  if (query_ms) *query_ms = b;
  return GPJSON_OK;
}

int gpjson_plan_fused(gpjson_plan* p, int* fused, int64_t* interpreted_lines) {
  if (!p) return fail(GPJSON_ERR_INVALID_ARG, "plan is NULL");
  if (fused) *fused = p->fuse.depth > 0 ? 1 : 0;
  if (interpreted_lines) {
    *interpreted_lines = -1;
    if (p->fuse.depth > 0) {
      GJ_HIP(hipSetDevice(p->ctx->device));
      GJ_HIP(hipStreamSynchronize(p->ctx->stream));
      unsigned long long c = 0;
      GJ_HIP(hipMemcpy(&c, p->fuse.count, 8, hipMemcpyDeviceToHost));
      *interpreted_lines = (int64_t)c;
    }
  }
  return GPJSON_OK;
}

int64_t gpjson_plan_lines(const gpjson_plan* p) { return p ? p->idx->lines : -1; }
int gpjson_plan_results_per_line(const gpjson_plan* p) { return p ? p->R : -1; }
const gpjson_index* gpjson_plan_index(const gpjson_plan* p) { return p ? p->idx : nullptr; }

int gpjson_plan_spans(gpjson_plan* p, int64_t* spans, int64_t cap_elems) {
  if (!p || !spans) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  GJ_HIP(hipSetDevice(p->ctx->device));
  GJ_HIP(hipStreamSynchronize(p->ctx->stream));
  // the line count of the last run (read back; it can only differ from the
  // first build if the buffer changed)
  int64_t lines = 0;
  GJ_HIP(hipMemcpy(&lines, p->idx->lines_dev, 8, hipMemcpyDeviceToHost));
  if (lines > p->idx->nl_cap) return fail(GPJSON_ERR_CAPACITY, "input changed: more lines than the plan holds");
  p->idx->lines = lines;
  const int64_t need = (int64_t)p->nq * lines * p->R * 2;
  if (cap_elems < need) return fail(GPJSON_ERR_CAPACITY, "span buffer too small");
  GJ_HIP(hipMemcpy(spans, p->spans, (size_t)need * 8, hipMemcpyDeviceToHost));
  return GPJSON_OK;
}

void gpjson_plan_free(gpjson_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  for (auto& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  if (p->spans) (void)hipFree(p->spans);
  if (p->list) (void)hipFree(p->list);
  p->table.release();
  if (p->idx) gpjson_index_free(p->idx);
  delete p;
}

int gpjson_time_plain(gpjson_ctx* ctx, gpjson_buf* buf, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
                      int iterations, double* ms_per_pass, double* index_ms, double* query_ms) {
  if (!ctx || !buf || !irs || !ir_lens || nq < 1 || iterations < 1)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  gpjson_plan* p = nullptr;
  int rc = gpjson_plan_create(ctx, buf, irs, ir_lens, nq, &p);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(GPJSON_ERR_HIP, "event");
  if (rc == GPJSON_OK) rc = gpjson_plan_run(p, 0);  // warm-up
  if (rc == GPJSON_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GPJSON_ERR_HIP, "sync");
  if (rc == GPJSON_OK) (void)hipEventRecord(e0, ctx->stream);
  for (int it = 0; it < iterations && rc == GPJSON_OK; it++) rc = gpjson_plan_run(p, it == iterations - 1);
  if (rc == GPJSON_OK) (void)hipEventRecord(e1, ctx->stream);
  if (rc == GPJSON_OK) {
    hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
  }
  if (rc == GPJSON_OK) {
    float total = 0;
    (void)hipEventElapsedTime(&total, e0, e1);
    if (ms_per_pass) *ms_per_pass = total / iterations;
    rc = gpjson_plan_timing(p, index_ms, query_ms);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  gpjson_plan_free(p);
  return rc;
}

}  // extern "C"
