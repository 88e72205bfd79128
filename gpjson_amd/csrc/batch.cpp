// batch.cpp -- batched execution: many files of uneven size indexed by one
// persistent index launch and queried by one interpreter launch.
//
// The reference handles one file per query() call (J/functions/QueryFunction.java:
// 48-265, J/ = gpjson/src/main/java/com/koenv/gpjson/); batching has no
// counterpart there.  Every file of a batch gets its own index arrays (sized
// for its own depth L, the largest of its queries) inside shared arenas.
//
// One run is:
//   reset kernel      lookback states + tile ticket
//   index kernel      persistent grid; tiles of all files are claimed from one
//                     atomic ticket (the device work queue of tile ranges); a
//                     tile finds its file by binary search over the files'
//                     first tile ids, and the lookback stops at the file's first
//                     tile, which starts from the initial carry.  A tile holding
//                     a newline speculates its carry-in (every '\n' of NDJSON
//                     lies outside strings at depth -1) and writes its newline
//                     offsets to its own scratch row
//   bookkeeping       four small kernels over all tiles: every carry-in checked
//                     against its predecessor's carry-out, per-file line counts,
//                     newline offsets to their slots; a wrong speculation (input
//                     that is not NDJSON) raises a flag that lets the same
//                     reset + index + bookkeeping run again without speculation
//                     (those launches return at once when the flag is clear)
//   batch plan kernel per-file line counts -> per-(file, query) work ranges
//   interpreter       one launch over every (file, query, line) item
// After the first (uncaptured) run has sized the arrays, the sequence is
// captured once into a hipGraph and replayed; nothing is allocated per run.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "runtime.h"

using namespace gj;

struct gpjson_batch {
  gpjson_ctx* ctx = nullptr;
  int nfiles = 0;
  int njobs = 0;
  std::vector<gpjson_buf*> bufs;
  std::vector<FileJob> files;       // host copy of the file table
  std::vector<QueryJob> qjobs;      // host copy of the job table (lines/item0 filled on device)
  std::vector<int> job_file, job_q, job_R;
  std::vector<int64_t> job_span_off;  // element offset of each job's spans in the span arena
  std::vector<int> file_first_job;
  // Level classes: with split_levels, the files whose level count needs a
  // smaller kernel variant than the deepest file's come first in tile order
  // and get an index launch of their own (tiles [0, split_tile)); the rest
  // take a second launch.  Files never span the two launches, so neither
  // launch waits on the other.  slot_file maps the device file table's order
  // (tile order) to the caller's file index.
  bool split_levels = false;
  int64_t split_tile = 0;
  int nlv_first = 2, nlv_all = 2;
  std::vector<int> slot_file, file_slot;
  int64_t total_tiles = 0;
  int64_t total_bytes = 0;
  int max_levels = 0;
  int64_t ir_total = 0;  // bytes of the distinct programs, contiguous in dev_ir
  std::vector<uint8_t> ir_host;    // the same bytes on the host (a new layout uploads them again)
  std::vector<int64_t> ir_lens_host;
  // device memory
  DevBuf arena_rec, arena_lb, arena_nl, arena_spans, dev_files, dev_qjobs, dev_misc, dev_ir, states;
  DevBuf arena_long;      // [2 counters, 64 B][long-line work items]
  DevBuf nl_scratch;      // per tile: a slot per word for its newline offsets
  DevBuf bk;              // bookkeeping: [flags, 128 B][block sums][per-file newline base]
  int64_t nl_scratch_cap = 0;
  int64_t long_cap = 0;
  int64_t* total_dev = nullptr;
  int64_t* lines_dev = nullptr;
  int* overflow_dev = nullptr;
  int* job_file_dev = nullptr;
  unsigned int* ticket = nullptr;
  unsigned long long* counter = nullptr;
  IndexTune itune;
  QueryTune qtune;
  int index_grid = 0;
  int query_grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool timed_graph = false;
  std::vector<int64_t> lines;  // per file, after a run
};

// dev_misc: line count total (0), overflow flag (8), index tickets
// (kMiscTicket), work counters (kMiscCounters), then per-file line counts and
// per-job file slots
static constexpr size_t kMiscTicket = 256;
static constexpr size_t kMiscCounters = kMiscTicket + kTicketBytes;
static constexpr size_t kMiscCtlBytes = (kMiscCounters + kQueryCounterWords * 8 + 4095) / 4096 * 4096;

static int levels_of(const uint8_t* ir, int64_t len, int* R) {
  int depth = 0, maxd = 0, stores = 0;
  for (int64_t pc = 0; pc < len;) {
    uint8_t op = ir[pc++];
    if (op == kMoveDown) { depth++; maxd = std::max(maxd, depth); }
    else if (op == kMoveUp) depth--;
    else if (op == kStoreResult) stores++;
    else if (op == kMoveToKey || op == kExpressionStringEquals || op == kMoveToIndex) {
      uint32_t v = 0;
      int sh = 0;
      while (pc < len) {
        uint8_t b = ir[pc++];
        v |= (uint32_t)(b & 0x7F) << sh;
        if (!(b & 0x80)) break;
        sh += 7;
      }
      if (op != kMoveToIndex) pc += v;
    }
  }
  *R = std::max(stores, 1);
  return maxd;
}

// Lays out the arenas for the current capacities and shapes and uploads the
// tables.
static int layout(gpjson_batch* b, const std::vector<int64_t>& nl_caps) {
  std::vector<const uint8_t*> irv(b->njobs);
  for (int64_t j = 0, off = 0; j < b->njobs; off += b->ir_lens_host[j], j++) irv[j] = b->ir_host.data() + off;
  const uint8_t* const* irs = irv.data();
  const int64_t* ir_lens = b->ir_lens_host.data();
  gpjson_ctx* ctx = b->ctx;
  const int64_t tile_bytes = index_tile_bytes(b->itune.block, b->itune.wpt);
  int64_t s_words = 0, lb_words = 0, nl_elems = 0, span_elems = 0, tiles = 0;
  std::vector<int64_t> s_off(b->nfiles), lb_off(b->nfiles), nl_off(b->nfiles);
  // tile order: the smaller level class first when the launch is split
  b->nlv_all = nlv_for_levels(b->max_levels);
  b->nlv_first = 2;
  b->slot_file.clear();
  for (int pass = 0; pass < 2; pass++)
    for (int f = 0; f < b->nfiles; f++) {
      const bool first = b->split_levels && nlv_for_levels(b->files[f].L) < b->nlv_all;
      if (first == (pass == 0)) b->slot_file.push_back(f);
      if (first && pass == 0) b->nlv_first = std::max(b->nlv_first, nlv_for_levels(b->files[f].L));
    }
  b->file_slot.assign(b->nfiles, 0);
  for (int k = 0; k < b->nfiles; k++) b->file_slot[b->slot_file[k]] = k;
  b->split_tile = 0;
  for (int k = 0; k < b->nfiles; k++) {
    const int f = b->slot_file[k];
    FileJob& J = b->files[f];
    const int64_t n = b->bufs[f]->n;
    J.N = n;
    J.W = (n + 63) / 64;
    J.Wp = (J.W + 7) / 8 * 8;
    J.nl_cap = nl_caps[f];
    J.ntiles = (n + tile_bytes - 1) / tile_bytes;
    J.tile0 = tiles;
    tiles += J.ntiles;
    if (b->split_levels && nlv_for_levels(J.L) < b->nlv_all) b->split_tile = tiles;
    s_off[f] = s_words;
    s_words += (J.W + 7) / 8 * 8;
    lb_off[f] = lb_words;
    lb_words += leveled_alloc_words(J.Wp, J.L);
    nl_off[f] = nl_elems;
    nl_elems += (J.nl_cap + 7) / 8 * 8;
  }
  b->job_span_off.assign(b->njobs, 0);
  for (int j = 0; j < b->njobs; j++) {
    b->job_span_off[j] = span_elems;
    span_elems += b->files[b->job_file[j]].nl_cap * b->job_R[j] * 2;
  }
  b->total_tiles = tiles;
  int rc = b->arena_rec.ensure((size_t)std::max<int64_t>(s_words, 1) * kRecWords * 8);
  if (!rc) rc = b->arena_lb.ensure((size_t)std::max<int64_t>(lb_words, 1) * 8);
  if (!rc) rc = b->arena_nl.ensure((size_t)std::max<int64_t>(nl_elems, 1) * 8);
  if (!rc) rc = b->arena_spans.ensure((size_t)std::max<int64_t>(span_elems, 1) * 8);
  b->long_cap = 0;
  for (int j = 0; j < b->njobs; j++) b->long_cap += b->bufs[b->job_file[j]]->n / kLongLineBytes + 1;
  if (!rc) rc = b->arena_long.ensure(64 + (size_t)b->long_cap * 8);
  if (!rc) rc = b->states.ensure((size_t)std::max<int64_t>(tiles, 1) * sizeof(TileState));
  b->nl_scratch_cap = tile_bytes / 64;
  if (!rc) rc = b->nl_scratch.ensure((size_t)std::max<int64_t>(tiles, 1) * (size_t)b->nl_scratch_cap * 8);
  if (!rc) rc = b->bk.ensure(128 + (size_t)(tiles / 256 + 2) * 8 + (size_t)b->nfiles * 8);
  if (!rc) rc = b->dev_files.ensure(sizeof(FileJob) * (size_t)b->nfiles);
  if (!rc) rc = b->dev_qjobs.ensure(sizeof(QueryJob) * (size_t)b->njobs);
  if (!rc) rc = b->dev_misc.ensure(kMiscCtlBytes + (size_t)b->nfiles * 8 + (size_t)b->njobs * 4);
  if (rc) return rc;
  for (int f = 0; f < b->nfiles; f++) {
    FileJob& J = b->files[f];
    J.F = b->bufs[f]->d;
    J.rec = (uint64_t*)b->arena_rec.p + s_off[f] * kRecWords;
    J.LB = (uint64_t*)b->arena_lb.p + lb_off[f];
    J.LS = summaries_of(J.LB, J.Wp, J.L);
    J.Ws = summary_stride(J.Wp);
    J.NL = (int64_t*)b->arena_nl.p + nl_off[f];
    J.E = nullptr;
    J.Q = nullptr;
  }
  // IR bytes, one copy per distinct program (files sharing a query set share
  // its programs), so a batch's programs usually fit the interpreter's LDS
  std::map<std::string, int64_t> ir_at;
  std::vector<int64_t> ir_off(b->njobs);
  int64_t ir_total = 0;
  for (int j = 0; j < b->njobs; j++) {
    const auto ins = ir_at.emplace(std::string((const char*)irs[j], (size_t)ir_lens[j]), ir_total);
    if (ins.second) ir_total += ir_lens[j];
    ir_off[j] = ins.first->second;
  }
  b->ir_total = ir_total;
  rc = b->dev_ir.ensure((size_t)ir_total + 16);
  if (rc) return rc;
  std::vector<uint8_t> flat((size_t)ir_total + 16);
  for (const auto& kv : ir_at) std::memcpy(flat.data() + kv.second, kv.first.data(), kv.first.size());
  for (int j = 0; j < b->njobs; j++) {
    QueryJob& q = b->qjobs[j];
    const FileJob& J = b->files[b->job_file[j]];
    q.F = J.F;
    q.N = J.N;
    q.NL = J.NL;
    q.lines = 0;
    q.rec = J.rec;
    q.LB = J.LB;
    q.LS = J.LS;
    q.Ws = J.Ws;
    q.Wp = J.Wp;
    q.L = J.L;
    q.R = b->job_R[j];
    q.ir = (const uint8_t*)b->dev_ir.p + ir_off[j];
    q.ir_len = ir_lens[j];
    q.out = (int64_t*)b->arena_spans.p + b->job_span_off[j];
    q.item0 = 0;
    q.qf = b->job_q[j];
    q.nqf = b->file_first_job[b->job_file[j] + 1] - b->file_first_job[b->job_file[j]];
  }
  uint8_t* misc = (uint8_t*)b->dev_misc.p;
  b->total_dev = (int64_t*)misc;
  b->overflow_dev = (int*)(misc + 8);
  b->ticket = (unsigned int*)(misc + kMiscTicket);
  b->counter = (unsigned long long*)(misc + kMiscCounters);
  b->lines_dev = (int64_t*)(misc + kMiscCtlBytes);
  b->job_file_dev = (int*)(misc + kMiscCtlBytes + (size_t)b->nfiles * 8);
  GJ_HIP(hipMemcpy(b->dev_ir.p, flat.data(), flat.size(), hipMemcpyHostToDevice));
  // the device tables in tile order: file table by slot, each job's file slot
  std::vector<FileJob> table(b->nfiles);
  for (int k = 0; k < b->nfiles; k++) table[k] = b->files[b->slot_file[k]];
  std::vector<int> job_slot(b->njobs);
  for (int j = 0; j < b->njobs; j++) job_slot[j] = b->file_slot[b->job_file[j]];
  GJ_HIP(hipMemcpy(b->dev_files.p, table.data(), sizeof(FileJob) * b->nfiles, hipMemcpyHostToDevice));
  GJ_HIP(hipMemcpy(b->dev_qjobs.p, b->qjobs.data(), sizeof(QueryJob) * b->njobs, hipMemcpyHostToDevice));
  GJ_HIP(hipMemcpy(b->job_file_dev, job_slot.data(), sizeof(int) * b->njobs, hipMemcpyHostToDevice));
  GJ_HIP(hipMemset(b->dev_misc.p, 0, kMiscCtlBytes));
  const int bpc = 2048 / b->itune.block;  // resident blocks per CU the grid aims for
  b->index_grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)ctx->num_cus * bpc));
  b->query_grid = std::max(1, ctx->num_cus * b->qtune.waves_per_cu / (b->qtune.block / 64));
  return GPJSON_OK;
}

// Enqueues one batched run on the context stream (capturable: kernels only).
static int enqueue_batch(gpjson_batch* b, hipEvent_t e0, hipEvent_t e1, hipEvent_t e2) {
  gpjson_ctx* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  IndexLaunch L{};
  L.speculate = 1;
  L.jobs = (const FileJob*)b->dev_files.p;
  L.njobs = b->nfiles;
  L.total_tiles = b->total_tiles;
  L.states = (TileState*)b->states.p;
  L.ticket = b->ticket;
  L.nl_scratch = (int64_t*)b->nl_scratch.p;
  L.nl_scratch_cap = b->nl_scratch_cap;
  // bookkeeping flags: [0] a speculated carry-in was wrong (it gates the second
  // pass), [8] a tile overflowed its scratch row, [16] the second pass's check
  // (not read)
  int* flags = (int*)b->bk.p;
  int64_t* block_sums = (int64_t*)((uint8_t*)b->bk.p + 128);
  int64_t* file_base = block_sums + b->total_tiles / 256 + 2;
  const int maxL = b->max_levels;
  const int nlv = nlv_for_levels(maxL);
  // the index launches of one pass: one over every tile, or one per level
  // class (the ticket restarts at the second class's first tile)
  const bool split = b->split_tile > 0 && b->split_tile < b->total_tiles;
  auto index_pass = [&](const IndexLaunch& P) -> int {
    if (b->total_tiles == 0) return GPJSON_OK;
    if (!split) {
      GJ_HIP(index_launch(P, b->itune.block, b->itune.wpt, nlv, false, b->index_grid, s));
      return GPJSON_OK;
    }
    IndexLaunch A = P;
    A.total_tiles = b->split_tile;
    GJ_HIP(index_launch(A, b->itune.block, b->itune.wpt, b->nlv_first, false,
                        (int)std::min<int64_t>(b->index_grid, b->split_tile), s));
    IndexLaunch B = P;
    B.tile_begin = b->split_tile;
    GJ_HIP(zero_u64((unsigned long long*)P.ticket, kTicketBytes / 8, s));
    GJ_HIP(index_launch(B, b->itune.block, b->itune.wpt, nlv, false,
                        (int)std::min<int64_t>(b->index_grid, b->total_tiles - b->split_tile), s));
    return GPJSON_OK;
  };
  GJ_HIP(index_reset(L.states, b->total_tiles, L.ticket, s));
  GJ_HIP(zero_u64((unsigned long long*)flags, 16, s));
  GJ_HIP(zero_u64(b->counter, kQueryCounterWords, s));
  if (e0) GJ_HIP(hipEventRecord(e0, s));
  int rc = index_pass(L);
  if (rc) return rc;
  GJ_HIP(index_finish_batch(L, block_sums, file_base, b->lines_dev, flags, flags + 8, s));
  IndexLaunch L2 = L;
  L2.speculate = 0;
  L2.run_if = flags;
  GJ_HIP(index_reset(L.states, b->total_tiles, L.ticket, s, flags));
  rc = index_pass(L2);
  if (rc) return rc;
  GJ_HIP(index_finish_batch(L2, block_sums, file_base, b->lines_dev, flags + 16, flags + 8, s));
  if (e1) GJ_HIP(hipEventRecord(e1, s));
  GJ_HIP(batch_plan((const FileJob*)b->dev_files.p, b->nfiles, (QueryJob*)b->dev_qjobs.p, b->job_file_dev,
                    b->njobs, b->total_dev, b->lines_dev, b->overflow_dev, s));
  QueryLaunch Q{};
  Q.jobs = (const QueryJob*)b->dev_qjobs.p;
  Q.njobs = b->njobs;
  Q.total_items = 0;
  Q.lines_dev = nullptr;
  Q.total_dev = b->total_dev;
  Q.counter = b->counter;
  Q.lines_per_wave = b->qtune.lines_per_wave;
  Q.ir_base = (const uint8_t*)b->dev_ir.p;
  Q.ir_bytes = b->ir_total;
  Q.qblock = b->qtune.qblock;
  const bool long_lines = b->qtune.mode == kLineLongMode && b->qtune.group != 64;
  if (long_lines) {
    Q.long_count = (unsigned long long*)b->arena_long.p;
    Q.long_items = (int64_t*)((uint8_t*)b->arena_long.p + 64);
    Q.long_cap = b->long_cap;
    Q.long_min = kLongLineBytes;
    GJ_HIP(zero_u64(Q.long_count, 2, s));
  }
  GJ_HIP(query_launch(Q, b->qtune.block, b->qtune.group, maxL, b->query_grid, s, ctx->num_cus * 2));
  if (e2) GJ_HIP(hipEventRecord(e2, s));
  return GPJSON_OK;
}

static int read_back_lines(gpjson_batch* b, bool* overflow) {
  int ov = 0;
  GJ_HIP(hipStreamSynchronize(b->ctx->stream));
  GJ_HIP(hipMemcpy(&ov, b->overflow_dev, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int64_t> by_slot(b->nfiles, 0);
  GJ_HIP(hipMemcpy(by_slot.data(), b->lines_dev, sizeof(int64_t) * b->nfiles, hipMemcpyDeviceToHost));
  b->lines.assign(b->nfiles, 0);
  for (int k = 0; k < b->nfiles; k++) b->lines[b->slot_file[k]] = by_slot[k];
  *overflow = ov != 0;
  return GPJSON_OK;
}

static void drop_graph(gpjson_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

extern "C" {

int gpjson_batch_create(gpjson_ctx* ctx, gpjson_buf* const* bufs, int nfiles, const uint8_t* const* irs,
                        const int64_t* ir_lens, const int* nq_per_file, gpjson_batch** out) {
  if (!ctx || !bufs || !irs || !ir_lens || !nq_per_file || !out || nfiles < 1)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  GJ_HIP(hipSetDevice(ctx->device));
  // a batch that fails half-built releases what it holds
  std::unique_ptr<gpjson_batch, void (*)(gpjson_batch*)> b(new gpjson_batch(), gpjson_batch_free);
  b->ctx = ctx;
  b->nfiles = nfiles;
  b->bufs.assign(bufs, bufs + nfiles);
  b->files.assign(nfiles, FileJob{});
  int j = 0;
  for (int f = 0; f < nfiles; f++) {
    if (!bufs[f]) return fail(GPJSON_ERR_INVALID_ARG, "NULL buffer");
    if (nq_per_file[f] < 1 || nq_per_file[f] > GPJSON_MAX_QUERIES)
      return fail(GPJSON_ERR_INVALID_ARG, "each file needs 1..255 queries");
    b->file_first_job.push_back(j);
    int L = 0;
    for (int q = 0; q < nq_per_file[f]; q++, j++) {
      if (!irs[j] || ir_lens[j] < 1) return fail(GPJSON_ERR_INVALID_ARG, "empty query program");
      int R = 1;
      int l = levels_of(irs[j], ir_lens[j], &R);
      L = std::max(L, l);
      b->job_file.push_back(f);
      b->job_q.push_back(q);
      b->job_R.push_back(R);
    }
    if (L > GPJSON_MAX_LEVELS) return fail(GPJSON_ERR_CAPACITY, "JSON path deeper than 64 levels");
    b->files[f].L = L;
  }
  b->file_first_job.push_back(j);
  b->njobs = j;
  b->qjobs.assign(j, QueryJob{});
  for (int k = 0; k < j; k++) {
    b->ir_host.insert(b->ir_host.end(), irs[k], irs[k] + ir_lens[k]);
    b->ir_lens_host.push_back(ir_lens[k]);
  }
  for (int f = 0; f < nfiles; f++) {
    b->total_bytes += bufs[f]->n;
    b->max_levels = std::max(b->max_levels, b->files[f].L);
  }
  // launch shapes: the tuning cache's for a batch of this size and depth, else the context's
  shapes_for(ctx, b->total_bytes, b->max_levels, 1, &b->itune, &b->qtune);
  for (int k = 0; k < 4; k++) GJ_HIP(hipEventCreate(&b->ev[k]));
  std::vector<int64_t> caps(nfiles);
  for (int f = 0; f < nfiles; f++) caps[f] = bufs[f]->n / 256 + 4096;
  int rc = layout(b.get(), caps);
  if (rc) return rc;
  // first run sizes the newline arrays exactly (grows them if a file overflowed)
  rc = enqueue_batch(b.get(), nullptr, nullptr, nullptr);
  bool ov = false;
  if (!rc) rc = read_back_lines(b.get(), &ov);
  if (!rc && ov) {
    // the bookkeeping counts every file's lines even when its newline array
    // overflowed: lay out again with exact capacities
    for (int f = 0; f < nfiles; f++) caps[f] = std::max<int64_t>(b->lines[f], caps[f]);
    rc = layout(b.get(), caps);
    if (!rc) rc = enqueue_batch(b.get(), nullptr, nullptr, nullptr);
    if (!rc) rc = read_back_lines(b.get(), &ov);
    if (!rc && ov) rc = fail(GPJSON_ERR_CAPACITY, "newline capacity still exceeded");
  }
  if (rc) return rc;
  *out = b.release();
  return GPJSON_OK;
}

int gpjson_batch_run(gpjson_batch* b, int use_graph, int timed) {
  if (!b) return fail(GPJSON_ERR_INVALID_ARG, "batch is NULL");
  gpjson_ctx* ctx = b->ctx;
  GJ_HIP(hipSetDevice(ctx->device));
  if (!use_graph) {
    b->timed_graph = false;
    return enqueue_batch(b, timed ? b->ev[0] : nullptr, timed ? b->ev[1] : nullptr, timed ? b->ev[2] : nullptr);
  }
  if (!b->exec) {
    GJ_HIP(hipStreamSynchronize(ctx->stream));
    GJ_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch(b, nullptr, nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  if (timed) GJ_HIP(hipEventRecord(b->ev[0], ctx->stream));
  GJ_HIP(hipGraphLaunch(b->exec, ctx->stream));
  if (timed) {
    GJ_HIP(hipEventRecord(b->ev[2], ctx->stream));
    b->timed_graph = true;
  }
  return GPJSON_OK;
}

int gpjson_batch_timing(gpjson_batch* b, double* index_ms, double* total_ms) {
  if (!b) return fail(GPJSON_ERR_INVALID_ARG, "batch is NULL");
  GJ_HIP(hipEventSynchronize(b->ev[2]));
  float a = -1, t = 0;
  if (!b->timed_graph) GJ_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
  GJ_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[2]));
  if (index_ms) *index_ms = a;
  if (total_ms) *total_ms = t;
  return GPJSON_OK;
}

int gpjson_batch_lines(gpjson_batch* b, int file, int64_t* lines) {
  if (!b || !lines || file < 0 || file >= b->nfiles) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  bool ov = false;
  int rc = read_back_lines(b, &ov);
  if (rc) return rc;
  if (ov) return fail(GPJSON_ERR_CAPACITY, "input changed: more lines than the batch holds");
  *lines = b->lines[file];
  return GPJSON_OK;
}

int gpjson_batch_spans(gpjson_batch* b, int file, int q, int64_t* spans, int64_t cap_elems) {
  if (!b || !spans || file < 0 || file >= b->nfiles) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  const int j = b->file_first_job[file] + q;
  if (q < 0 || j >= b->file_first_job[file + 1]) return fail(GPJSON_ERR_INVALID_ARG, "query out of range");
  int64_t lines = 0;
  int rc = gpjson_batch_lines(b, file, &lines);
  if (rc) return rc;
  const int64_t need = lines * b->job_R[j] * 2;
  if (cap_elems < need) return fail(GPJSON_ERR_CAPACITY, "span buffer too small");
  GJ_HIP(hipMemcpy(spans, (int64_t*)b->arena_spans.p + b->job_span_off[j], (size_t)need * 8,
                   hipMemcpyDeviceToHost));
  return GPJSON_OK;
}

int gpjson_batch_results_per_line(const gpjson_batch* b, int file, int q) {
  if (!b || file < 0 || file >= b->nfiles) return -1;
  const int j = b->file_first_job[file] + q;
  if (q < 0 || j >= b->file_first_job[file + 1]) return -1;
  return b->job_R[j];
}

// Times the batch's own runs (eager, with events: index + bookkeeping, then
// plan + interpreter) under every index tile shape with the current
// interpreter shape, then every interpreter shape with the fastest tile, and
// keeps the fastest pair: the arenas are laid out again for it, the captured
// graph (if any) is dropped so the next graph run captures the new launches,
// and the pair goes into the context's tuning cache under (bucket of the
// batch's total bytes, its deepest level, batch).  Each candidate is run once
// untimed and `iterations` times timed.
int gpjson_batch_autotune(gpjson_batch* b, int iterations, char* report, size_t report_cap) {
  if (!b || iterations < 1) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  gpjson_ctx* ctx = b->ctx;
  GJ_HIP(hipSetDevice(ctx->device));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(b);
  b->timed_graph = false;
  std::vector<int64_t> caps(b->nfiles);
  for (int f = 0; f < b->nfiles; f++) caps[f] = b->files[f].nl_cap;
  const IndexTune keep_i = b->itune;
  const QueryTune keep_q = b->qtune;
  // lays out for the current shapes and times them
  auto time_shapes = [&](double* index_ms, double* query_ms) -> int {
    int rc = layout(b, caps);
    if (rc) return rc;
    rc = enqueue_batch(b, nullptr, nullptr, nullptr);
    double si = 0, sq = 0;
    for (int it = 0; it < iterations && rc == GPJSON_OK; it++) {
      rc = enqueue_batch(b, b->ev[0], b->ev[1], b->ev[2]);
      if (rc) break;
      GJ_HIP(hipEventSynchronize(b->ev[2]));
      float a = 0, t = 0;
      GJ_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
      GJ_HIP(hipEventElapsedTime(&t, b->ev[1], b->ev[2]));
      si += a;
      sq += t;
    }
    *index_ms = si / iterations;
    *query_ms = sq / iterations;
    return rc;
  };
  std::string rep;
  char line[256];
  double best = 1e30;
  IndexTune best_i = keep_i;
  int rc = GPJSON_OK;
  for (const IndexTune& c : index_candidates()) {
    b->itune = c;
    double ims = 0, qms = 0;
    rc = time_shapes(&ims, &qms);
    if (rc) break;
    std::snprintf(line, sizeof(line), "batch index block=%d wpt=%d tile=%lld B: %.4f ms\n", c.block, c.wpt,
                  (long long)index_tile_bytes(c.block, c.wpt), ims);
    rep += line;
    if (ims < best) {
      best = ims;
      best_i = c;
    }
  }
  b->itune = best_i;
  // one index launch for all files, or one per level class (when some files
  // need a smaller kernel variant than the deepest file's)
  const bool keep_split = b->split_levels;
  bool best_split = keep_split;
  bool two_classes = false;
  for (int f = 0; f < b->nfiles; f++)
    two_classes |= nlv_for_levels(b->files[f].L) < nlv_for_levels(b->max_levels) && b->bufs[f]->n > 0;
  if (rc == GPJSON_OK && two_classes) {
    double t[2] = {0, 0};
    for (int split = 0; split < 2 && rc == GPJSON_OK; split++) {
      b->split_levels = split != 0;
      double qms = 0;
      rc = time_shapes(&t[split], &qms);
      if (rc) break;
      std::snprintf(line, sizeof(line), "batch index (best shape) %s: %.4f ms\n",
                    split ? "one launch per level class" : "one launch", t[split]);
      rep += line;
    }
    if (rc == GPJSON_OK) {
      best_split = t[1] < t[0];
      best = best_split ? t[1] : t[0];
    }
  }
  b->split_levels = best_split;
  double bestq = 1e30;
  QueryTune best_q = keep_q;
  if (rc == GPJSON_OK) {
    for (const QueryTune& c : query_candidates(keep_q)) {
      b->qtune = c;
      double ims = 0, qms = 0;
      rc = time_shapes(&ims, &qms);
      if (rc) break;
      std::snprintf(line, sizeof(line),
                    "batch query mode=%d block=%d group=%d waves_per_cu=%d lines_per_group=%d: %.4f ms\n", c.mode,
                    c.block, c.group, c.waves_per_cu, c.lines_per_wave, qms);
      rep += line;
      if (qms < bestq) {
        bestq = qms;
        best_q = c;
      }
    }
  }
  // the work order: the queries of a file taking turns over blocks of lines
  if (rc == GPJSON_OK) {
    for (int qb : {16, 64, 256, 1024, 4096}) {
      QueryTune c = best_q;
      c.qblock = qb;
      b->qtune = c;
      double ims = 0, qms = 0;
      rc = time_shapes(&ims, &qms);
      if (rc) break;
      std::snprintf(line, sizeof(line), "batch query (best shape) queries taking turns over blocks of %d lines: %.4f ms\n",
                    qb, qms);
      rep += line;
      if (qms < bestq) {
        bestq = qms;
        best_q = c;
      }
    }
  }
  if (rc) {
    // back to the shapes the batch had, laid out as before
    b->itune = keep_i;
    b->qtune = keep_q;
    b->split_levels = keep_split;
    (void)layout(b, caps);
    return rc;
  }
  b->qtune = best_q;
  rc = layout(b, caps);
  if (!rc) rc = enqueue_batch(b, nullptr, nullptr, nullptr);  // the spans of the chosen shapes
  if (rc) return rc;
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  const TuneKey key{size_bucket(b->total_bytes), b->max_levels, 1};
  ctx->tune_cache[key] = TuneEntry{best_i, best_q, best + bestq};
  std::snprintf(line, sizeof(line),
                "batch best index %dx%d%s (%.4f ms), query mode %d block %d group %d wpc %d lpg %d qblock %d "
                "(%.4f ms); cached for batches of [2^%d, 2^%d) bytes at %d levels\n",
                best_i.block, best_i.wpt, best_split ? " per level class" : "", best, best_q.mode, best_q.block,
                best_q.group, best_q.waves_per_cu, best_q.lines_per_wave, best_q.qblock, bestq, key.bucket,
                key.bucket + 1, b->max_levels);
  rep += line;
  if (report && report_cap) {
    const size_t n = std::min(report_cap - 1, rep.size());
    std::copy(rep.begin(), rep.begin() + n, report);
    report[n] = 0;
  }
  return GPJSON_OK;
}

int gpjson_batch_query_order(const gpjson_batch* b) { return b ? b->qtune.qblock : -1; }

int gpjson_batch_set_level_split(gpjson_batch* b, int enable) {
  if (!b) return fail(GPJSON_ERR_INVALID_ARG, "batch is NULL");
  GJ_HIP(hipSetDevice(b->ctx->device));
  GJ_HIP(hipStreamSynchronize(b->ctx->stream));
  drop_graph(b);
  b->timed_graph = false;
  b->split_levels = enable != 0;
  std::vector<int64_t> caps(b->nfiles);
  for (int f = 0; f < b->nfiles; f++) caps[f] = b->files[f].nl_cap;
  int rc = layout(b, caps);
  if (!rc) rc = enqueue_batch(b, nullptr, nullptr, nullptr);  // the spans of the new layout
  if (rc) return rc;
  GJ_HIP(hipStreamSynchronize(b->ctx->stream));
  return GPJSON_OK;
}

int gpjson_batch_get_level_split(const gpjson_batch* b) {
  if (!b) return -1;
  return b->split_tile > 0 && b->split_tile < b->total_tiles ? 1 : 0;
}

int gpjson_batch_get_config(const gpjson_batch* b, int* index_block, int* index_wpt, int* query_mode,
                            int* query_block, int* query_group, int* query_waves_per_cu, int* lines_per_wave) {
  if (!b) return fail(GPJSON_ERR_INVALID_ARG, "batch is NULL");
  if (index_block) *index_block = b->itune.block;
  if (index_wpt) *index_wpt = b->itune.wpt;
  if (query_mode) *query_mode = b->qtune.mode;
  if (query_block) *query_block = b->qtune.block;
  if (query_group) *query_group = b->qtune.group;
  if (query_waves_per_cu) *query_waves_per_cu = b->qtune.waves_per_cu;
  if (lines_per_wave) *lines_per_wave = b->qtune.lines_per_wave;
  return GPJSON_OK;
}

void gpjson_batch_free(gpjson_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  DevBuf* bufs[] = {&b->arena_rec, &b->arena_lb, &b->arena_nl, &b->arena_spans,
                    &b->dev_files, &b->dev_qjobs, &b->dev_misc, &b->dev_ir, &b->states, &b->arena_long,
                    &b->nl_scratch, &b->bk};
  for (DevBuf* d : bufs) d->release();
  delete b;
}

}  // extern "C"
