// autotune.cpp -- launch-shape search for the index and interpreter kernels,
// and the tuning cache that keeps its results.
//
// The reference hard-codes 8x1024 threads for the index kernels and 512x1024
// for find_value (J/stages/CombinedIndex.java:11-14, J/functions/
// QueryFunction.java:229; J/ = gpjson/src/main/java/com/koenv/gpjson/).  Here
// the tile shape (threads per block x 64-byte words per thread, i.e. the chunk
// of input one tile covers), the interpreter's block size, its persistent grid
// size and lines claimed per wavefront are measured on the caller's own input.
// The fastest shapes go into the context's tuning cache under (input size
// bucket, levels, plain or batch); every later plain pass, one-shot query and
// batch with the same key runs with them.  The cache can be saved to a text
// file and loaded into another context of the same architecture.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.h"

using namespace gj;

namespace gj {

const std::vector<IndexTune>& index_candidates() {
  static const std::vector<IndexTune> c = {{256, 1}, {256, 2}, {256, 3}, {256, 4}, {512, 1},
                                           {512, 2}, {128, 2}, {128, 3}, {128, 4}};
  return c;
}

// Interpreter shapes: line mode over lane groups, grid sizes and claim sizes,
// and long lines to a wavefront each.
std::vector<QueryTune> query_candidates(const QueryTune& base) {
  std::vector<QueryTune> c;
  auto add = [&](int mode, int block, int group, int wpc, int lpw) {
    QueryTune q = base;
    q.mode = mode;
    q.block = block;
    q.group = group;
    q.waves_per_cu = wpc;
    q.lines_per_wave = lpw;
    q.qblock = 0;  // query after query; the block orders are tried on the best shape
    c.push_back(q);
  };
  for (int group : {1, 4, 16, 64})
    for (int wpc : {16, 32})
      for (int lpw : {1, 4}) add(kLineMode, 256, group, wpc, lpw);
  for (int group : {1, 4})
    for (int wpc : {16, 32}) add(kLineLongMode, 256, group, wpc, group == 1 ? 4 : 1);
  return c;
}

int size_bucket(int64_t n) { return n > 1 ? 63 - __builtin_clzll((unsigned long long)n) : 0; }

void shapes_for(const gpjson_ctx* ctx, int64_t n, int L, int batch, IndexTune* it, QueryTune* qt) {
  IndexTune i = ctx->itune;
  QueryTune q = ctx->qtune;
  if (!ctx->tuning) {
    const auto e = ctx->tune_cache.find(TuneKey{size_bucket(n), L, batch});
    if (e != ctx->tune_cache.end()) {
      i = e->second.itune;
      q = e->second.qtune;
      q.fuse = ctx->qtune.fuse;
    }
  }
  if (it) *it = i;
  if (qt) *qt = q;
}

}  // namespace gj

namespace {

void copy_report(const std::string& rep, char* report, size_t report_cap) {
  if (report && report_cap) {
    const size_t n = std::min(report_cap - 1, rep.size());
    std::copy(rep.begin(), rep.begin() + n, report);
    report[n] = 0;
  }
}

// Restores the context's tuning state on every exit of an autotuner.
struct TuningScope {
  gpjson_ctx* ctx;
  IndexTune keep_i;
  QueryTune keep_q;
  explicit TuningScope(gpjson_ctx* c) : ctx(c), keep_i(c->itune), keep_q(c->qtune) { ctx->tuning = true; }
  ~TuningScope() {
    ctx->itune = keep_i;
    ctx->qtune = keep_q;
    ctx->tuning = false;
  }
};

}  // namespace

extern "C" {

int gpjson_get_config(gpjson_ctx* ctx, int* index_block, int* index_wpt, int* query_mode, int* query_block,
                      int* query_group, int* query_waves_per_cu, int* lines_per_wave) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  if (query_mode) *query_mode = ctx->qtune.mode;
  if (index_block) *index_block = ctx->itune.block;
  if (index_wpt) *index_wpt = ctx->itune.wpt;
  if (query_block) *query_block = ctx->qtune.block;
  if (query_group) *query_group = ctx->qtune.group;
  if (query_waves_per_cu) *query_waves_per_cu = ctx->qtune.waves_per_cu;
  if (lines_per_wave) *lines_per_wave = ctx->qtune.lines_per_wave;
  return GPJSON_OK;
}

int gpjson_set_config(gpjson_ctx* ctx, int index_block, int index_wpt, int query_mode, int query_block,
                      int query_group, int query_waves_per_cu, int lines_per_wave) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  if (!index_config_supported(index_block, index_wpt))
    return fail(GPJSON_ERR_INVALID_ARG, "unsupported index shape " + std::to_string(index_block) + "x" +
                                            std::to_string(index_wpt));
  if (!query_config_supported(query_block, query_group))
    return fail(GPJSON_ERR_INVALID_ARG, "query block must be 64, 128 or 256 and group 1, 4, 16 or 64");
  if (query_waves_per_cu < 1 || query_waves_per_cu > 64 || lines_per_wave < 1 || lines_per_wave > 1024)
    return fail(GPJSON_ERR_INVALID_ARG, "query shape out of range");
  if (query_mode != kLineMode && query_mode != kLineLongMode)
    return fail(GPJSON_ERR_INVALID_ARG, "query mode must be 0 (line) or 3 (line, long lines to a wavefront)");
  GJ_HIP(hipSetDevice(ctx->device));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->itune.block = index_block;
  ctx->itune.wpt = index_wpt;
  ctx->qtune.mode = query_mode;
  ctx->qtune.block = query_block;
  ctx->qtune.group = query_group;
  ctx->qtune.waves_per_cu = query_waves_per_cu;
  ctx->qtune.lines_per_wave = lines_per_wave;
  ctx->tune_cache.clear();  // shapes set by hand apply to every input
  return GPJSON_OK;
}

int gpjson_set_query_order(gpjson_ctx* ctx, int qblock) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  if (qblock < 0 || qblock > (1 << 20)) return fail(GPJSON_ERR_INVALID_ARG, "qblock must be in [0, 2^20]");
  GJ_HIP(hipSetDevice(ctx->device));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->qtune.qblock = qblock;
  ctx->tune_cache.clear();  // shapes set by hand apply to every input
  return GPJSON_OK;
}

int gpjson_get_query_order(gpjson_ctx* ctx, int* qblock) {
  if (!ctx || !qblock) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  *qblock = ctx->qtune.qblock;
  return GPJSON_OK;
}

int gpjson_set_fusion(gpjson_ctx* ctx, int enable) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  ctx->qtune.fuse = enable ? 1 : 0;
  return GPJSON_OK;
}

int gpjson_get_fusion(gpjson_ctx* ctx, int* enabled) {
  if (!ctx || !enabled) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  *enabled = ctx->qtune.fuse;
  return GPJSON_OK;
}

// Times every candidate shape on buf with the given queries (index kernel and
// interpreter separately, `iterations` timed runs each after one warm-up) and
// records the fastest in the tuning cache under (bucket of buf's size, the
// queries' depth, plain).  `report` (optional) receives one line per candidate.
int gpjson_autotune(gpjson_ctx* ctx, gpjson_buf* buf, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
                    int iterations, char* report, size_t report_cap) {
  if (!ctx || !buf || !irs || !ir_lens || nq < 1 || iterations < 1)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  GJ_HIP(hipSetDevice(ctx->device));
  int L = 0, R = 1;
  ir_shape(irs, ir_lens, nq, &L, &R);
  TuningScope scope(ctx);
  std::string rep;
  char line[256];
  // one plan per candidate (a plan caches its launch shape); each is timed
  // `iterations` times after a warm-up run
  auto time_plan = [&](double* index_ms, double* query_ms) -> int {
    gpjson_plan* p = nullptr;
    int rc = gpjson_plan_create(ctx, buf, irs, ir_lens, nq, &p);
    if (rc) return rc;
    double si = 0, sq = 0;
    rc = gpjson_plan_run(p, 0);
    for (int it = 0; it < iterations && rc == GPJSON_OK; it++) {
      double a = 0, b = 0;
      rc = gpjson_plan_run(p, 1);
      if (rc == GPJSON_OK) rc = gpjson_plan_timing(p, &a, &b);
      si += a;
      sq += b;
    }
    gpjson_plan_free(p);
    *index_ms = si / iterations;
    *query_ms = sq / iterations;
    return rc;
  };
  double best = 1e30;
  IndexTune best_i = ctx->itune;
  for (const IndexTune& c : index_candidates()) {
    ctx->itune = c;
    double ims = 0, qms = 0;
    const int rc = time_plan(&ims, &qms);
    if (rc) return rc;
    std::snprintf(line, sizeof(line), "index block=%d wpt=%d tile=%lld B: %.4f ms\n", c.block, c.wpt,
                  (long long)index_tile_bytes(c.block, c.wpt), ims);
    rep += line;
    if (ims < best) {
      best = ims;
      best_i = c;
    }
  }
  ctx->itune = best_i;
  double bestq = 1e30;
  QueryTune best_q = ctx->qtune;
  for (const QueryTune& c : query_candidates(ctx->qtune)) {
    ctx->qtune = c;
    double ims = 0, qms = 0;
    const int rc = time_plan(&ims, &qms);
    if (rc) return rc;
    std::snprintf(line, sizeof(line), "query mode=%d block=%d group=%d waves_per_cu=%d lines_per_group=%d: %.4f ms\n",
                  c.mode, c.block, c.group, c.waves_per_cu, c.lines_per_wave, qms);
    rep += line;
    if (qms < bestq) {
      bestq = qms;
      best_q = c;
    }
  }
  if (nq > 1) {
    // several queries: their work order (taking turns over blocks of lines)
    for (int qb : {16, 64, 256, 1024, 4096}) {
      QueryTune c = best_q;
      c.qblock = qb;
      ctx->qtune = c;
      double ims = 0, qms = 0;
      const int rc = time_plan(&ims, &qms);
      if (rc) return rc;
      std::snprintf(line, sizeof(line), "query (best shape) queries taking turns over blocks of %d lines: %.4f ms\n",
                    qb, qms);
      rep += line;
      if (qms < bestq) {
        bestq = qms;
        best_q = c;
      }
    }
  }
  const TuneKey key{size_bucket(buf->n), L, 0};
  ctx->tune_cache[key] = TuneEntry{best_i, best_q, best + bestq};
  std::snprintf(line, sizeof(line),
                "best index %dx%d (%.4f ms), query mode %d block %d group %d wpc %d lpg %d qblock %d (%.4f ms); "
                "cached for inputs of [2^%d, 2^%d) bytes at %d levels\n",
                best_i.block, best_i.wpt, best, best_q.mode, best_q.block, best_q.group, best_q.waves_per_cu,
                best_q.lines_per_wave, best_q.qblock, bestq, key.bucket, key.bucket + 1, L);
  rep += line;
  copy_report(rep, report, report_cap);
  return GPJSON_OK;
}

int gpjson_tune_cache_clear(gpjson_ctx* ctx) {
  if (!ctx) return fail(GPJSON_ERR_INVALID_ARG, "ctx is NULL");
  ctx->tune_cache.clear();
  return GPJSON_OK;
}

int64_t gpjson_tune_cache_size(const gpjson_ctx* ctx) { return ctx ? (int64_t)ctx->tune_cache.size() : -1; }

// Text format: a header line, then one line per entry:
//   arch bucket levels batch index_block index_wpt query_mode query_block
//   query_group waves_per_cu lines_per_wave qblock ms
int gpjson_tune_cache_save(const gpjson_ctx* ctx, const char* path) {
  if (!ctx || !path) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(GPJSON_ERR_IO, std::string("Failed to open file ") + path);
  std::fprintf(f, "gpjson-tune-cache 2\n");
  for (const auto& kv : ctx->tune_cache) {
    const TuneKey& k = kv.first;
    const TuneEntry& e = kv.second;
    std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %d %.6f\n", ctx->arch.c_str(), k.bucket, k.levels, k.batch,
                 e.itune.block, e.itune.wpt, e.qtune.mode, e.qtune.block, e.qtune.group, e.qtune.waves_per_cu,
                 e.qtune.lines_per_wave, e.qtune.qblock, e.ms);
  }
  const bool ok = std::fflush(f) == 0;
  std::fclose(f);
  return ok ? GPJSON_OK : fail(GPJSON_ERR_IO, std::string("Failed to write file ") + path);
}

// Loads the entries of ctx's architecture (others are skipped) over the cache's
// own; *loaded (optional) gets their number.  Entries with shapes this build
// does not support are rejected.
int gpjson_tune_cache_load(gpjson_ctx* ctx, const char* path, int* loaded) {
  if (!ctx || !path) return fail(GPJSON_ERR_INVALID_ARG, "NULL argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(GPJSON_ERR_IO, std::string("Failed to open file ") + path);
  char buf[512];
  int n = 0, rc = GPJSON_OK;
  if (!std::fgets(buf, sizeof(buf), f) || std::strncmp(buf, "gpjson-tune-cache 2", 19) != 0)
    rc = fail(GPJSON_ERR_INVALID_ARG, std::string("not a tuning cache file: ") + path);
  while (rc == GPJSON_OK && std::fgets(buf, sizeof(buf), f)) {
    char arch[128];
    TuneKey k;
    TuneEntry e;
    if (std::sscanf(buf, "%127s %d %d %d %d %d %d %d %d %d %d %d %lf", arch, &k.bucket, &k.levels, &k.batch,
                    &e.itune.block, &e.itune.wpt, &e.qtune.mode, &e.qtune.block, &e.qtune.group,
                    &e.qtune.waves_per_cu, &e.qtune.lines_per_wave, &e.qtune.qblock, &e.ms) != 13) {
      rc = fail(GPJSON_ERR_INVALID_ARG, std::string("malformed tuning cache line: ") + buf);
      break;
    }
    if (ctx->arch != arch) continue;
    if (!index_config_supported(e.itune.block, e.itune.wpt) || !query_config_supported(e.qtune.block, e.qtune.group) ||
        (e.qtune.mode != kLineMode && e.qtune.mode != kLineLongMode) || e.qtune.waves_per_cu < 1 ||
        e.qtune.waves_per_cu > 64 || e.qtune.lines_per_wave < 1 || e.qtune.lines_per_wave > 1024 ||
        e.qtune.qblock < 0 || e.qtune.qblock > (1 << 20) || k.levels < 0 ||
        k.levels > GPJSON_MAX_LEVELS || k.bucket < 0 || k.bucket > 62 || (k.batch != 0 && k.batch != 1)) {
      rc = fail(GPJSON_ERR_INVALID_ARG, std::string("unsupported shapes in tuning cache line: ") + buf);
      break;
    }
    ctx->tune_cache[k] = e;
    n++;
  }
  std::fclose(f);
  if (loaded) *loaded = n;
  return rc;
}

}  // extern "C"
