// autotune.cpp -- launch-shape search for the index and interpreter kernels.
//
// The reference hard-codes 8x1024 threads for the index kernels and 512x1024
// for find_value (J/stages/CombinedIndex.java:11-14, J/functions/
// QueryFunction.java:229; J/ = gpjson/src/main/java/com/koenv/gpjson/).  Here
// the tile shape (threads per block x 64-byte words per thread, i.e. the chunk
// of input one tile covers), the interpreter's block size, its persistent grid
// size and lines claimed per wavefront are measured on the caller's own input
// and the fastest shape is installed in the context.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "runtime.h"

using namespace gj;

namespace {

struct IdxCand {
  int block, wpt;
};
const IdxCand kIdx[] = {{256, 1}, {256, 2}, {256, 3}, {256, 4}, {512, 1}, {512, 2}, {128, 2}, {128, 4}};

struct QCand {
  int mode, block, group, wpc, lpw;
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
  if (query_mode < kLineMode || query_mode > kLineLongMode)
    return fail(GPJSON_ERR_INVALID_ARG, "query mode must be 0 (line), 1 (tiled), 2 (chain) or 3 (line, long lines to a wavefront)");
  GJ_HIP(hipSetDevice(ctx->device));
  GJ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->itune.block = index_block;
  ctx->itune.wpt = index_wpt;
  ctx->qtune.mode = query_mode;
  ctx->qtune.block = query_block;
  ctx->qtune.group = query_group;
  ctx->qtune.waves_per_cu = query_waves_per_cu;
  ctx->qtune.lines_per_wave = lines_per_wave;
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
// installs the fastest.  `report` (optional) receives one line per candidate.
int gpjson_autotune(gpjson_ctx* ctx, gpjson_buf* buf, const uint8_t* const* irs, const int64_t* ir_lens, int nq,
                    int iterations, char* report, size_t report_cap) {
  if (!ctx || !buf || !irs || !ir_lens || nq < 1 || iterations < 1)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  GJ_HIP(hipSetDevice(ctx->device));
  std::string rep;
  const IndexTune keep_i = ctx->itune;
  const QueryTune keep_q = ctx->qtune;
  double best = 1e30;
  IndexTune best_i = keep_i;
  char line[256];
  for (const IdxCand& c : kIdx) {
    ctx->itune.block = c.block;
    ctx->itune.wpt = c.wpt;
    gpjson_plan* p = nullptr;
    int rc = gpjson_plan_create(ctx, buf, irs, ir_lens, nq, &p);
    if (rc) {
      ctx->itune = keep_i;
      ctx->qtune = keep_q;
      return rc;
    }
    double sum = 0;
    rc = gpjson_plan_run(p, 0);
    for (int it = 0; it < iterations && rc == GPJSON_OK; it++) {
      double a = 0, b = 0;
      rc = gpjson_plan_run(p, 1);
      if (rc == GPJSON_OK) rc = gpjson_plan_timing(p, &a, &b);
      sum += a;
    }
    gpjson_plan_free(p);
    if (rc) {
      ctx->itune = keep_i;
      ctx->qtune = keep_q;
      return rc;
    }
    double ms = sum / iterations;
    std::snprintf(line, sizeof(line), "index block=%d wpt=%d tile=%lld B: %.4f ms\n", c.block, c.wpt,
                  (long long)index_tile_bytes(c.block, c.wpt), ms);
    rep += line;
    if (ms < best) {
      best = ms;
      best_i = ctx->itune;
    }
  }
  ctx->itune = best_i;
  std::vector<QCand> qc;
  for (int group : {1, 4, 16})
    for (int block : {128, 256}) qc.push_back({1, block, group, 0, 1});
  for (int group : {1, 4, 16, 64})
    for (int wpc : {16, 32})
      for (int lpw : {1, 4}) qc.push_back({0, 256, group, wpc, lpw});
  // long lines to a wavefront each
  for (int group : {1, 4})
    for (int wpc : {16, 32}) qc.push_back({3, 256, group, wpc, group == 1 ? 4 : 1});
  // member-chain evaluator (used only when every query is a member chain)
  for (int group : {1, 4, 16})
    for (int wpc : {8, 16, 32}) qc.push_back({2, 256, group, wpc, 1});
  int rc = GPJSON_OK;
  double bestq = 1e30;
  QueryTune best_q = keep_q;
  for (const QCand& c : qc) {
    ctx->qtune.mode = c.mode;
    ctx->qtune.block = c.block;
    ctx->qtune.group = c.group;
    if (c.wpc) ctx->qtune.waves_per_cu = c.wpc;
    ctx->qtune.lines_per_wave = c.lpw;
    // a plan caches the launch shape: prepare one per candidate
    gpjson_plan* p = nullptr;
    rc = gpjson_plan_create(ctx, buf, irs, ir_lens, nq, &p);
    if (rc) break;
    double sum = 0;
    rc = gpjson_plan_run(p, 0);
    for (int it = 0; it < iterations && rc == GPJSON_OK; it++) {
      double a = 0, b = 0;
      rc = gpjson_plan_run(p, 1);
      if (rc == GPJSON_OK) rc = gpjson_plan_timing(p, &a, &b);
      sum += b;
    }
    gpjson_plan_free(p);
    if (rc) break;
    double ms = sum / iterations;
    std::snprintf(line, sizeof(line), "query mode=%d block=%d group=%d waves_per_cu=%d lines_per_group=%d: %.4f ms\n",
                  c.mode, c.block, c.group, c.wpc, c.lpw, ms);
    rep += line;
    if (ms < bestq) {
      bestq = ms;
      best_q = ctx->qtune;
    }
  }
  if (rc) {
    ctx->itune = keep_i;
    ctx->qtune = keep_q;
    return rc;
  }
  ctx->qtune = best_q;
  std::snprintf(line, sizeof(line),
                "best index %dx%d (%.4f ms), query mode %d block %d group %d wpc %d lpg %d (%.4f ms)\n", best_i.block,
                best_i.wpt, best, best_q.mode, best_q.block, best_q.group, best_q.waves_per_cu,
                best_q.lines_per_wave, bestq);
  rep += line;
  if (report && report_cap) {
    size_t n = std::min(report_cap - 1, rep.size());
    std::copy(rep.begin(), rep.begin() + n, report);
    report[n] = 0;
  }
  return GPJSON_OK;
}

}  // extern "C"
