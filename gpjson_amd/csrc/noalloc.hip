// noalloc.hip -- gpjson_query_noalloc: the plain path (one index build, one
// interpreter launch) for callers that cannot allocate device memory.
//
// It runs the same kernels as gpjson_plan_create + gpjson_plan_run, but every
// array those kernels use beyond the caller's input and outputs -- tile
// states, ticket and flags, newline scratch rows, records, leveled bitmaps and
// their summaries, tile line table, the query job table and the IR bytes --
// is carved out of one file-scope device array, `scratch`, of
// GPJSON_NOALLOC_SCRATCH_BYTES.  Each GPU holds its own copy of the array from
// the first context created on it until the process ends.  The entry point
// calls no allocator: the job table and IR bytes reach the device as the
// arguments of a small kernel, not through a staging buffer.
//
// Every call on one GPU shares that GPU's array and keeps the previous call's
// contents; what must start at zero (tile states, ticket, flags, the work
// counter) is cleared by kernels at the start of each call.  Calls that overlap
// on one GPU (two contexts or two threads) race on the array.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "runtime.h"

namespace gj {

namespace {

constexpr int64_t kScratchBytes = GPJSON_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxQ = GPJSON_NOALLOC_MAX_QUERIES;
constexpr int kMaxIr = GPJSON_NOALLOC_MAX_IR_BYTES;

// Host code must never pass `scratch` itself to a kernel or a copy: in host
// code the name is a host-side stand-in, not a device address.  Only the
// per-device address from hipGetSymbolAddress (g_ptr, set once per device) is used.
__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The job table and IR bytes of one call, passed by value as kernel arguments.
struct JobTable {
  QueryJob jobs[kMaxQ];
  uint8_t ir[kMaxIr];
  int nq;
  int ir_bytes;
};
static_assert(sizeof(JobTable) <= 3072, "kernel arguments are limited to 4 KiB");

__global__ void put_jobs_kernel(JobTable t, QueryJob* jobs, uint8_t* ir) {
  const int i = threadIdx.x;
  if (i < t.nq) jobs[i] = t.jobs[i];
  for (int k = i; k < t.ir_bytes; k += blockDim.x) ir[k] = t.ir[k];
}

// The index's line count, and the lines the interpreter may take: none when
// they exceed what the caller's newline and span buffers hold.
__global__ void clamp_lines_kernel(const int64_t* lines, int64_t cap_lines, int64_t* query_lines, int64_t* info) {
  if (threadIdx.x != 0) return;
  const int64_t l = *lines;
  const bool over = l > cap_lines;
  info[0] = l;
  info[1] = over ? 1 : 0;
  *query_lines = over ? 0 : l;
}

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// control block: index tickets (0), bookkeeping flags (kCtlFlags), line count
// (kCtlLines), query line count (kCtlLines + 8), work counters (kCtlCounters)
constexpr int64_t kCtlFlags = kTicketBytes;
constexpr int64_t kCtlLines = kCtlFlags + 512;
constexpr int64_t kCtlCounters = kCtlLines + 512;
constexpr int64_t kCtlBytes = (kCtlCounters + kQueryCounterWords * 8 + 4095) / 4096 * 4096;

// Offsets of one call's arrays in the scratch array.
struct Layout {
  int64_t ctl;         // kCtlBytes of control block
  int64_t states;      // TileState[ntiles]
  int64_t nl_scratch;  // int64[ntiles * tile words] + block sums
  int64_t rec;         // uint64[W * kRecWords]
  int64_t lb;          // leveled bitmaps + summaries
  int64_t tile_lines;  // int64[ntiles + 2]
  int64_t jobs;        // QueryJob[nq]
  int64_t ir;          // IR bytes
  int64_t end;
};

Layout layout_for(int64_t n, int L, const IndexTune& t, int nq, int64_t ir_bytes, int64_t* ntiles_out) {
  const int64_t tile_bytes = index_tile_bytes(t.block, t.wpt);
  const int64_t ntiles = (n + tile_bytes - 1) / tile_bytes;
  const int64_t W = (n + 63) / 64, Wp = (W + 7) / 8 * 8;
  const int64_t cap = tile_bytes / 64;
  Layout o{};
  int64_t p = 0;
  o.ctl = p;
  p += kCtlBytes;
  o.states = p;
  p = align256(p + std::max<int64_t>(ntiles, 1) * (int64_t)sizeof(TileState));
  o.nl_scratch = p;
  p = align256(p + (std::max<int64_t>(ntiles, 1) * cap + ntiles / 256 + 2) * 8);
  o.rec = p;
  p = align256(p + W * kRecWords * 8);
  o.lb = p;
  p = align256(p + leveled_alloc_words(Wp, L) * 8);
  o.tile_lines = p;
  p = align256(p + (ntiles + 2) * 8);
  o.jobs = p;
  p = align256(p + (int64_t)nq * (int64_t)sizeof(QueryJob));
  o.ir = p;
  p = align256(p + ir_bytes + 1);
  o.end = p;
  *ntiles_out = ntiles;
  return o;
}

}  // namespace

int noalloc_scratch_setup(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return fail(GPJSON_ERR_INVALID_ARG, "device id out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    // each GPU holds its own copy of the array at its own address
    GJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(GPJSON_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[device] = p;
  }
  *out = g_ptr[device];
  return GPJSON_OK;
}

}  // namespace gj

using namespace gj;

extern "C" int gpjson_query_noalloc(gpjson_ctx* ctx, const uint8_t* input, int64_t n, const uint8_t* const* irs,
                                    const int64_t* ir_lens, int nq, int64_t* newlines, int64_t newline_cap,
                                    int64_t* spans, int64_t span_cap, int64_t* info) {
  if (!ctx || !irs || !ir_lens || !newlines || !spans || !info || n < 0 || (n > 0 && !input))
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  if (nq < 1 || nq > kMaxQ)
    return fail(GPJSON_ERR_INVALID_ARG, "nq must be in [1, " + std::to_string(kMaxQ) + "]");
  if (newline_cap < 1 || span_cap < 0) return fail(GPJSON_ERR_INVALID_ARG, "bad buffer capacity");
  if (!ctx->noalloc_scratch) return fail(GPJSON_ERR_HIP, "no scratch array on this device");
  int64_t ir_total = 0;
  for (int q = 0; q < nq; q++) {
    if (ir_lens[q] < 0 || (ir_lens[q] > 0 && !irs[q])) return fail(GPJSON_ERR_INVALID_ARG, "bad IR");
    ir_total += ir_lens[q];
  }
  if (ir_total > kMaxIr)
    return fail(GPJSON_ERR_CAPACITY, "IR bytes exceed " + std::to_string(kMaxIr) + " (" + std::to_string(ir_total) + ")");
  int L = 0, R = 1;
  ir_shape(irs, ir_lens, nq, &L, &R);
  const IndexTune t = ctx->itune;
  int64_t ntiles = 0;
  const Layout o = layout_for(n, L, t, nq, ir_total, &ntiles);
  if (o.end > kScratchBytes)
    return fail(GPJSON_ERR_CAPACITY, "input of " + std::to_string(n) + " bytes at " + std::to_string(L) +
                                         " levels needs " + std::to_string(o.end) + " bytes of scratch; the limit is " +
                                         std::to_string(kScratchBytes));
  GJ_HIP(hipSetDevice(ctx->device));
  uint8_t* base = (uint8_t*)ctx->noalloc_scratch;
  hipStream_t stream = ctx->stream;
  const int64_t W = (n + 63) / 64, Wp = (W + 7) / 8 * 8;
  int64_t* lines_dev = (int64_t*)(base + o.ctl + kCtlLines);
  int64_t* query_lines = (int64_t*)(base + o.ctl + kCtlLines + 8);
  unsigned long long* counter = (unsigned long long*)(base + o.ctl + kCtlCounters);

  // ---- index ----
  FileJob j{};
  j.F = input;
  j.N = n;
  j.W = W;
  j.Wp = Wp;
  j.rec = (uint64_t*)(base + o.rec);
  j.LB = (uint64_t*)(base + o.lb);
  j.Ws = summary_stride(Wp);
  j.LS = summaries_of(j.LB, Wp, L);
  j.NL = newlines;
  j.nl_cap = newline_cap;
  j.tile_lines = (int64_t*)(base + o.tile_lines);
  j.tile0 = 0;
  j.ntiles = ntiles;
  j.L = L;
  IndexLaunch il{};
  il.one = j;
  il.njobs = 1;
  il.total_tiles = ntiles;
  il.states = (TileState*)(base + o.states);
  il.ticket = (unsigned int*)(base + o.ctl);
  il.speculate = 1;
  il.nl_scratch_cap = index_tile_bytes(t.block, t.wpt) / 64;
  il.nl_scratch = (int64_t*)(base + o.nl_scratch);
  int rc = enqueue_index_launches(il, t, L, false, (int*)(base + o.ctl + kCtlFlags), lines_dev, stream);
  if (rc) return rc;

  // ---- lines the outputs hold ----
  const int64_t per_line = (int64_t)nq * R * 2;
  const int64_t cap_lines = std::min<int64_t>(newline_cap, span_cap / per_line);
  hipLaunchKernelGGL(clamp_lines_kernel, dim3(1), dim3(64), 0, stream, (const int64_t*)lines_dev, cap_lines,
                     query_lines, info);
  GJ_HIP(hipGetLastError());

  // ---- job table and IR bytes, as kernel arguments ----
  JobTable tab;
  std::memset(&tab, 0, sizeof(tab));
  tab.nq = nq;
  tab.ir_bytes = (int)ir_total;
  uint8_t* ir_dev = base + o.ir;
  int64_t off = 0;
  for (int q = 0; q < nq; q++) {
    QueryJob& qj = tab.jobs[q];
    qj.F = input;
    qj.N = n;
    qj.NL = newlines;
    qj.lines = 0;  // read on the device (query_lines)
    qj.rec = j.rec;
    qj.LB = j.LB;
    qj.LS = j.LS;
    qj.Ws = j.Ws;
    qj.Wp = Wp;
    qj.L = L;
    qj.R = R;
    qj.ir = ir_dev + off;
    qj.ir_len = ir_lens[q];
    qj.out = spans;  // job q writes at spans + q * lines * R * 2
    qj.item0 = 0;
    if (ir_lens[q]) std::memcpy(tab.ir + off, irs[q], (size_t)ir_lens[q]);
    off += ir_lens[q];
  }
  QueryJob* jobs_dev = (QueryJob*)(base + o.jobs);
  hipLaunchKernelGGL(put_jobs_kernel, dim3(1), dim3(256), 0, stream, tab, jobs_dev, ir_dev);
  GJ_HIP(hipGetLastError());

  // ---- one interpreter launch (line mode, the context's shape) ----
  GJ_HIP(zero_u64(counter, kQueryCounterWords, stream));
  QueryLaunch ql{};
  ql.jobs = jobs_dev;
  ql.njobs = nq;
  ql.lines_dev = query_lines;
  ql.counter = counter;
  ql.lines_per_wave = ctx->qtune.lines_per_wave;
  ql.ir_base = ir_dev;
  ql.ir_bytes = ir_total;
  const int block = ctx->qtune.block, group = ctx->qtune.group;
  const int grid = std::max(1, ctx->num_cus * ctx->qtune.waves_per_cu / (block / 64));
  GJ_HIP(query_launch(ql, block, group, L, grid, stream));
  return GPJSON_OK;
}
