// query_common.h -- host/device shared definitions of the per-line interpreter.
#pragma once
#include <stdint.h>

#include "index_common.h"

namespace gj {

// IR opcodes: ordinal of the reference's Opcode enum
// (J/jsonpath/Opcode.java:3-11, K/find_value.cu:4-10).
enum Opcode : uint8_t {
  kEnd = 0,
  kStoreResult = 1,
  kMoveUp = 2,
  kMoveDown = 3,
  kMoveToKey = 4,
  kMoveToIndex = 5,
  kExpressionStringEquals = 6,
};

// One (file, query) pair of an interpreter launch.
struct QueryJob {
  const uint8_t* F;
  int64_t N;
  const int64_t* NL;  // newline index
  int64_t lines;
  const uint64_t* rec; // records: string, colon (key candidates), raw-quote and level-0 words per input word
  const uint64_t* LB; // leveled bitmaps of levels 1 .. L-1, level stride Wp
  const uint64_t* LS; // level summaries (bit per leveled word), stride Ws; may be null
  int64_t Ws;
  int64_t Wp;
  int L;              // levels present (0 in the records, the rest in LB)
  int R;              // result slots per line
  const uint8_t* ir;  // device copy of the IR bytes
  int64_t ir_len;
  int64_t* out;       // int64[lines][R][2]
  int64_t item0;      // first global work item (line) of this job
  // batched launches (QueryLaunch.qblock): this job is query qf of the nqf
  // queries of its file, whose jobs are consecutive and share the file's lines
  int qf;
  int nqf;
};

struct QueryLaunch {
  const QueryJob* jobs;           // device array
  int64_t njobs;
  int64_t total_items;            // sum of lines over jobs (when lines_dev is null)
  // When set, every job runs over the same index whose line count is read from
  // device memory (no host round trip between index and query): item g is
  // line g % lines of job g / lines, and job q writes its spans at
  // jobs[q].out + q * lines * R * 2.
  const int64_t* lines_dev;
  // When set (and lines_dev is null), the job table's lines / item0 were
  // written on the device and the item count is read from here.
  const int64_t* total_dev;
  unsigned long long* counter;    // kQueryCounterWords words, zeroed before launch
  int lines_per_wave;             // lines each lane group takes per wavefront claim
  // Long lines (line mode): a lane group that claims a line longer than
  // long_min bytes appends the work item to long_items (up to long_cap items)
  // instead of walking it alone; a second launch then evaluates each of those
  // lines with a whole wavefront.  long_items == nullptr: no deferral.
  // long_count[0]: items appended; long_count[1]: the second launch's claims.
  int64_t* long_items;
  unsigned long long* long_count;
  int64_t long_cap;
  int64_t long_min;
  // When set (one job, lines_dev set): the launch evaluates only the lines
  // line_list[0 .. min(*list_count, list_cap)), in any order (the lines the
  // index kernel left to the interpreter after evaluating the others itself).
  const int64_t* line_list;
  const unsigned long long* list_count;
  int64_t list_cap;
  // Every job's IR lies in [ir_base, ir_base + ir_bytes).  When ir_bytes fits
  // kIrLdsBytes, each block copies it to LDS once and interprets from there.
  const uint8_t* ir_base;
  int64_t ir_bytes;
  // Work order of several queries over one file's lines (line mode).  0: query
  // after query (items q * lines + i).  B > 0: the queries take turns over
  // blocks of B lines (items of block c: query 0 on lines [cB, cB + B), then
  // query 1 on the same lines, ...; the last block holds the lines left), so
  // the queries of a line run close together and share its cached words.
  int qblock;
};

// Item r of a file's nq * lines work items when its queries take turns over
// blocks of B lines: the query q and line i it stands for.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline void split_item(int64_t r, int64_t lines, int64_t nq, int64_t B, int64_t* q, int64_t* i) {
  const int64_t full = lines / B;  // whole blocks
  const int64_t span = nq * B;     // items of a whole block
  if (r < full * span) {
    const int64_t c = r / span, w = r - c * span;
    *q = w / B;
    *i = c * B + (w - *q * B);
  } else {
    const int64_t t = lines - full * B;  // lines of the last block (> 0 when items are left)
    const int64_t w = r - full * span;
    *q = w / t;
    *i = full * B + (w - *q * t);
  }
}

constexpr int kIrLdsBytes = 4096;
// Work counters of an interpreter launch: kQuerySegments counters,
// kQueryCounterStride words (256 bytes) apart, at QueryLaunch.counter; the
// host zeroes kQueryCounterWords words there before each launch.  A launch
// splits its work items into up to kQuerySegments ranges with a counter each,
// so its claims do not all queue on one address.
constexpr int kQuerySegments = 8;
constexpr int kQueryCounterStride = 32;
constexpr int kQueryCounterWords = kQuerySegments * kQueryCounterStride;
static_assert(kQuerySegments < 32 && (kQuerySegments & (kQuerySegments - 1)) == 0,
              "a power of two of ranges, tracked in a 32-bit mask");
// Lines at least this long are evaluated by a whole wavefront (line mode).
constexpr int64_t kLongLineBytes = 8192;

}  // namespace gj
