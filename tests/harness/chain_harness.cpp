// CPU build of the member-chain evaluator's state machine
// (gpjson_amd/csrc/chain_step.h), run line by line with a group of G lanes
// emulated in order (lane loads, lane parts, the group's combine, the update),
// so its results can be compared with the oracle's interpreter without a GPU.
#include <cstddef>
#include <vector>

#include "chain_step.h"

extern "C" {
// Evaluates every line with one group of G lanes; returns the number of steps
// taken, or -1 if a line needed more than max_steps steps.
int64_t ch_run(const uint8_t* F, int64_t N, const int64_t* NL, int64_t lines, const uint64_t* S, const uint64_t* LB,
               int64_t Wp, int L, const uint64_t* RC, const uint64_t* RQ, const uint8_t* ir, int R, int64_t* out,
               int G, int64_t max_steps) {
  // the records the index kernel writes: string, colon and raw-quote words
  const int64_t W = (N + 63) / 64;
  std::vector<uint64_t> rec((std::size_t)(W + 1) * gj::kRecWords, 0);
  for (int64_t w = 0; w < W; w++) {
    rec[(std::size_t)w * gj::kRecWords] = S[w];
    rec[(std::size_t)w * gj::kRecWords + 1] = RC[w];
    rec[(std::size_t)w * gj::kRecWords + 2] = RQ[w];
  }
  gj::ChainJobView J{F, NL, rec.data(), LB, out, N, lines, Wp, L, R, 0};
  std::vector<gj::ChainLoads> ld(G);
  int64_t total = 0;
  for (int64_t i = 0; i < lines; i++) {
    gj::ChainLane s{};
    s.phase = gj::kInit;
    s.i = i;
    int64_t steps = 0;
    while (s.phase != gj::kNeedLine) {
      gj::ChainPart p{-1, 0, 0, true};
      bool have = false;
      for (int r = 0; r < G; r++) {
        gj::chain_load(s, J, r, ld[r]);
        const gj::ChainPart part = gj::chain_lane(s, J, r, ld[r], ir);
        if (!have && part.pos >= 0) {
          p.pos = part.pos;
          p.ss = part.ss;
          p.kind = part.kind;
          have = true;
        }
        p.same = p.same && part.same;
      }
      gj::chain_update(s, J, ld[0], p, G, ir, true);
      if (++steps > max_steps) return -1;
    }
    total += steps;
  }
  return total;
}
}
