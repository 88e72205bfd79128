// CPU build of the member-chain evaluator's per-lane state machine
// (gpjson_amd/csrc/chain_step.h), run line by line so its results can be
// compared with the oracle's interpreter without a GPU.
#include "chain_step.h"

extern "C" {
// Evaluates every line with one lane; returns the number of steps taken, or -1
// if a line needed more than max_steps steps.
int64_t ch_run(const uint8_t* F, int64_t N, const int64_t* NL, int64_t lines, const uint64_t* S, const uint64_t* LB,
               int64_t Wp, int L, const uint64_t* RC, const uint64_t* RQ, const uint8_t* ir, int R, int64_t* out,
               int64_t max_steps) {
  gj::ChainJobView J{F, NL, S, LB, RC, RQ, out, N, lines, Wp, L, R, 0};
  int64_t total = 0;
  for (int64_t i = 0; i < lines; i++) {
    gj::ChainLane s{};
    s.phase = gj::kInit;
    s.i = i;
    int64_t steps = 0;
    while (s.phase != gj::kNeedLine) {
      gj::ChainLoads ld;
      gj::chain_load(s, J, ld);
      gj::chain_advance(s, J, ld, ir);
      if (++steps > max_steps) return -1;
    }
    total += steps;
  }
  return total;
}
}
