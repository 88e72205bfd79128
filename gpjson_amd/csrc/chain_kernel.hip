// chain_kernel.hip -- member-chain evaluator for gfx950.
//
// The general interpreter (query_kernel.hip) runs lines through a switch over
// opcodes; lanes at different opcodes and scan positions serialize, and every
// lane's scattered 8-byte loads pull whole cache lines that are evicted before
// the rest of them is used.  For member chains -- the programs
// (KEY k DOWN)^D STORE END of $.k0.k1...k(D-1) -- this kernel runs the state
// machine of chain_step.h with a group of G lanes per line: every iteration
// each lane loads kChainWords consecutive words of the bitmaps its group's
// phase needs (a group step reads whole cache lines of each bitmap), the
// group combines the lanes' partial results with a ballot, and every lane
// applies the same update.  A group that finishes a line takes a new one at
// the next iteration (one atomic per wavefront for all idle groups), so
// groups never wait for each other's lines.
#include <hip/hip_runtime.h>

#include "chain_step.h"
#include "query_common.h"

namespace gj {

namespace {

constexpr int kItemLdsJobs = 1024;  // job item0 table kept in LDS up to this many jobs

__device__ __forceinline__ unsigned long long shfl64(unsigned long long v, int src) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl(lo, src);
  hi = __shfl(hi, src);
  return ((unsigned long long)hi << 32) | lo;
}

}  // namespace

template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK) void chain_kernel(QueryLaunch launch) {
  __shared__ uint8_t s_ir[kIrLdsBytes];
  __shared__ int64_t s_item0[kItemLdsJobs];
  for (int k = threadIdx.x; k < launch.ir_bytes; k += BLOCK) s_ir[k] = launch.ir_base[k];
  const bool item_lds = launch.njobs <= kItemLdsJobs && !launch.lines_dev;
  if (item_lds)
    for (int k = threadIdx.x; k < launch.njobs; k += BLOCK) s_item0[k] = launch.jobs[k].item0;
  __syncthreads();

  const int lane = __lane_id();
  const int r = lane % G;         // lane within the group
  const int gbase = lane - r;     // first lane of the group
  const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << gbase);
  const int64_t common = launch.lines_dev ? *launch.lines_dev : 0;
  const int64_t total = launch.lines_dev ? common * launch.njobs
                                         : (launch.total_dev ? *launch.total_dev : launch.total_items);

  ChainLane s{};
  s.phase = kNeedLine;
  ChainJobView J{};
  int job = -1;
  int64_t item0 = 0;

  for (;;) {
    // ---- claim lines for idle groups: one atomic per wavefront ----
    const unsigned long long idle = __ballot(s.phase == kNeedLine && r == 0);
    if (idle) {
      const int leader = __builtin_ctzll(idle);
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(launch.counter, (unsigned long long)__popcll(idle));
      base = shfl64(base, leader);
      if (s.phase == kNeedLine) {
        const int64_t item = (int64_t)(base + __popcll(idle & ((1ull << gbase) - 1ull)));
        if (item >= total) {
          s.phase = kDone;
        } else {
          // map the item to (job, line); reload the job's arrays when it changes
          int q;
          if (launch.lines_dev) {
            q = (int)(item / common);
            s.i = item - (int64_t)q * common;
          } else {
            int lo = (job >= 0 && item0 <= item) ? job : 0, hi = (int)launch.njobs - 1;
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              const int64_t m0 = item_lds ? s_item0[mid] : launch.jobs[mid].item0;
              if (m0 <= item) lo = mid; else hi = mid - 1;
            }
            q = lo;
          }
          if (q != job) {
            const QueryJob& Q = launch.jobs[q];
            job = q;
            item0 = Q.item0;
            J.F = Q.F;
            J.NL = Q.NL;
            J.rec = Q.rec;
            J.LB = Q.LB;
            J.N = Q.N;
            J.Wp = Q.Wp;
            J.L = Q.L;
            J.R = Q.R;
            J.ir0 = (int)(Q.ir - launch.ir_base);
            J.lines = launch.lines_dev ? common : Q.lines;
            J.out_base = launch.lines_dev ? Q.out + (int64_t)q * common * Q.R * 2 : Q.out;
          }
          if (!launch.lines_dev) s.i = item - item0;
          s.phase = kInit;
        }
      }
    }
    if (__ballot(s.phase != kDone) == 0) break;

    // ---- one step: loads, this lane's part, group combine, update ----
    ChainLoads ld;
    ChainPart part{-1, 0, 0, true};
    if (s.phase != kDone) {
      chain_load(s, J, r, ld);
      part = chain_lane(s, J, r, ld, s_ir);
    }
    const unsigned long long has = __ballot(part.pos >= 0) & gmask;
    const unsigned long long differ = __ballot(!part.same) & gmask;
    ChainPart p{-1, 0, 0, differ == 0};
    if (has) {
      const int src = __builtin_ctzll(has);
      p.pos = (int64_t)shfl64((unsigned long long)part.pos, src);
      p.ss = (int64_t)shfl64((unsigned long long)part.ss, src);
      p.kind = __shfl(part.kind, src);
    }
    if (s.phase != kDone) chain_update(s, J, ld, p, G, s_ir, r == 0);
  }
}

// Shape check for a member chain: (KEY len bytes DOWN)^D STORE END with D >= 1.
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

template <int BLOCK>
static hipError_t chain_launch_b(const QueryLaunch& L, int group, int grid, hipStream_t stream) {
  switch (group) {
    case 1: hipLaunchKernelGGL((chain_kernel<BLOCK, 1>), dim3(grid), dim3(BLOCK), 0, stream, L); break;
    case 4: hipLaunchKernelGGL((chain_kernel<BLOCK, 4>), dim3(grid), dim3(BLOCK), 0, stream, L); break;
    case 16: hipLaunchKernelGGL((chain_kernel<BLOCK, 16>), dim3(grid), dim3(BLOCK), 0, stream, L); break;
    case 64: hipLaunchKernelGGL((chain_kernel<BLOCK, 64>), dim3(grid), dim3(BLOCK), 0, stream, L); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// group: lanes per line (1, 4, 16 or 64).
hipError_t chain_launch(const QueryLaunch& L, int block, int group, int grid, hipStream_t stream) {
  switch (block) {
    case 64: return chain_launch_b<64>(L, group, grid, stream);
    case 128: return chain_launch_b<128>(L, group, grid, stream);
    case 256: return chain_launch_b<256>(L, group, grid, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gj
