// chain_step.h -- the member-chain evaluator's state machine, one step at a
// time.  Header-only: compiled by hipcc into chain_kernel.hip and by g++ into
// the CPU harness that checks it line by line against the oracle.
//
// A member chain $.k0.k1...k(D-1) compiles to (KEY k DOWN)^D STORE END
// (J/jsonpath/IRBuilder.java:19-81).  For such a program the interpreter of
// K/find_value.cu:21-263 (== J/sequential/Sequential.java:146-337) reduces,
// per line, to:
//
//   hi = line end; j = line start
//   if !(j < hi && j < n): no result
//   for d in 0..D-1:
//     p = first ':' at level d in [j, min(hi, n)) whose key is k_d   (none: no result)
//     j = p + 1;  if !(j < hi && j < n): no result
//     hi = first level-d bit in (j, hi), or -1                         (MOVE_DOWN)
//   skip blanks from j (j < n); result (j, hi)                        (STORE)
//
// A level d >= L holds no event, and a search over an empty window (hi == -1
// included) finds none, exactly as in the general interpreter.  The key test
// at a ':' at p is the reference's: e = the last string bit in (start, p),
// ss = e - len, ss >= start, F[ss] == '"' (raw-quote bitmap), then the len
// bytes after ss equal the key.
//
// A group of G lanes evaluates one line.  Each step every lane of the group
// loads kChainWords consecutive words of each bitmap it needs (lane r the words
// wbase + r*kChainWords ...), so a group step reads whole cache lines; each
// lane then computes its partial result (chain_lane), the group combines them
// (first lane with a result, or AND for a byte compare), and every lane applies
// the same update (chain_update).  The state is uniform within the group.
#pragma once
#include <stdint.h>

#include "bitops.h"
#include "query_common.h"

namespace gj {

constexpr int kChainWords = 4;  // bitmap words a lane loads per step

enum ChainPhase : int {
  kNeedLine = 0,  // idle: claim a line
  kInit = 1,      // load the line's newline pair
  kKey = 2,       // scan level d for a ':' whose key matches
  kVerify = 3,    // compare the candidate key's bytes
  kDelim = 4,     // MOVE_DOWN: scan level d for the first bit in (j, hi)
  kStore = 5,     // skip blanks at j, write the result
  kDone = 6,      // every line claimed
};

// The arrays of the job a group is working on.
struct ChainJobView {
  const uint8_t* F;
  const int64_t* NL;
  const uint64_t* rec;  // string, colon and raw-quote words of each input word
  const uint64_t* LB;
  int64_t* out_base;  // int64[lines][R][2]
  int64_t N, lines, Wp;
  int L, R, ir0;      // ir0: offset of the job's IR in the IR bytes
};

struct ChainLane {
  int phase;
  // j: the interpreter's position (moved by events only); dpos: where the
  // MOVE_DOWN scan continues; cand / ss: key candidate ':' and its quote
  int64_t i, start, j, hi, cand, ss, dpos;
  int d, koff, klen, vdone;
};

struct ChainLoads {
  int64_t nl0, nl1, w0;  // w0: first word this lane loaded
  uint64_t lw[kChainWords], cw[kChainWords];
  uint64_t sw[kChainWords + 1], qw[kChainWords + 1];  // [0]: the word before w0
  uint8_t fb[8];
};

// A lane's contribution to the group's step.
struct ChainPart {
  int64_t pos;  // kKey: candidate ':'; kDelim: first bit; kStore: first non-blank; -1: none
  int64_t ss;   // kKey: the candidate's quote position
  int kind;     // kKey: 1 match (empty key), 2 needs a byte compare
  bool same;    // kVerify: this lane's key bytes are equal
};

GJ_HD uint64_t chain_mask_from(int lo_bit) { return ~0ull << lo_bit; }
GJ_HD uint64_t chain_mask_upto(int hi_bit) { return hi_bit >= 63 ? ~0ull : ((2ull << hi_bit) - 1ull); }
GJ_HD bool chain_blank(uint8_t c) { return c == '\n' || c == '\r' || c == '\t' || c == ' '; }

// a[idx] through selects (a dynamic index would move the array to scratch
// memory on the GPU); 0 when idx is out of range.
GJ_HD uint64_t chain_pick(const uint64_t (&a)[kChainWords + 1], int idx) {
  uint64_t r = 0;
#pragma unroll
  for (int k = 0; k <= kChainWords; k++) r = idx == k ? a[k] : r;
  return r;
}

GJ_HD int chain_varint(const uint8_t* ir, int* pc) {
  int v = 0, shift = 0;
  for (;;) {
    const uint8_t b = ir[(*pc)++];
    v |= (int)(b & 0x7F) << shift;
    if (!(b & 0x80) || shift > 21) break;
    shift += 7;
  }
  return v;
}

// First word of the group's window this step.
GJ_HD int64_t chain_wbase(const ChainLane& s) { return (s.phase == kKey ? s.j : s.dpos) >> 6; }

// Lane r's loads for the group's current phase.
GJ_HD void chain_load(const ChainLane& s, const ChainJobView& J, int r, ChainLoads& ld) {
  ld.nl0 = ld.nl1 = ld.w0 = 0;
#pragma unroll
  for (int u = 0; u < kChainWords; u++) ld.lw[u] = ld.cw[u] = 0;
#pragma unroll
  for (int u = 0; u <= kChainWords; u++) ld.sw[u] = ld.qw[u] = 0;
#pragma unroll
  for (int u = 0; u < 8; u++) ld.fb[u] = 0;
  if (s.phase == kInit) {
    ld.nl0 = J.NL[s.i];
    ld.nl1 = s.i + 1 < J.lines ? J.NL[s.i + 1] : J.N;
  } else if (s.phase == kKey || s.phase == kDelim) {
    const int64_t w0 = chain_wbase(s) + (int64_t)r * kChainWords;
    ld.w0 = w0;
    const uint64_t* lv = J.LB + (int64_t)s.d * J.Wp;
    const int64_t wlast = ((s.hi < J.N ? s.hi : J.N) - 1) >> 6;  // last word of the window
#pragma unroll
    for (int u = 0; u < kChainWords; u++) {
      const int64_t w = w0 + u;
      if (w <= wlast) {
        ld.lw[u] = lv[w];
        if (s.phase == kKey) {
          ld.cw[u] = rec_colon(J.rec, w);
          ld.sw[u + 1] = rec_s(J.rec, w);
          ld.qw[u + 1] = rec_quote(J.rec, w);
        }
      }
    }
    if (s.phase == kKey && w0 > (s.start >> 6) && w0 - 1 <= wlast) {
      ld.sw[0] = rec_s(J.rec, w0 - 1);
      ld.qw[0] = rec_quote(J.rec, w0 - 1);
    }
  } else if (s.phase == kVerify) {
    const int b0 = s.vdone + 8 * r;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (b0 + u < s.klen) ld.fb[u] = J.F[s.ss + 1 + b0 + u];
  } else if (s.phase == kStore) {
    const int64_t b0 = s.j + 8 * (int64_t)r;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (b0 + u < J.N) ld.fb[u] = J.F[b0 + u];
  }
}

// Lane r's partial result of the step.
GJ_HD ChainPart chain_lane(const ChainLane& s, const ChainJobView& J, int r, const ChainLoads& ld,
                           const uint8_t* ir) {
  ChainPart out{-1, 0, 0, true};
  const int64_t N = J.N;
  if (s.phase == kKey) {
    const int64_t to = s.hi < N ? s.hi : N;
    const int64_t wlast = (to - 1) >> 6;
    const int64_t wstart = s.start >> 6;
#pragma unroll
    for (int u = 0; u < kChainWords; u++) {
      const int64_t w = ld.w0 + u;
      uint64_t v = ld.lw[u] & ld.cw[u];
      if (w == (s.j >> 6)) v &= chain_mask_from((int)(s.j & 63));
      if (w == wlast) v &= chain_mask_upto((int)((to - 1) & 63));
      if (w > wlast) v = 0;
      while (v && out.pos < 0) {
        const int64_t p = (w << 6) + ctz64(v);
        v &= v - 1;
        const int64_t k = p - 1;
        if (k <= s.start) continue;
        // e = the last string bit in (start, k]: the loaded words first
        int64_t e = -1;
        int64_t wi = k >> 6;
        int idx = (int)(wi - ld.w0) + 1;  // sw[] slot of word wi
        bool slow = idx < 0 || idx > kChainWords;
        uint64_t sv = slow ? 0 : chain_pick(ld.sw, idx);
        sv &= chain_mask_upto((int)(k & 63));
        while (!slow) {
          if (wi == wstart) sv &= ~chain_mask_upto((int)(s.start & 63));
          if (sv) {
            e = (wi << 6) + 63 - clz64(sv);
            break;
          }
          if (wi <= wstart) break;
          wi--;
          idx--;
          if (idx < 0) {
            slow = true;
            break;
          }
          sv = chain_pick(ld.sw, idx);
        }
        if (slow) {
          // the string run starts before the loaded words: read on
          wi = k >> 6;
          sv = rec_s(J.rec, wi) & chain_mask_upto((int)(k & 63));
          for (;;) {
            if (wi == wstart) sv &= ~chain_mask_upto((int)(s.start & 63));
            if (sv) {
              e = (wi << 6) + 63 - clz64(sv);
              break;
            }
            if (wi <= wstart) break;
            wi--;
            sv = rec_s(J.rec, wi);
          }
        }
        if (e < 0) continue;
        const int64_t sc = e - s.klen;
        if (sc < s.start) continue;
        const int64_t qi = sc >> 6;
        const int qidx = (int)(qi - ld.w0) + 1;
        const uint64_t q = (qidx >= 0 && qidx <= kChainWords) ? chain_pick(ld.qw, qidx) : rec_quote(J.rec, qi);
        if (!((q >> (sc & 63)) & 1ull)) continue;
        out.pos = p;
        out.ss = sc;
        out.kind = s.klen == 0 ? 1 : 2;
      }
    }
  } else if (s.phase == kVerify) {
    const int b0 = s.vdone + 8 * r;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const bool in = b0 + u < s.klen;
      const uint8_t kb = ir[in ? s.koff + b0 + u : s.koff];
      out.same = out.same && (!in || ld.fb[u] == kb);
    }
  } else if (s.phase == kDelim) {
    // first level-d bit in [dpos, hi); hi > j >= 0 here
    const int64_t whi = (s.hi - 1) >> 6;
#pragma unroll
    for (int u = kChainWords - 1; u >= 0; u--) {
      const int64_t w = ld.w0 + u;
      uint64_t v = ld.lw[u];
      if (w == (s.dpos >> 6)) v &= chain_mask_from((int)(s.dpos & 63));
      if (w < (s.dpos >> 6)) v = 0;
      if (w == whi) v &= chain_mask_upto((int)((s.hi - 1) & 63));
      if (w > whi) v = 0;
      if (v) out.pos = (w << 6) + ctz64(v);
    }
  } else if (s.phase == kStore) {
    const int64_t b0 = s.j + 8 * (int64_t)r;
#pragma unroll
    for (int u = 7; u >= 0; u--)
      if (b0 + u < N && !chain_blank(ld.fb[u])) out.pos = b0 + u;
  }
  return out;
}

// Sets up the KEY op at IR offset kpc (ir[kpc] == KEY) or the final STORE.
GJ_HD void chain_next_op(ChainLane& s, const uint8_t* ir, int kpc) {
  if (ir[kpc] == kStoreResult) {
    s.klen = -1;
    return;
  }
  int pc = kpc + 1;
  s.klen = chain_varint(ir, &pc);
  s.koff = pc;
}

// The group's update (identical in every lane).  `p` is the combined result:
// the part of the first lane (in lane order) with pos >= 0 (all -1 if none),
// and for kVerify the AND of every lane's `same`.  G: lanes in the group.
// ld: this lane's loads (kInit uses the newline pair, loaded alike by all
// lanes).  Writes the -1 fill of a new line and the result to J.out_base;
// write_out selects the one lane that writes.
GJ_HD void chain_update(ChainLane& s, const ChainJobView& J, const ChainLoads& ld, const ChainPart& p, int G,
                        const uint8_t* ir, bool write_out) {
  const int64_t N = J.N;
  const int64_t span = (int64_t)G * kChainWords;  // words the group loaded
  if (s.phase == kInit) {
    s.start = ld.nl0;
    if (write_out) {
      int64_t* out = J.out_base + s.i * (int64_t)J.R * 2;
      for (int k = 0; k < J.R * 2; k++) out[k] = -1;
    }
    s.j = s.start;
    s.hi = ld.nl1;
    s.d = 0;
    chain_next_op(s, ir, J.ir0);
    s.phase = (s.start < s.hi && s.start < N && s.d < J.L && s.klen >= 0) ? kKey : kNeedLine;
  } else if (s.phase == kKey) {
    const int64_t to = s.hi < N ? s.hi : N;
    if (p.pos >= 0 && p.kind == 1) {
      s.j = p.pos + 1;
      s.dpos = s.j + 1;
      s.phase = (s.j < s.hi && s.j < N) ? kDelim : kNeedLine;
    } else if (p.pos >= 0) {
      s.cand = p.pos;
      s.ss = p.ss;
      s.vdone = 0;
      s.phase = kVerify;
    } else {
      s.j = (chain_wbase(s) + span) << 6;
      if (s.j >= to) s.phase = kNeedLine;  // no matching key in the window
    }
  } else if (s.phase == kVerify) {
    s.vdone += 8 * G;
    if (!p.same) {
      s.j = s.cand + 1;  // keep scanning after the candidate
      s.phase = kKey;
    } else if (s.vdone >= s.klen) {
      s.j = s.cand + 1;
      s.dpos = s.j + 1;
      s.phase = (s.j < s.hi && s.j < N) ? kDelim : kNeedLine;
    }
  } else if (s.phase == kDelim) {
    const int64_t lim = s.hi < N ? s.hi : N;  // words at or beyond lim were not loaded
    const int64_t next = (chain_wbase(s) + span) << 6;
    const bool resolved = p.pos >= 0 || s.dpos >= s.hi || next >= lim;
    if (!resolved) {
      s.dpos = next;
    } else {
      s.hi = p.pos;  // -1: no delimiter at this level
      s.d++;
      chain_next_op(s, ir, s.koff + s.klen + 1);  // skip the DOWN byte
      if (s.klen < 0) s.phase = kStore;
      else s.phase = (s.d < J.L && s.hi > s.j && s.j < N) ? kKey : kNeedLine;
    }
  } else if (s.phase == kStore) {
    const int64_t next = s.j + 8 * (int64_t)G;
    if (p.pos >= 0 || next >= N) {
      if (write_out) {
        int64_t* out = J.out_base + s.i * (int64_t)J.R * 2;
        out[0] = p.pos >= 0 ? p.pos : N;
        out[1] = s.hi;
      }
      s.phase = kNeedLine;
    } else {
      s.j = next;
    }
  }
}

}  // namespace gj
