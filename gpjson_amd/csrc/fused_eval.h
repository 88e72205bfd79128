// fused_eval.h -- member-chain queries evaluated inside the index kernel, from
// one index tile's masks.  Header-only: compiled by hipcc for gfx950 (LDS
// views of the tile) and by g++ for the CPU harness that checks it against the
// oracle (array views).
//
// A member chain is the IR (KEY k DOWN)^D STORE END ($.a.b.c).  Its semantics
// are the reference interpreter's (K/find_value.cu:21-263 ==
// J/sequential/Sequential.java:146-337; K/ = gpjson/src/main/resources/com/
// koenv/gpjson/kernels/, J/ = gpjson/src/main/java/com/koenv/gpjson/) restated
// for that shape: a line [start, end) is walked level by level,
//   c_0     = the first ':' at level 0 in [start, end) whose key is k_0,
//   le_1    = the first structural byte at level 0 in (c_0 + 1, end), or -1,
//   c_l     = the first ':' at level l in [c_{l-1} + 1, le_l) whose key is k_l,
//   le_{l+1}= the first structural byte at level l in (c_l + 1, le_l), or -1,
// with no result as soon as a c_l is missing, c_l + 1 >= le_l (the walk's
// cursor leaves the level), or a level after the first has le_l = -1; the
// result is (the first non-blank byte at or after c_{D-1} + 1, le_D), where
// le_D may be -1.  The key test of a ':' at p is the reference's: e = the
// last in-string byte in (start, p); the key starts with a '"' at ss = e -
// klen (ss >= start) and the klen bytes after ss equal the key.
//
// Tile-local positions: bit b of tile word w is byte 64 w + b of the tile.
// Lines that start and end inside one tile are evaluated here; a line that
// starts before the tile's first newline (it began in an earlier tile) or
// runs past the tile's last newline is left to the interpreter (it crosses a
// tile boundary), except the file's first and last lines, whose start / end
// is the file's.
#pragma once
#include <stdint.h>

#include "bitops.h"

// member functions (GJ_HD is `static inline` on the host)
#if defined(__HIPCC__)
#define GJ_FE_M __host__ __device__ __forceinline__
#else
#define GJ_FE_M inline
#endif

namespace gj {

// Most lines (newlines) a tile of tw words evaluates here.  The tile's
// newline scratch row (tw slots) then holds its newline offsets at [0, c) and
// the lines' spans at [tw / 3, tw / 3 + 2 (c + 1)).
GJ_HD int fused_line_cap(int tw) { return tw / 3 - 1; }
GJ_HD int fused_span_base(int tw) { return tw / 3; }

GJ_HD uint64_t fe_mask_from(int b) { return ~0ull << b; }
GJ_HD uint64_t fe_mask_upto(int b) { return b >= 63 ? ~0ull : ((2ull << b) - 1ull); }

// First half of the key test of the ':' at tile-local position p: finds e and
// ss = e - klen from the tile's string (S), newline (NL) and raw-quote (Q)
// words, and checks ss >= start and the '"' at ss.  tile0: the tile starts the
// file (its first line starts at byte 0, exclusive in the search for e).
// Returns ss (tile-local), or -1 when the key cannot match (also for a ':'
// whose line starts before the tile: such lines are not evaluated here).  The
// key bytes are compared afterwards (fused_key_bytes), so that a tile can
// issue every candidate's compare at once.
template <class T>
GJ_HD int fused_key_locate(const T& t, int p, bool tile0, int klen) {
  const int k = p - 1;
  if (k < 0) return -1;
  int wi = k >> 6;
  uint64_t sw = t.S(wi) & fe_mask_upto(k & 63);
  uint64_t nw = t.NL(wi) & fe_mask_upto(k & 63);
  int start = -1;  // the line's start, once the walk has passed it
  int se;
  for (;;) {
    if (tile0 && wi == 0) sw &= ~1ull;  // e > 0 on the file's first line
    if (nw) {
      start = (wi << 6) + 63 - clz64(nw);
      sw &= ~fe_mask_upto(start & 63);
      if (!sw) return -1;
    }
    if (sw) {
      se = (wi << 6) + 63 - clz64(sw);
      break;
    }
    if (wi == 0) return -1;  // no in-string byte after the line's start (or the line began earlier)
    wi--;
    sw = t.S(wi);
    nw = t.NL(wi);
  }
  const int ss = se - klen;
  if (ss < 0) return -1;  // the key would start before the tile: its line began in an earlier tile
  if (start < 0) {
    // the line's start lies before se's word (no newline in [64 (se >> 6), p)):
    // ss >= start iff no newline in (ss, se).  (A ':' whose line began in an
    // earlier tile may pass; such lines are not evaluated here.)
    for (int w = ss >> 6; w <= (se >> 6); w++) {
      uint64_t m = t.NL(w);
      if (w == (ss >> 6)) m &= fe_mask_from(ss & 63) & ~(1ull << (ss & 63));
      if (w == (se >> 6)) m &= ~fe_mask_from(se & 63);
      if (m) return -1;
    }
  } else if (ss < start) {
    return -1;
  }
  if (!((t.Q(ss >> 6) >> (ss & 63)) & 1ull)) return -1;
  return ss;
}

// fused_key_locate for the ':' at bit b of tile word w from registers alone:
// the string, newline and raw-quote words of word w (S1, NL1, Q1) and of word
// w - 1 (S0, NL0, Q0; w >= 1).  Returns ss (tile-local) or -1 as
// fused_key_locate does, or -2 when the answer needs words before w - 1 (then
// fused_key_locate decides).
GJ_HD int fused_key_locate_near(uint64_t S1, uint64_t NL1, uint64_t Q1, uint64_t S0, uint64_t NL0, uint64_t Q0,
                                int w, int b, bool tile0, int klen) {
  // window positions: word w - 1 is [0, 64), word w is [64, 128); the ':' is at 64 + b
  int start = -1, se = -1;
  if (b > 0) {
    uint64_t sw = S1 & fe_mask_upto(b - 1);
    const uint64_t nw = NL1 & fe_mask_upto(b - 1);
    if (nw) {
      const int st = 63 - clz64(nw);
      start = 64 + st;
      sw &= ~fe_mask_upto(st);
      if (!sw) return -1;
    }
    if (sw) se = 64 + 63 - clz64(sw);
  }
  if (se < 0) {
    uint64_t sw = S0;
    if (tile0 && w == 1) sw &= ~1ull;  // e > 0 on the file's first line
    if (NL0) {
      const int st = 63 - clz64(NL0);
      start = st;
      sw &= ~fe_mask_upto(st);
      if (!sw) return -1;
    }
    if (!sw) return -2;
    se = 63 - clz64(sw);
  }
  const int ss = se - klen;
  if (ss < 0) return -2;
  if (start < 0) {
    // no newline in (ss, se): bits ss+1 .. se-1 of the window
    const int lo = ss + 1, hi = se - 1;
    if (lo <= hi) {
      uint64_t m0 = 0, m1 = 0;
      if (lo < 64) m0 = NL0 & fe_mask_from(lo) & (hi < 64 ? fe_mask_upto(hi) : ~0ull);
      if (hi >= 64) m1 = NL1 & (lo >= 64 ? fe_mask_from(lo - 64) : ~0ull) & fe_mask_upto(hi - 64);
      if (m0 | m1) return -1;
    }
  } else if (ss < start) {
    return -1;
  }
  const uint64_t q = ss < 64 ? (Q0 >> ss) : (Q1 >> (ss - 64));
  if (!(q & 1ull)) return -1;
  return ((w - 1) << 6) + ss;
}

// The usual layout, tested for every ':' of tile word w >= 1 at once: a key
// whose closing quote sits right before its ':' (an unescaped quote at p - 1,
// so the in-string byte p - 2 is the last one before p) has its opening quote
// at ss = p - 2 - klen.  c[l]: the word's ':' at level l (l < D); S1, Q1 and
// S0, Q0: string and raw-quote words of word w and w - 1.  fast: the ':' that
// passed (ss as above); slow: the ':' that need fused_key_locate's walk (no
// unescaped quote right before them, or a key too long for the two words).
// The rest cannot match.  Valid for keys without a newline byte: the key
// bytes, compared afterwards, then prove ss >= the line's start.
GJ_HD void fused_key_split(const uint64_t c[3], int D, const int klen[3], uint64_t S1, uint64_t S0, uint64_t Q1,
                           uint64_t Q0, uint64_t* fast, uint64_t* slow) {
  const uint64_t qb = (Q1 << 1) | (Q0 >> 63);  // bit p: a quote at p - 1
  const uint64_t sb = (S1 << 2) | (S0 >> 62);  // bit p: p - 2 is in a string
  const uint64_t usual = qb & sb;
  uint64_t f = 0, sl = 0;
  for (int l = 0; l < 3; l++) {
    if (l >= D) break;
    const int sh = klen[l] + 2;
    if (sh >= 64) {
      sl |= c[l];
      continue;
    }
    const uint64_t open = (Q1 << sh) | (Q0 >> (64 - sh));  // bit p: a quote at p - 2 - klen
    f |= c[l] & usual & open;
    sl |= c[l] & ~usual;
  }
  *fast = f;
  *slow = sl;
}

// Second half: the klen bytes after the '"' at file position q equal the key.
// Up to 8 bytes are loaded per round trip, all before any compare.
GJ_HD bool fused_key_bytes(const uint8_t* F, int64_t q, const uint8_t* key, int klen) {
  const uint8_t* f = F + q + 1;
  for (int i0 = 0; i0 < klen; i0 += 8) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = i0 + i < klen ? f[i0 + i] : 0;
    bool same = true;
    for (int i = 0; i < 8; i++) same &= !(i0 + i < klen) || b[i] == key[i0 + i];
    if (!same) return false;
  }
  return true;
}

// The whole key test of the ':' at tile-local position p (pos0: the tile's
// first byte in the file F).
template <class T>
GJ_HD bool fused_key_test(const T& t, int p, bool tile0, const uint8_t* F, int64_t pos0, const uint8_t* key,
                          int klen) {
  const int ss = fused_key_locate(t, p, tile0, klen);
  return ss >= 0 && fused_key_bytes(F, pos0 + ss, key, klen);
}

// First tile-local position in [from, to) whose bit is set in words x.w(i);
// x.sum(s) has bit k set iff word 64 s + k is not zero.  -1 if none.
template <class X>
GJ_HD int fused_first_bit(const X& x, int from, int to) {
  if (from >= to) return -1;
  const int wf = from >> 6, wt = (to - 1) >> 6;
  // the first word and the summary after it are read together
  uint64_t v = x.w(wf) & fe_mask_from(from & 63);
  const int w1 = wf + 1;
  const uint64_t sm1 = w1 <= wt ? x.sum(w1 >> 6) : 0;
  if (wf == wt) v &= fe_mask_upto((to - 1) & 63);
  if (v) return (wf << 6) + ctz64(v);
  int w = w1;
  bool first = true;
  while (w <= wt) {
    const int s = w >> 6;
    const uint64_t sm = (first ? sm1 : x.sum(s)) & fe_mask_from(w & 63);
    first = false;
    if (!sm) {
      w = (s + 1) << 6;
      continue;
    }
    const int w2 = (s << 6) + ctz64(sm);
    if (w2 > wt) return -1;
    uint64_t u = x.w(w2);
    if (w2 == wt) u &= fe_mask_upto((to - 1) & 63);
    if (u) return (w2 << 6) + ctz64(u);
    w = w2 + 1;
  }
  return -1;
}

// Views over a tile for fused_first_bit: level l's words (and its summary), or
// level l's matching ':' (match words & level words, and their summary).
template <class T>
struct FeLevel {
  const T& t;
  int l;
  GJ_FE_M uint64_t w(int i) const { return t.LB(l, i); }
  GJ_FE_M uint64_t sum(int s) const { return t.SL(l, s); }
};
template <class T>
struct FeMatch {
  const T& t;
  int l;
  GJ_FE_M uint64_t w(int i) const { return t.M(i) & t.LB(l, i); }
  GJ_FE_M uint64_t sum(int s) const { return t.SM(l, s); }
};

// Spans of one member chain of depth D on the line [a, b) (tile-local; the
// line starts at a, which is its first byte, and ends before b).  t gives the
// leveled words LB(l, w), the match words M(w) (bit set at every ':' at level
// l < D whose key test against k_l passed) and their summaries.  out[0..1]
// get the file spans before the result's leading blanks are skipped (the byte
// after the last key's ':'; fused_skip_blanks finishes it), or -1, -1.
template <class T>
GJ_HD void fused_chain_line(const T& t, int D, int a, int b, int64_t pos0, int64_t out[2]) {
  out[0] = -1;
  out[1] = -1;
  int c = -1;
  for (int l = 0; l < D; l++) {
    if (b < 0) return;  // the level's end was never found: the walk stops
    c = fused_first_bit(FeMatch<T>{t, l}, a, b);
    if (c < 0 || c + 1 >= b) return;
    const int nb = fused_first_bit(FeLevel<T>{t, l}, c + 2, b);
    a = c + 1;
    b = nb;
  }
  out[0] = pos0 + c + 1;
  out[1] = b < 0 ? -1 : pos0 + b;
}

// The STORE's blank skip: the first byte at or after j (j < 0: no result) that
// is not ' ', '\t', '\r' or '\n', or N.
GJ_HD int64_t fused_skip_blanks(const uint8_t* F, int64_t N, int64_t j) {
  if (j < 0) return j;
  while (j < N && (F[j] == ' ' || F[j] == '\t' || F[j] == '\r' || F[j] == '\n')) j++;
  return j;
}

}  // namespace gj
