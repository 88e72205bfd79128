// query_kernel.hip -- per-line JSONPath IR interpreter for gfx950.
//
// Semantics are those of the reference's find_value kernel
// (K/find_value.cu:21-263, == J/sequential/Sequential.java:146-337; K/ =
// gpjson/src/main/resources/com/koenv/gpjson/kernels/, J/ = gpjson/src/main/
// java/com/koenv/gpjson/): a line is evaluated with a byte cursor j that only
// moves forward while the interpreter waits for a structural event at the
// current level (a ':' whose key matches, or the n-th ','); IR ops run between
// events.
//
// Mapping: a group of G lanes (G = 1, 4, 16 or 64) evaluates one line; the
// IR, the cursor and the per-level state are uniform within the group.  A wait
// scans the current level's bitmap G words (G*64 bytes of text) per step: each
// lane tests the candidate positions in its word and the earliest event wins
// by a group ballot, so the first match in byte order is the one taken, as in
// the sequential loop.  Level ends (MOVE_DOWN) are found the same way.  Small
// G keeps many lines in flight (latency-bound short lines); G = 64 suits long
// lines.  With one query and a lane per line, a lane starts its next line as
// soon as its current one ends (a per-wavefront pool of claimed lines);
// otherwise groups claim lines_per_wave lines at a time from an atomic
// counter.  Per-level state lives in LDS: levels+2 slots per group.
#include <hip/hip_runtime.h>
#include "index_common.h"
#include "query_common.h"

namespace gj {

namespace {

__device__ __forceinline__ bool is_blank(uint8_t c) { return c == '\n' || c == '\r' || c == '\t' || c == ' '; }

// Loads of the index arrays and the input through global-segment pointers.
// The jobs' pointers come from a table in device memory and stay generic, and
// generic loads compile to FLAT instructions, which also count against the
// LDS wait counter: every read of the level slots in LDS would wait for them.
typedef __attribute__((address_space(1))) const uint64_t g_cu64_t;
typedef __attribute__((address_space(1))) const int64_t g_ci64_t;
typedef __attribute__((address_space(1))) const uint8_t g_cu8_t;
typedef unsigned int q_u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const q_u32x4_t g_cu32x4_t;
__device__ __forceinline__ uint64_t gld64(const uint64_t* p) { return *(const g_cu64_t*)p; }
__device__ __forceinline__ int64_t gldi64(const int64_t* p) { return *(const g_ci64_t*)p; }
__device__ __forceinline__ uint8_t gld8(const uint8_t* p) { return *(const g_cu8_t*)p; }
__device__ __forceinline__ uint4 gld16(const uint8_t* p) {
  const q_u32x4_t v = *(const g_cu32x4_t*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int ctz64(unsigned long long x) { return __builtin_ctzll(x); }
// Bits [lo_bit, 63] of a word.
__device__ __forceinline__ uint64_t mask_from(int lo_bit) { return ~0ull << lo_bit; }
// Bits [0, hi_bit] of a word.
__device__ __forceinline__ uint64_t mask_upto(int hi_bit) { return hi_bit >= 63 ? ~0ull : ((2ull << hi_bit) - 1ull); }

__device__ __forceinline__ unsigned long long shfl64_lane(unsigned long long v, int src) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl(lo, src);
  hi = __shfl(hi, src);
  return ((unsigned long long)hi << 32) | lo;
}

// A launch's work items [0, total) lie in 2^lg ranges (at most
// kQuerySegments), each with a counter of its own, kQueryCounterStride words
// from the next (atomics on one cache line take turns).  A wavefront claims up
// to n consecutive items of range seg; once seg is used up it moves on to the
// lowest range not yet seen used up (live).  Wavefront-uniform; false once
// every range is used up.  (Reading every counter at once when a range runs
// out, instead of trying the ranges one by one, ends a launch sooner but
// costs the interpreter spilled registers, and is slower overall.)
__device__ __forceinline__ bool claim_items(unsigned long long* counter, int64_t total, int lg, int64_t n, int lane,
                                            int& seg, unsigned& live, int64_t& lo, int64_t& hi) {
  for (;;) {
    if (live >> seg & 1u) {
      unsigned long long c = 0;
      if (lane == 0) c = atomicAdd(counter + seg * kQueryCounterStride, (unsigned long long)n);
      const int64_t f = ((total * seg) >> lg) + (int64_t)shfl64_lane(c, 0);
      const int64_t e = (total * (seg + 1)) >> lg;
      if (f < e) {
        lo = f;
        hi = min(f + n, e);
        return true;
      }
      live &= ~(1u << seg);
    }
    if (!live) return false;
    seg = __builtin_ctz(live);
  }
}

// Lanes [base, base+G) of the wavefront work on one line.
template <int G>
struct Group {
  int lane, r, base;
  __device__ Group() {
    lane = __lane_id();
    r = lane % G;
    base = lane - r;
  }
  __device__ __forceinline__ unsigned long long ballot(bool p) const {
    if (G == 1) return p ? 1ull : 0ull;
    unsigned long long b = __ballot(p);
    if (G == 64) return b;
    return (b >> base) & ((1ull << G) - 1ull);
  }
  __device__ __forceinline__ int64_t shfl_i64(int64_t v, int src) const {
    if (G == 1) return v;
    return (int64_t)shfl64_lane((unsigned long long)v, base + src);
  }
  __device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) const {
    if (G == 1) return v;
    return shfl64_lane(v, base + src);
  }
  __device__ __forceinline__ int shfl(int v, int src) const {
    if (G == 1) return v;
    return __shfl(v, base + src);
  }
  // inclusive prefix sum over the group
  __device__ __forceinline__ int scan_add(int v) const {
    if (G == 1) return v;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      int y = __shfl_up(v, off, G);
      if (r >= off) v += y;
    }
    return v;
  }
};

// Where the leveled and string words come from: device memory.
struct GlobalWords {
  const uint64_t* LB;   // levels 1 .. L-1, stride Wp
  int64_t Wp;
  const uint64_t* rec;  // records: string, colon, raw-quote and level-0 words
  const uint64_t* LS;   // level summaries (may be null: no skipping)
  int64_t Ws;
  __device__ __forceinline__ uint64_t lb(int lvl, int64_t wi) const {
    return lvl == 0 ? gld64(rec + wi * kRecWords + kRecLevel0) : gld64(LB + (int64_t)(lvl - 1) * Wp + wi);
  }
  __device__ __forceinline__ uint64_t s(int64_t wi) const { return gld64(rec + wi * kRecWords); }
  __device__ __forceinline__ uint64_t colon(int64_t wi) const { return gld64(rec + wi * kRecWords + 1); }
  __device__ __forceinline__ uint64_t quote(int64_t wi) const { return gld64(rec + wi * kRecWords + 2); }
  __device__ __forceinline__ bool has_summary() const { return LS != nullptr; }
  __device__ __forceinline__ uint64_t ls(int lvl, int64_t si) const { return gld64(LS + (int64_t)lvl * Ws + si); }
};

// Summary words of level lvl for the blocks of 64 words that hold word nb and
// the one after, loaded together with a scan step's words (so they cost no
// extra round trip); blocks past wend read as empty.
template <class Words>
__device__ __forceinline__ void load_summary(const Words& W, int lvl, int64_t nb, int64_t wend, uint64_t& sm0,
                                             uint64_t& sm1) {
  sm0 = 0;
  sm1 = 0;
  if (!W.has_summary()) return;
  if (nb <= wend) sm0 = W.ls(lvl, nb >> 6);
  if ((nb >> 6) + 1 <= (wend >> 6)) sm1 = W.ls(lvl, (nb >> 6) + 1);
}

// The next word at or after nb whose level word may have a bit: words the
// summaries show empty are skipped (up to the end of the block after nb's).
template <class Words>
__device__ __forceinline__ int64_t skip_empty(const Words& W, int64_t nb, uint64_t sm0, uint64_t sm1) {
  if (!W.has_summary()) return nb;
  const uint64_t rest = sm0 & mask_from((int)(nb & 63));
  if (rest) return (nb & ~63ll) + ctz64(rest);
  const int64_t next = (nb | 63) + 1;
  return sm1 ? next + ctz64(sm1) : next + 64;
}

// Word wi of level lvl restricted to positions [from, to).
template <class Words>
__device__ __forceinline__ uint64_t window_word(const Words& W, int lvl, int64_t wi, int64_t wbeg, int64_t wend,
                                                int64_t from, int64_t to) {
  if (wi > wend) return 0;
  uint64_t v = W.lb(lvl, wi);
  if (wi == wbeg) v &= mask_from((int)(from & 63));
  if (wi == wend) v &= mask_upto((int)((to - 1) & 63));
  return v;
}

// Words a lane loads per step in the scans below: independent loads issued
// together, so one memory round trip covers that many words of a line (key
// search: level, colon, string and raw-quote words; level-bit search: level
// words only).  Measured on 4 GiB of tweets ($.user.lang / $.id): 1 word
// 3.35 / 0.90 ms, 2 words 2.71 / 0.91 ms, 4 words 2.86 / 1.06 ms, 8 words
// 3.29 / 1.35 ms.
#ifndef GJ_KEY_WORDS
#define GJ_KEY_WORDS 2
#endif
#ifndef GJ_BIT_WORDS
#define GJ_BIT_WORDS 2
#endif
constexpr int kKeyWords = GJ_KEY_WORDS;
constexpr int kBitWords = GJ_BIT_WORDS;

// First set bit of level lvl in [from, to), or -1.
template <int G, class Words>
__device__ int64_t group_first_bit(const Group<G>& g, const Words& W, int lvl, int64_t from, int64_t to) {
  if (from >= to) return -1;
  const int64_t wbeg = from >> 6, wend = (to - 1) >> 6;
  for (int64_t b0 = wbeg; b0 <= wend;) {
    const int64_t w0 = b0 + (int64_t)g.r * kBitWords;
    const int64_t nb = b0 + (int64_t)G * kBitWords;
    uint64_t v[kBitWords], sm0, sm1;
#pragma unroll
    for (int u = 0; u < kBitWords; u++) v[u] = window_word(W, lvl, w0 + u, wbeg, wend, from, to);
    load_summary(W, lvl, nb, wend, sm0, sm1);
    int64_t found = -1;
#pragma unroll
    for (int u = kBitWords - 1; u >= 0; u--)
      if (v[u]) found = ((w0 + u) << 6) + ctz64(v[u]);
    const unsigned long long m = g.ballot(found >= 0);
    if (m) return g.shfl_i64(found, ctz64(m));
    b0 = skip_empty(W, nb, sm0, sm1);
  }
  return -1;
}

// Key test at a ':' at position p (K/find_value.cu:227-250): e = the last
// in-string byte in (start, p); the key must start with a '"' at ss = e - klen
// (ss >= start) and the klen bytes after it must equal the key.  The quote test
// reads the raw-quote bitmap, so input bytes are read only for candidates
// whose length already fits.  sw / qw: string and raw-quote words of p's word.
template <class Words>
__device__ bool key_matches(const uint8_t* F, const uint64_t* rec, const Words& W, int64_t start, int64_t p,
                            uint64_t sw_p, uint64_t qw_p, const uint8_t* key, int klen) {
  const int64_t k = p - 1;
  if (k <= start) return false;
  const int64_t wp = p >> 6;
  int64_t wi = k >> 6;
  const int64_t wstart = start >> 6;
  uint64_t sw = (wi == wp ? sw_p : W.s(wi)) & mask_upto((int)(k & 63));
  int64_t se;
  for (;;) {
    if (wi == wstart) sw &= ~mask_upto((int)(start & 63));
    if (sw) { se = (wi << 6) + 63 - __clzll(sw); break; }
    if (wi <= wstart) return false;
    wi--;
    sw = W.s(wi);
  }
  const int64_t ss = se - klen;
  if (ss < start) return false;
  const int64_t qi = ss >> 6;
  const uint64_t qw = qi == wp ? qw_p : W.quote(qi);
  if (!((qw >> (ss & 63)) & 1ull)) return false;
  // compare 8 bytes per round trip (the loads of a chunk are independent)
  for (int i0 = 0; i0 < klen; i0 += 8) {
    uint8_t f[8];
#pragma unroll
    for (int i = 0; i < 8; i++) f[i] = i0 + i < klen ? gld8(F + ss + 1 + i0 + i) : 0;
    bool same = true;
#pragma unroll
    for (int i = 0; i < 8; i++) same &= !(i0 + i < klen) || f[i] == key[i0 + i];
    if (!same) return false;
  }
  return true;
}

// First structural ':' in [from, to) at this level whose key matches, or -1.
// Candidates are the level's bits that are ':' bytes (colon bitmap), so no
// input byte is read to find them.
template <int G, class Words>
__device__ int64_t group_find_key(const Group<G>& g, const Words& W, const uint64_t* rec, int lvl,
                                  const uint8_t* F, int64_t start, int64_t from, int64_t to, const uint8_t* key,
                                  int klen) {
  if (from >= to) return -1;
  const int64_t wbeg = from >> 6, wend = (to - 1) >> 6;
  for (int64_t b0 = wbeg; b0 <= wend;) {
    const int64_t w0 = b0 + (int64_t)g.r * kKeyWords;
    const int64_t nb = b0 + (int64_t)G * kKeyWords;
    uint64_t v[kKeyWords], sw[kKeyWords], qw[kKeyWords], sm0, sm1;
#pragma unroll
    for (int u = 0; u < kKeyWords; u++) {
      const int64_t wi = w0 + u;
      v[u] = 0;
      sw[u] = 0;
      qw[u] = 0;
      if (wi <= wend) {
        v[u] = window_word(W, lvl, wi, wbeg, wend, from, to) & W.colon(wi);
        sw[u] = W.s(wi);
        qw[u] = W.quote(wi);
      }
    }
    load_summary(W, lvl, nb, wend, sm0, sm1);
    int64_t found = -1;
#pragma unroll
    for (int u = 0; u < kKeyWords; u++) {
      uint64_t t = found < 0 ? v[u] : 0;
      while (t) {
        const int64_t p = ((w0 + u) << 6) + ctz64(t);
        if (key_matches(F, rec, W, start, p, sw[u], qw[u], key, klen)) { found = p; break; }
        t &= t - 1;
      }
    }
    const unsigned long long m = g.ballot(found >= 0);
    if (m) return g.shfl_i64(found, ctz64(m));
    b0 = skip_empty(W, nb, sm0, sm1);
  }
  return -1;
}

// Bit per byte of the 64-byte word wi of F that equals ',' (byte-parallel
// compare of 16-byte loads; the file's last partial word byte by byte).
__device__ __forceinline__ uint64_t comma_word(const uint8_t* F, int64_t n, int64_t wi) {
  const int64_t off = wi << 6;
  uint32_t d[16];
  if (off + 64 <= n) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 x = gld16(F + off + 16 * q);
      d[4 * q] = x.x;
      d[4 * q + 1] = x.y;
      d[4 * q + 2] = x.z;
      d[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) d[k] = 0;
    for (int b = 0; off + b < n && b < 64; b++) d[b >> 2] |= (uint32_t)gld8(F + off + b) << (8 * (b & 3));
  }
  uint64_t m = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t t = d[k] ^ 0x2C2C2C2Cu;
    // bit 7 of a byte: that byte of t is zero (no carry crosses bytes)
    const uint32_t z = ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u;
    // gather the four bit-7 flags into bits 21..24 (the partial products do
    // not overlap, so the multiply adds no carries)
    const uint32_t four = (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
    m |= (uint64_t)four << (4 * k);
  }
  return m;
}

// Walks structural ',' in [from, to), counting from *count; returns the
// position of the comma at which the count becomes want, or -1 (count advanced
// by every comma in range).  A lane tests kBitWords words per step: their
// level words and text are loaded together, and the text is compared whole.
template <int G, class Words>
__device__ int64_t group_find_comma(const Group<G>& g, const Words& W, int lvl, const uint8_t* F, int64_t n,
                                    int64_t from, int64_t to, int* count, int want) {
  if (from >= to) return -1;
  const int64_t wbeg = from >> 6, wend = (to - 1) >> 6;
  int cnt0 = *count;
  for (int64_t b0 = wbeg; b0 <= wend;) {
    const int64_t w0 = b0 + (int64_t)g.r * kBitWords;
    const int64_t nb = b0 + (int64_t)G * kBitWords;
    uint64_t v[kBitWords], fm[kBitWords], sm0, sm1;
#pragma unroll
    for (int u = 0; u < kBitWords; u++) {
      v[u] = window_word(W, lvl, w0 + u, wbeg, wend, from, to);
      fm[u] = w0 + u <= wend ? comma_word(F, n, w0 + u) : 0;
    }
    load_summary(W, lvl, nb, wend, sm0, sm1);
    int c = 0;
#pragma unroll
    for (int u = 0; u < kBitWords; u++) {
      v[u] &= fm[u];
      c += __popcll(v[u]);
    }
    const int incl = g.scan_add(c);
    const int total = g.shfl(incl, G - 1);
    const int need = want - cnt0;  // 1-based rank of the event comma in this window
    const bool hit = need >= 1 && need <= total && incl >= need && (incl - c) < need;
    const unsigned long long m = g.ballot(hit);
    if (m) {
      const int l = ctz64(m);
      int64_t pos = -1;
      if (g.r == l) {
        int k = need - (incl - c);  // 1-based rank within this lane's words
#pragma unroll
        for (int u = 0; u < kBitWords; u++) {
          const int cu = __popcll(v[u]);
          if (pos < 0 && k <= cu) {
            uint64_t cm = v[u];
            for (int t = 1; t < k; t++) cm &= cm - 1;
            pos = ((w0 + u) << 6) + ctz64(cm);
          }
          k -= cu;
        }
      }
      *count = want;
      return g.shfl_i64(pos, l);
    }
    cnt0 += total;
    b0 = skip_empty(W, nb, sm0, sm1);
  }
  *count = cnt0;
  return -1;
}

__device__ __forceinline__ int read_varint(const uint8_t* ir, int64_t ir_len, int64_t* pc) {
  int v = 0, shift = 0;
  while (*pc < ir_len) {
    const uint8_t b = ir[(*pc)++];
    v |= (int)(b & 0x7F) << shift;
    if (!(b & 0x80)) break;
    shift += 7;
    if (shift > 28) break;
  }
  return v;
}

// Interpreter state of one line (uniform within its group).
struct LineState {
  int64_t start;           // first byte of the line
  int64_t j;               // byte cursor
  int64_t pc;              // IR position
  int64_t* out;            // the line's R span slots
  const uint8_t* want;     // key of the pending MOVE_TO_KEY
  int lvl, r, op, want_len, want_idx;
  bool exec;               // run IR ops before the next wait
};

// Starts line i of job J with group g.  le/ci: this group's nslots level
// slots in LDS, written by the group's first lane and read by all of them (LDS
// operations of one wavefront complete in program order).
template <int G>
__device__ __forceinline__ void line_begin(const Group<G>& g, const QueryJob& J, int64_t lines, int64_t i,
                                           int64_t* out, int64_t* le, int* ci, int nslots, LineState& st) {
  const bool lead = g.r == 0;
  const int64_t start = gldi64(J.NL + i);
  const int64_t end = (i + 1 < lines) ? gldi64(J.NL + i + 1) : J.N;
  if (lead)
    for (int k = 0; k < J.R * 2; k++) out[k] = -1;
  for (int k = g.r; k < nslots; k += G) { le[k] = -1; ci[k] = -1; }
  if (lead) le[0] = end;
  st.start = start;
  st.j = start;
  st.pc = 0;
  st.out = out;
  st.want = nullptr;
  st.lvl = 0;
  st.r = 0;
  st.op = kEnd;
  st.want_len = 0;
  st.want_idx = 0;
  st.exec = true;
}

// One round of a line: the IR ops up to the next wait, then the wait (a scan
// of the current level for the event).  Returns false once the line is done.
template <int G, class Words>
__device__ bool line_step(const Group<G>& g, const QueryJob& J, const Words& W, const uint8_t* ir, LineState& st,
                          int64_t* le, int* ci, int nslots) {
  const bool lead = g.r == 0;
  const uint8_t* F = J.F;
  const int64_t n = J.N;
  int lvl = st.lvl;
  int64_t j = st.j;
  int64_t pc = st.pc;
  int r = st.r;
  int op = st.op;
  int want_len = st.want_len, want_idx = st.want_idx;
  const uint8_t* want = st.want;
  bool exec = st.exec;
  int64_t* out = st.out;

  if (lvl < 0) return false;
  if (!(j < le[lvl] && j < n)) return false;
  while (exec) {
    if (pc >= J.ir_len) return false;
    op = ir[pc++];
    switch (op) {
      case kEnd:
        return false;
      case kStoreResult: {
        while (j < n && is_blank(gld8(F + j))) j++;
        if (r < J.R && lead) {
          out[r * 2] = j;
          out[r * 2 + 1] = le[lvl];
        }
        r++;
        break;
      }
      case kMoveUp:
        j = le[lvl];
        if (lead) le[lvl] = -1;
        lvl--;
        // a level whose end was never found (-1) ends the line: the walk
        // would otherwise continue from before the line's first byte
        if (lvl < 0 || j < 0) return false;
        break;
      case kMoveDown: {
        lvl++;
        if (lvl >= nslots) return false;
        if (le[lvl] == -1) {
          int64_t e = -1;
          if (lvl - 1 < J.L) e = group_first_bit(g, W, lvl - 1, j + 1, le[lvl - 1]);
          if (lead) le[lvl] = e;
        }
        break;
      }
      case kMoveToKey:
        want_len = read_varint(ir, J.ir_len, &pc);
        want = ir + pc;
        pc += want_len;
        exec = false;
        break;
      case kMoveToIndex: {
        want_idx = read_varint(ir, J.ir_len, &pc);
        if (ci[lvl] == -1) {
          while (j < n && gld8(F + j) != '[') {
            if (!is_blank(gld8(F + j))) return false;  // not an array: the line has no further results
            j++;
          }
          if (j >= n) return false;
          if (lead) ci[lvl] = 0;
        }
        if (want_idx == 0) { j++; exec = true; } else { exec = false; }
        break;
      }
      case kExpressionStringEquals: {
        want_len = read_varint(ir, J.ir_len, &pc);
        want = ir + pc;
        pc += want_len;
        while (j < n && is_blank(gld8(F + j))) j++;
        if (le[lvl] - j != want_len) return false;
        for (int k = 0; k < want_len; k++)
          if (want[k] != gld8(F + j + k)) return false;
        break;
      }
      default:
        return false;
    }
  }
  if (!(j < n)) return false;
  // wait for the next event at this level
  const int64_t lim = le[lvl];
  const int64_t to = lim < n ? lim : n;
  int64_t p = -1;
  if (lvl < J.L) {
    if (op == kMoveToKey) {
      p = group_find_key(g, W, J.rec, lvl, F, st.start, j, to, want, want_len);
    } else if (op == kMoveToIndex) {
      int cnt = ci[lvl];
      p = group_find_comma(g, W, lvl, F, n, j, to, &cnt, want_idx);
      if (lead) ci[lvl] = cnt;
    }
  }
  if (p < 0) return false;
  st.lvl = lvl;
  st.j = p + 1;
  st.pc = pc;
  st.r = r;
  st.op = op;
  st.want_len = want_len;
  st.want_idx = want_idx;
  st.want = want;
  st.exec = true;
  return true;
}

// Evaluates line i of job J with group g.
template <int G, class Words>
__device__ void eval_line(const Group<G>& g, const QueryJob& J, const Words& W, const uint8_t* ir, int64_t lines,
                          int64_t i, int64_t* out, int64_t* le, int* ci, int nslots) {
  LineState st;
  line_begin(g, J, lines, i, out, le, ci, nslots, st);
  while (line_step(g, J, W, ir, st, le, ci, nslots)) {
  }
}

}  // namespace

// Registers held to 128 so that 4 wavefronts fit per SIMD: the walk is a chain
// of dependent loads, and the fourth wavefront hides more of it than the few
// spilled registers cost (4 GiB of tweets, four one-query paths: 3.64 against
// 3.96 ms at 3 wavefronts per SIMD).
template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK, 4) void query_kernel(QueryLaunch launch, int nslots) {
  constexpr int NG = BLOCK / G;
  extern __shared__ int64_t q_lds[];
  __shared__ uint8_t s_ir[kIrLdsBytes];
  const Group<G> g;
  const int grp = threadIdx.x / G;
  int64_t* le = q_lds + (int64_t)grp * nslots;
  int* ci = reinterpret_cast<int*>(q_lds + (int64_t)NG * nslots) + (int64_t)grp * nslots;
  // the IR of every job, read from LDS for the rest of the kernel
  const bool ir_lds = launch.ir_base && launch.ir_bytes <= kIrLdsBytes;
  if (ir_lds)
    for (int k = threadIdx.x; k < launch.ir_bytes; k += BLOCK) s_ir[k] = launch.ir_base[k];
  __syncthreads();
  auto ir_of = [&](const QueryJob& J) { return ir_lds ? s_ir + (J.ir - launch.ir_base) : J.ir; };
  const int64_t common = launch.lines_dev ? *launch.lines_dev : 0;
  int64_t total = launch.lines_dev ? common * launch.njobs
                                   : (launch.total_dev ? *launch.total_dev : launch.total_items);
  if (launch.line_list) total = min((int64_t)*launch.list_count, launch.list_cap);
  // Several lines per group amortize a claim over short lines, but with few
  // (long) lines they serialize work one group could share: halve the claim
  // until every group of the grid gets at least one full claim.
  int lpw = launch.lines_per_wave > 0 ? launch.lines_per_wave : 1;
  const int64_t groups = (int64_t)gridDim.x * NG;
  while (lpw > 1 && total < groups * lpw) lpw >>= 1;
  // One atomic per wavefront claims lpw lines for each of its 64/G groups: the
  // wavefront runs until its slowest lane is done anyway, and the counter sees
  // 64/G times fewer atomics than with per-group claims.
  constexpr int kGroupsPerWave = 64 / G;
  const int lane = __lane_id();
  // Ranges of work items: 2^lg of them, one per kSegItems items or more, at
  // most kQuerySegments (few items make few claims, which one counter
  // serves); the wavefront starts in range wavefront % 2^lg.
  constexpr int64_t kSegItems = 65536;
  int lg = 0;
  while ((2 << lg) <= kQuerySegments && total >= (kSegItems << (lg + 1))) lg++;
  int seg = (int)((blockIdx.x * (BLOCK / 64) + threadIdx.x / 64) & ((1 << lg) - 1));
  unsigned live = (1u << (1 << lg)) - 1u;
  // defers a long line to the wavefront-per-line launch (false: list full)
  auto defer = [&](const QueryJob& J, int64_t lines, int64_t i, int64_t it) -> bool {
    if (!launch.long_items || G == 64) return false;
    const int64_t start = J.NL[i], end = (i + 1 < lines) ? J.NL[i + 1] : J.N;
    if (end - start < launch.long_min) return false;
    int64_t slot = 0;
    if (g.r == 0) slot = (int64_t)atomicAdd(&launch.long_count[0], 1ull);
    slot = g.shfl_i64(slot, 0);
    if (slot >= launch.long_cap) return false;
    if (g.r == 0) launch.long_items[slot] = it;
    return true;
  };
  if (G == 1 && launch.lines_dev && launch.njobs == 1) {
    // One query, a lane per line: a lane starts its next line as soon as its
    // current one ends, so the wavefront's lanes stay busy until the lines run
    // out instead of waiting at the end of every claim for the slowest one.
    // Every round each lane with a line runs its IR ops up to the next wait
    // and that wait; lanes without one take the next items of the wavefront's
    // pool, which claims 64 items when it runs short.
    const QueryJob J = launch.jobs[0];
    const GlobalWords W{J.LB, J.Wp, J.rec, J.LS, J.Ws};
    const uint8_t* ir = ir_of(J);
    const unsigned long long below = (1ull << lane) - 1ull;
    int64_t pool = 0, pool_end = 0;  // items claimed, not yet handed out (wavefront-uniform)
    bool has = false, done = false;
    LineState st;
    for (;;) {
      const unsigned long long need = __ballot(!has && !done);
      if (need) {
        const int k = __popcll(need);
        const int64_t avail = pool_end - pool;
        int64_t fresh = 0, fresh_end = 0;  // a new claim [fresh, fresh_end)
        const bool claimed = avail < k && claim_items(launch.counter, total, lg, 64, lane, seg, live, fresh, fresh_end);
        if (!has && !done) {
          const int rank = __popcll(need & below);
          if (rank < avail || (claimed && fresh + (rank - avail) < fresh_end)) {
            const int64_t item = rank < avail ? pool + rank : fresh + (rank - avail);
            const int64_t i = launch.line_list ? launch.line_list[item] : item;
            if (i >= 0 && i < common && !defer(J, common, i, item)) {
              line_begin(g, J, common, i, J.out + i * (int64_t)J.R * 2, le, ci, nslots, st);
              has = true;
            }
          } else if (avail < k && !claimed) {
            done = true;  // every range is used up
          }
        }
        if (avail < k) {
          pool = min(fresh + (k - avail), fresh_end);
          pool_end = fresh_end;
        } else {
          pool += k;
        }
      }
      if (!__ballot(has || !done)) break;  // every lane is past the last item
      if (has) has = line_step(g, J, W, ir, st, le, ci, nslots);
    }
    return;
  }
  for (;;) {
    int64_t w0 = 0, wend = 0;  // the wavefront's claim [w0, wend)
    if (!claim_items(launch.counter, total, lg, (int64_t)kGroupsPerWave * lpw, lane, seg, live, w0, wend)) break;
    const int64_t g0 = w0 + (int64_t)(lane / G) * lpw;
    if (g0 >= wend) continue;
    if (G > 1 && launch.lines_dev && launch.njobs == 1) {  // (G = 1: the refill above)
      // one query: its job once, in scalar registers, not reloaded per line
      const QueryJob J = launch.jobs[0];
      const GlobalWords W{J.LB, J.Wp, J.rec, J.LS, J.Ws};
      const uint8_t* ir = ir_of(J);
      for (int64_t it = g0; it < g0 + lpw && it < wend; it++) {
        const int64_t i = launch.line_list ? launch.line_list[it] : it;
        if (i < 0 || i >= common) continue;
        if (defer(J, common, i, it)) continue;
        eval_line(g, J, W, ir, common, i, J.out + i * (int64_t)J.R * 2, le, ci, nslots);
      }
      continue;
    }
    // several jobs: each lane keeps its current job in registers and reloads
    // it only when its items cross into another job (items of a claim are
    // consecutive), so a line's walk does not start with a load of its job.
    // Work item -> (job, line): by the launch's shared line count (one index,
    // several queries); or by the jobs' first items (several files), where a
    // file's queries may take turns over blocks of its lines (its items are
    // [base, base + nq * lines), its jobs j0 .. j0 + nq - 1).  One loop and
    // one call of the line walk serve all three.
    const bool by_lines = launch.lines_dev != nullptr;
    auto split = [&](int64_t it, int64_t& q, int64_t& i) {
      if (launch.qblock > 0) {
        split_item(it, common, launch.njobs, launch.qblock, &q, &i);
      } else {
        q = it / common;
        i = it - q * common;
      }
    };
    int64_t jcur = 0;
    if (by_lines) {
      int64_t i0;
      split((int64_t)g0, jcur, i0);
    } else {
      int lo = 0, hi = (int)launch.njobs - 1;
      while (lo < hi) {  // job of the first item
        const int mid = (lo + hi + 1) >> 1;
        if (launch.jobs[mid].item0 <= (int64_t)g0) lo = mid; else hi = mid - 1;
      }
      jcur = lo;
    }
    QueryJob J = launch.jobs[jcur];
    const bool turns = !by_lines && launch.qblock > 0 && J.nqf > 0;
    int64_t j0 = 0, nq = 0, flines = 0, base = 0, next0 = INT64_MAX;
    if (turns) {
      j0 = jcur - J.qf;
      nq = J.nqf;
      flines = J.lines;
      base = launch.jobs[j0].item0;
    } else if (!by_lines && jcur + 1 < launch.njobs) {
      next0 = launch.jobs[jcur + 1].item0;  // first item of the next job
    }
    for (int64_t it = g0; it < g0 + lpw && it < wend; it++) {
      int64_t jq, i = 0, lines = 0, outrow = 0;
      if (by_lines) {
        split(it, jq, i);
        lines = common;
        outrow = jq * common + i;
      } else if (turns) {
        while (it >= base + nq * flines && j0 + nq < launch.njobs) {  // into the next file
          j0 += nq;
          const QueryJob& F0 = launch.jobs[j0];
          nq = F0.nqf;
          flines = F0.lines;
          base = F0.item0;
        }
        int64_t q;
        split_item(it - base, flines, nq, launch.qblock, &q, &i);
        jq = j0 + q;
        lines = flines;
        outrow = i;
      } else {
        jq = jcur;
        if (it >= next0) {
          while (jq + 1 < launch.njobs && launch.jobs[jq + 1].item0 <= it) jq++;
          next0 = jq + 1 < launch.njobs ? launch.jobs[jq + 1].item0 : INT64_MAX;
        }
      }
      if (jq != jcur) {
        jcur = jq;
        J = launch.jobs[jq];
      }
      if (!by_lines && !turns) {
        i = it - J.item0;
        lines = J.lines;
        outrow = i;
      }
      if (defer(J, lines, i, it)) continue;
      eval_line(g, J, GlobalWords{J.LB, J.Wp, J.rec, J.LS, J.Ws}, ir_of(J), lines, i,
                J.out + outrow * (int64_t)J.R * 2, le, ci, nslots);
    }
  }
}

// Second launch of line mode: the long lines the first launch deferred, one
// wavefront per line (64 words per scan step).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void query_long_kernel(QueryLaunch launch, int nslots) {
  constexpr int G = 64;
  extern __shared__ int64_t q_lds[];
  __shared__ uint8_t s_ir[kIrLdsBytes];
  const Group<G> g;
  const int grp = threadIdx.x / G;
  constexpr int NG = BLOCK / G;
  int64_t* le = q_lds + (int64_t)grp * nslots;
  int* ci = reinterpret_cast<int*>(q_lds + (int64_t)NG * nslots) + (int64_t)grp * nslots;
  const bool ir_lds = launch.ir_base && launch.ir_bytes <= kIrLdsBytes;
  if (ir_lds)
    for (int k = threadIdx.x; k < launch.ir_bytes; k += BLOCK) s_ir[k] = launch.ir_base[k];
  __syncthreads();
  auto ir_of = [&](const QueryJob& J) { return ir_lds ? s_ir + (J.ir - launch.ir_base) : J.ir; };
  const int64_t common = launch.lines_dev ? *launch.lines_dev : 0;
  int64_t count = (int64_t)launch.long_count[0];
  if (count > launch.long_cap) count = launch.long_cap;
  for (;;) {
    int64_t k = 0;
    if (g.r == 0) k = (int64_t)atomicAdd(&launch.long_count[1], 1ull);
    k = g.shfl_i64(k, 0);
    if (k >= count) break;
    const int64_t it = launch.long_items[k];
    if (launch.lines_dev) {
      int64_t q, i;
      if (launch.qblock > 0) {
        split_item(it, common, launch.njobs, launch.qblock, &q, &i);
      } else {
        q = it / common;
        i = it - q * common;
      }
      const QueryJob& J = launch.jobs[q];
      int64_t* out = J.out + (q * common + i) * (int64_t)J.R * 2;
      eval_line(g, J, GlobalWords{J.LB, J.Wp, J.rec, J.LS, J.Ws}, ir_of(J), common, i, out, le, ci, nslots);
    } else {
      int lo = 0, hi = (int)launch.njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (launch.jobs[mid].item0 <= it) lo = mid; else hi = mid - 1;
      }
      int64_t i = it - launch.jobs[lo].item0;
      if (launch.qblock > 0 && launch.jobs[lo].nqf > 0) {  // the same split as the first launch
        const int64_t j0 = lo - launch.jobs[lo].qf;
        int64_t q;
        split_item(it - launch.jobs[j0].item0, launch.jobs[lo].lines, launch.jobs[lo].nqf, launch.qblock, &q, &i);
        lo = (int)(j0 + q);
      }
      const QueryJob& J = launch.jobs[lo];
      eval_line(g, J, GlobalWords{J.LB, J.Wp, J.rec, J.LS, J.Ws}, ir_of(J), J.lines, i, J.out + i * (int64_t)J.R * 2, le, ci,
                nslots);
    }
  }
}


// Batched runs: fills each query job's line count and first work item from the
// per-file line counts of the newline bookkeeping, so the interpreter launch
// needs no host round trip.  One block of kPlanBlock threads: the jobs' line
// counts are scanned kPlanBlock at a time.
constexpr int kPlanBlock = 256;

__global__ __launch_bounds__(kPlanBlock) void batch_plan_kernel(const FileJob* files, int nfiles, QueryJob* qjobs,
                                                                const int* job_file, int njobs, int64_t* total_out,
                                                                const int64_t* lines_in, int* overflow) {
  __shared__ int64_t s_scan[kPlanBlock];
  const int tid = threadIdx.x;
  for (int f = tid; f < nfiles; f += kPlanBlock)
    if (lines_in[f] > files[f].nl_cap) *overflow = 1;
  int64_t carry = 0;
  for (int j0 = 0; j0 < njobs; j0 += kPlanBlock) {
    const int j = j0 + tid;
    int64_t lines = 0;
    if (j < njobs) {
      const int f = job_file[j];
      lines = lines_in[f] < files[f].nl_cap ? lines_in[f] : files[f].nl_cap;  // jobs stop at what the newline array holds
    }
    s_scan[tid] = lines;
    __syncthreads();
    for (int off = 1; off < kPlanBlock; off <<= 1) {  // inclusive scan in place
      const int64_t x = tid >= off ? s_scan[tid - off] : 0;
      __syncthreads();
      s_scan[tid] += x;
      __syncthreads();
    }
    if (j < njobs) {
      qjobs[j].lines = lines;
      qjobs[j].item0 = carry + s_scan[tid] - lines;
    }
    carry += s_scan[kPlanBlock - 1];
    __syncthreads();  // the next chunk overwrites s_scan
  }
  if (tid == 0) *total_out = carry;
}

hipError_t batch_plan(const FileJob* files, int nfiles, QueryJob* qjobs, const int* job_file, int njobs,
                      int64_t* total_out, const int64_t* lines, int* overflow, hipStream_t stream) {
  hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(kPlanBlock), 0, stream, files, nfiles, qjobs, job_file, njobs,
                     total_out, lines, overflow);
  return hipGetLastError();
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

bool query_config_supported(int block, int group) {
  return (block == 64 || block == 128 || block == 256) && (group == 1 || group == 4 || group == 16 || group == 64);
}

template <int BLOCK, int G>
static hipError_t launch_bg(const QueryLaunch& L, int nslots, int grid, hipStream_t stream) {
  const size_t slots = (size_t)(BLOCK / G) * nslots * 12 + 8;
  hipLaunchKernelGGL((query_kernel<BLOCK, G>), dim3(grid), dim3(BLOCK), slots, stream, L, nslots);
  return hipGetLastError();
}

template <int BLOCK>
static hipError_t launch_b(const QueryLaunch& L, int group, int nslots, int grid, hipStream_t stream) {
  switch (group) {
    case 1: return launch_bg<BLOCK, 1>(L, nslots, grid, stream);
    case 4: return launch_bg<BLOCK, 4>(L, nslots, grid, stream);
    case 16: return launch_bg<BLOCK, 16>(L, nslots, grid, stream);
    case 64: return launch_bg<BLOCK, 64>(L, nslots, grid, stream);
    default: return hipErrorInvalidValue;
  }
}

// levels: the deepest level any job's IR reaches (slots levels+2 per group).
// In line mode with a long-line list, a second launch (long_grid blocks of 256
// threads, a wavefront per line) follows on the same stream.
hipError_t query_launch(const QueryLaunch& L, int block, int group, int levels, int grid, hipStream_t stream,
                        int long_grid) {
  int nslots = levels + 2;
  if (nslots > 66) nslots = 66;
  hipError_t e = hipErrorInvalidValue;
  switch (block) {
    case 64: e = launch_b<64>(L, group, nslots, grid, stream); break;
    case 128: e = launch_b<128>(L, group, nslots, grid, stream); break;
    case 256: e = launch_b<256>(L, group, nslots, grid, stream); break;
    default: break;
  }
  if (e != hipSuccess || !L.long_items || group == 64) return e;
  const size_t slots = (size_t)(256 / 64) * nslots * 12 + 8;
  hipLaunchKernelGGL(query_long_kernel<256>, dim3(long_grid > 0 ? long_grid : 1), dim3(256), slots, stream, L,
                     nslots);
  return hipGetLastError();
}

hipError_t zero_u64(unsigned long long* p, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  int grid = (int)((n + 255) / 256);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace gj
