// index_kernel.hip -- fused structural index for gfx950.
//
// One pass over the input builds what the reference builds with six kernels
// and three host round trips (J/stages/CombinedIndex.java:28-151,
// J/stages/LeveledBitmapsIndex.java:40-103; kernels K/create_combined_*.cu,
// K/create_quote_index.cu, K/create_string_index.cu, K/create_leveled_*.cu):
//   newline index   int64[lines]   ([0] = 0, then every '\n')
//   records         uint64[W][4]   (in-string bits, raw ':' and '"' bytes, level 0)
//   leveled bitmaps uint64[L-1][Wp] (structural chars per depth level 1 .. L-1)
//   level summaries uint64[L][Ws]  (a bit per leveled word that is not zero)
//   escape / quote  uint64[W]      (only in verify mode)
//
// Work decomposition.  A tile is ROWS rows of BLOCK consecutive 64-byte words;
// lane t of the block owns word t of every row, so the lanes of a wavefront own
// consecutive words: each lane loads its 64 bytes straight into registers
// (16-byte loads), classifies them with byte-parallel (SWAR) compares into
// 64-bit masks, and every index store of the wavefront is coalesced.  All
// further bit arithmetic is word-parallel across the wavefront.  Tiles are
// claimed in increasing order per block through atomic tickets (one counter
// per residue class of tiles, so the claims do not all queue on one address),
// so the lowest unfinished tile is always one a block is working on.  Scans
// run in row-major word order.
//
// Cross-word state is a transfer function, not a value: a word's escape carry
// out is fixed unless the word is all backslashes, and given the escape carry,
// its effect on (in-string, depth) is (q, d0, tot): string parity flips by q,
// depth moves by d0 if the word starts outside a string and by tot-d0 if it
// starts inside one.  These compose associatively, so the tile computes its
// aggregate with a block scan, publishes it, and resolves its carry-in with a
// decoupled lookback over earlier tiles' published aggregates / inclusive
// prefixes.  The only dependence of a tile on the escape carry coming from the
// previous tile is through its first word that is not all backslashes (a quote
// right after a leading backslash run flips), so a tile publishes its
// aggregate for both carry values.
#include <hip/hip_runtime.h>
#include "bitops.h"
#include "fused_eval.h"
#include "index_common.h"

namespace gj {

namespace {

constexpr unsigned long long kValid = 1ull << 63;

struct M4 {
  int q;    // string-parity flip (0/1)
  int d0;   // depth change when entering outside a string
  int tot;  // opens minus closes, strings ignored
  int nl;   // newlines
};

__device__ __forceinline__ M4 m_id() { return M4{0, 0, 0, 0}; }

__device__ __forceinline__ M4 m_cat(const M4& x, const M4& y) {
  return M4{x.q ^ y.q, x.d0 + (x.q ? (y.tot - y.d0) : y.d0), x.tot + y.tot, x.nl + y.nl};
}

// y such that m_cat(x, y) == z
__device__ __forceinline__ M4 m_left_div(const M4& x, const M4& z) {
  M4 y;
  y.q = x.q ^ z.q;
  y.tot = z.tot - x.tot;
  y.nl = z.nl - x.nl;
  int dd = z.d0 - x.d0;
  y.d0 = x.q ? (y.tot - dd) : dd;
  return y;
}

// The same word evaluated with the opposite escape carry at its leading
// backslash run, when the byte after that run is a quote.
__device__ __forceinline__ M4 m_flip(const M4& x) { return M4{x.q ^ 1, x.tot - x.d0, x.tot, x.nl}; }

// DPP lane moves of an M4 (GFX9 controls: row_shr:n = 0x110 + n, row_bcast:15 =
// 0x142, row_bcast:31 = 0x143, wave_shr:1 = 0x138).  Lanes with no source lane,
// or in rows ROWMASK leaves out, get 0, which is the identity m_id().
template <int CTRL, int ROWMASK>
__device__ __forceinline__ M4 dpp_m(const M4& v) {
  return M4{__builtin_amdgcn_update_dpp(0, v.q, CTRL, ROWMASK, 0xF, false),
            __builtin_amdgcn_update_dpp(0, v.d0, CTRL, ROWMASK, 0xF, false),
            __builtin_amdgcn_update_dpp(0, v.tot, CTRL, ROWMASK, 0xF, false),
            __builtin_amdgcn_update_dpp(0, v.nl, CTRL, ROWMASK, 0xF, false)};
}

// The value of the lane below (lane 0: the identity).
__device__ __forceinline__ M4 lane_below_m(const M4& x) { return dpp_m<0x138, 0xF>(x); }

// Inclusive prefix sum of an int within each row of 16 lanes (DPP row shifts).
__device__ __forceinline__ int row16_add_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
  return x;
}

// Inclusive prefix sum of an int over the wavefront in lane order (DPP row
// shifts, then row broadcasts).
__device__ __forceinline__ int wave_add_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
  return x;
}

// Transfer functions composed in lane order, over the wavefront (or, ROW16,
// within each row of 16 lanes): returns this lane's exclusive prefix, *incl
// its inclusive one.  Add-scans instead of a scan of M4: the first sums quote
// parities and newlines (packed in one int: quote counts below 2^8, newlines
// at most 2^16 per tile), which gives each word the string state it is entered
// with from the start of the scan; the depth change for that state is then
// summed with the net depth.  Entered outside a string, the prefix's d0 is
// exactly that sum.  Over a wavefront both depths stay within +-2^12 and share
// one int (low half sign-extended); over a tile's rows they are scanned apart.
template <bool ROW16>
__device__ __forceinline__ M4 scan_words(const M4& w, M4* incl) {
  auto scan = [](int x) { return ROW16 ? row16_add_scan(x) : wave_add_scan(x); };
  const int a = (w.q & 1) + (w.nl << 8);
  const int ia = scan(a);
  const int ea = ia - a;
  const int d = (ea & 1) ? (w.tot - w.d0) : w.d0;  // entered inside a string: the other depth change
  int it, et, id, ed;
  if (ROW16) {
    it = scan(w.tot);
    et = it - w.tot;
    id = scan(d);
    ed = id - d;
  } else {
    const int b = w.tot + d * 65536;
    const int ib = scan(b);
    const int eb = ib - b;
    it = (int)(short)(ib & 0xFFFF);
    id = (ib - it) >> 16;
    et = (int)(short)(eb & 0xFFFF);
    ed = (eb - et) >> 16;
  }
  *incl = M4{ia & 1, id, it, (int)((unsigned)ia >> 8)};
  return M4{ea & 1, ed, et, (int)((unsigned)ea >> 8)};
}

// Carry state at a position.
struct Carry {
  int p;          // escape carry
  int s;          // inside a string
  int64_t d;      // depth
  int64_t lines;  // newline-index slot of the next '\n'
};

// Tile aggregate as a function of the incoming escape carry.
struct TileAgg {
  int esc;  // 0: passes the carry through, 2/3: sets it to 0/1
  M4 m[2];
};

__device__ __forceinline__ Carry apply_agg(const TileAgg& a, const Carry& c) {
  const M4 m = (c.p & 1) ? a.m[1] : a.m[0];
  Carry r;
  r.p = a.esc ? (a.esc - 2) : c.p;
  r.s = c.s ^ m.q;
  r.d = c.d + (c.s ? (int64_t)(m.tot - m.d0) : (int64_t)m.d0);
  r.lines = c.lines + m.nl;
  return r;
}

__device__ __forceinline__ TileAgg agg_cat(const TileAgg& x, const TileAgg& y) {
  TileAgg z;
  z.esc = y.esc ? y.esc : x.esc;
  // carry out of x for carry-in P: fixed by x unless x passes it through
  const M4 y0 = x.esc == 3 ? y.m[1] : y.m[0];
  const M4 y1 = x.esc == 2 ? y.m[0] : y.m[1];
  z.m[0] = m_cat(x.m[0], y0);
  z.m[1] = m_cat(x.m[1], y1);
  return z;
}

__device__ __forceinline__ int64_t sext(unsigned long long v, int bits) {
  return ((int64_t)(v << (64 - bits))) >> (64 - bits);
}

// Accesses to the index arrays and the input through global-segment pointers.
// A file's pointers reach a tile through LDS (its FileJob), which leaves them
// generic, and generic accesses compile to FLAT instructions: those count
// against the LDS wait counter as well, so every LDS wait would also wait for
// them.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4_t g_u32x4_t;
typedef __attribute__((address_space(1))) unsigned long long g_u64_t;
typedef __attribute__((address_space(1))) long long g_i64_t;

__device__ __forceinline__ uint4 gload16(const uint8_t* p) {
  const u32x4_t v = *(const g_u32x4_t*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore64(uint64_t* p, uint64_t v) { *(g_u64_t*)p = v; }
__device__ __forceinline__ void gstore64i(int64_t* p, int64_t v) { *(g_i64_t*)p = v; }
typedef __attribute__((address_space(1))) u32x4_t g_u32x4w_t;
// 16-byte store of two words (p 16-byte aligned).
__device__ __forceinline__ void gstore128(uint64_t* p, uint64_t a, uint64_t b) {
  u32x4_t v;
  v.x = (unsigned)a;
  v.y = (unsigned)(a >> 32);
  v.z = (unsigned)b;
  v.w = (unsigned)(b >> 32);
  *(g_u32x4w_t*)p = v;
}

// Two state words, each a relaxed atomic at device scope: every CU of every
// XCD sees the words other CUs published, and a store does not wait for its
// write to complete (a volatile access would, and the publishing wavefront
// would hold its block's other wavefronts at the next barrier meanwhile).  No
// ordering is needed: each word validates itself (bit 63), and a reader takes
// a pair only when both words are valid.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u64x2 ld_state2(const unsigned long long* p) {
  unsigned long long* q = const_cast<unsigned long long*>(p);
  u64x2 v;
  v.x = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v.y = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return v;
}

__device__ __forceinline__ void st_state2(unsigned long long* p, unsigned long long w0, unsigned long long w1) {
  __hip_atomic_store(p, w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, w1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Aggregate of a tile (at most 64 KiB) or of a group of 64 tiles (at most
// 4 MiB) in two words: counts in 24-bit fields.
//   w0: valid | carry passes (62) | carry out 1 (61) | q0 (60) | q1 (59) | tot (24..47) | newlines (0..23)
//   w1: valid | d0 for carry-in 1 (24..47) | d0 for carry-in 0 (0..23)
constexpr unsigned long long kField24 = (1ull << 24) - 1;

__device__ __forceinline__ void pack_agg(const TileAgg& a, unsigned long long& w0, unsigned long long& w1) {
  w0 = kValid | ((unsigned long long)(a.esc == 0) << 62) | ((unsigned long long)(a.esc == 3) << 61) |
       ((unsigned long long)(a.m[0].q & 1) << 60) | ((unsigned long long)(a.m[1].q & 1) << 59) |
       (((unsigned long long)a.m[0].tot & kField24) << 24) | ((unsigned long long)a.m[0].nl & kField24);
  w1 = kValid | ((unsigned long long)a.m[0].d0 & kField24) | (((unsigned long long)a.m[1].d0 & kField24) << 24);
}

__device__ __forceinline__ TileAgg decode_agg(unsigned long long w0, unsigned long long w1) {
  TileAgg a;
  a.esc = ((w0 >> 62) & 1) ? 0 : (2 + (int)((w0 >> 61) & 1));
  const int nl = (int)(w0 & kField24);
  const int tot = (int)sext((w0 >> 24) & kField24, 24);
  a.m[0] = M4{(int)((w0 >> 60) & 1), (int)sext(w1 & kField24, 24), tot, nl};
  a.m[1] = M4{(int)((w0 >> 59) & 1), (int)sext((w1 >> 24) & kField24, 24), tot, nl};
  return a;
}

__device__ void publish_agg(TileState* st, const TileAgg& a) {
  unsigned long long w0, w1;
  pack_agg(a, w0, w1);
  st_state2(&st->agg0, w0, w1);
}

// The aggregate of the 64 tiles of the group that st's tile ends.
__device__ void publish_group(TileState* st, const TileAgg& a) {
  unsigned long long w0, w1;
  pack_agg(a, w0, w1);
  st_state2(&st->pad[0], w0, w1);
}

__device__ void publish_inc(TileState* st, const Carry& c) {
  const unsigned long long w0 = kValid | ((unsigned long long)(c.p & 1) << 62) |
                                ((unsigned long long)(c.s & 1) << 61) |
                                ((unsigned long long)c.d & 0xFFFFFFFFFFFFull);
  const unsigned long long w1 = kValid | ((unsigned long long)c.lines & ~kValid);
  st_state2(&st->inc0, w0, w1);
}

// The carry-in a tile was indexed with (deferred bookkeeping), in the words
// after its group aggregate: valid | p (62) | s (61) | speculated (60) | depth
// (48 bits).
__device__ void store_carry_in(TileState* st, const Carry& c, bool spec) {
  const unsigned long long w0 = kValid | ((unsigned long long)(c.p & 1) << 62) |
                                ((unsigned long long)(c.s & 1) << 61) | ((unsigned long long)spec << 60) |
                                ((unsigned long long)c.d & 0xFFFFFFFFFFFFull);
  st_state2(&st->pad[2], w0, 0);
}

__device__ __forceinline__ TileAgg agg_id() {
  TileAgg a;
  a.esc = 0;
  a.m[0] = m_id();
  a.m[1] = m_id();
  return a;
}

__device__ __forceinline__ TileAgg shfl_down_agg(const TileAgg& a, int off) {
  TileAgg b;
  b.esc = __shfl_down(a.esc, off);
#pragma unroll
  for (int P = 0; P < 2; P++) {
    b.m[P].q = __shfl_down(a.m[P].q, off);
    b.m[P].d0 = __shfl_down(a.m[P].d0, off);
    b.m[P].tot = __shfl_down(a.m[P].tot, off);
    b.m[P].nl = __shfl_down(a.m[P].nl, off);
  }
  return b;
}

__device__ __forceinline__ TileAgg shfl_agg(const TileAgg& a, int src) {
  TileAgg b;
  b.esc = __shfl(a.esc, src);
#pragma unroll
  for (int P = 0; P < 2; P++) {
    b.m[P].q = __shfl(a.m[P].q, src);
    b.m[P].d0 = __shfl(a.m[P].d0, src);
    b.m[P].tot = __shfl(a.m[P].tot, src);
    b.m[P].nl = __shfl(a.m[P].nl, src);
  }
  return b;
}

__device__ __forceinline__ int64_t shfl_i64(int64_t v, int src) {
  unsigned lo = (unsigned)(uint64_t)v, hi = (unsigned)((uint64_t)v >> 32);
  lo = __shfl(lo, src);
  hi = __shfl(hi, src);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Lane l holds the aggregate of tile k_hi - l; returns (to every lane) the
// composition of lanes [0, 64) in tile order, i.e. x_63 then ... then x_0.
__device__ __forceinline__ TileAgg reduce_tiles(TileAgg x, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    TileAgg y = shfl_down_agg(x, off);
    if (lane + off < 64) x = agg_cat(y, x);
  }
  return shfl_agg(x, 0);
}

// Exclusive carry of local tile lt of a file whose tiles start at st0,
// resolved by one wavefront.  Tiles form groups of 64 (tiles 64g .. 64g+63);
// the last tile of a group publishes the group's aggregate (in its state's pad
// words) as soon as the group's other tiles have published theirs, and a
// group's inclusive prefix is that of its last tile.  The walk reads, per
// round trip, one unit per lane nearest first: the tiles of the cursor tile's
// group at or below it, then whole groups below.  It stops at the first unit
// holding an inclusive prefix (done) or not ready: a tile not ready is waited
// for, a group without its aggregate is walked tile by tile.  Tile 0 of a file
// always publishes its inclusive prefix, so the walk ends there at the latest.
__device__ Carry lookback_wave(TileState* st0, int64_t lt, int lane, const TileAgg& own) {
  TileAgg acc = agg_id();  // units between the cursor and lt, already folded
  int64_t c = lt - 1;      // cursor: the nearest tile not folded yet
  bool group_todo = (lt & 63) == 63;
  const Carry file_start{0, 0, -1, 1};  // the empty prefix before tile 0
  for (;;) {
    if (c < 0) return apply_agg(acc, file_start);
    const int64_t g = c >> 6;
    const int p = (int)(c - (g << 6)) + 1;  // tiles of the cursor's group at or below it
    int status = 0;  // 0 not ready, 1 aggregate, 2 inclusive
    bool agg_ok = false;
    TileAgg a = agg_id();
    Carry cc{0, 0, 0, 0};
    const TileState* s = nullptr;
    bool is_group = false;
    if (lane < p) {
      s = st0 + (c - lane);
    } else if (g - 1 - (lane - p) >= 0) {
      s = st0 + ((g - 1 - (lane - p)) * 64 + 63);
      is_group = true;
    } else {
      status = 2;  // before the file: the empty prefix
      cc = file_start;
    }
    if (s) {
      // both halves in one round trip (each word validates itself)
      const u64x2 iv = ld_state2(&s->inc0);
      const u64x2 av = ld_state2(is_group ? &s->pad[0] : &s->agg0);
      agg_ok = (av.x & kValid) && (av.y & kValid);
      if (agg_ok) a = decode_agg(av.x, av.y);
      if ((iv.x & kValid) && (iv.y & kValid)) {
        status = 2;
        cc.p = (int)((iv.x >> 62) & 1);
        cc.s = (int)((iv.x >> 61) & 1);
        cc.d = sext(iv.x & 0xFFFFFFFFFFFFull, 48);
        cc.lines = (int64_t)(iv.y & ~kValid);
      } else if (agg_ok) {
        status = 1;
      }
    }
    if (group_todo) {
      // the group's last tile waits for the aggregates of its other 63 tiles
      // (lanes 0..62, nearest first; the cursor stays at lt - 1 meanwhile) and
      // publishes the group's: they are all running and publish without waiting
      const unsigned long long m_ok = __ballot(lane < 63 && agg_ok);
      if (m_ok != (~0ull >> 1)) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      const TileAgg grp = agg_cat(reduce_tiles(lane < 63 ? a : agg_id(), lane), own);
      if (lane == 0) publish_group(st0 + lt, grp);
      group_todo = false;
    }
    const unsigned long long m_inc = __ballot(status == 2);
    const unsigned long long m_nr = __ballot(status == 0);
    const int a_inc = m_inc ? __builtin_ctzll(m_inc) : 64;
    const int a_nr = m_nr ? __builtin_ctzll(m_nr) : 64;
    if (a_inc < a_nr) {
      // units [0, a_inc) are aggregates, unit a_inc the nearest inclusive prefix
      if (lane >= a_inc) a = agg_id();
      const TileAgg part = reduce_tiles(a, lane);
      Carry base;
      base.p = __shfl(cc.p, a_inc);
      base.s = __shfl(cc.s, a_inc);
      base.d = shfl_i64(cc.d, a_inc);
      base.lines = shfl_i64(cc.lines, a_inc);
      return apply_agg(agg_cat(part, acc), base);
    }
    if (a_nr == 0) {  // unit 0 is always a tile: wait for it
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    // fold the ready aggregates [0, a_nr) and continue at the unit that stopped
    // (all 64 units aggregates: continue below them)
    if (lane >= a_nr) a = agg_id();
    acc = agg_cat(reduce_tiles(a, lane), acc);
    if (a_nr < p) c -= a_nr;                       // a tile
    else if (a_nr < 64) c = (g - 1 - (a_nr - p)) * 64 + 63;  // a group without its aggregate: its tiles
    else c = (g - (64 - p)) * 64 - 1;               // below the last group read
  }
}

}  // namespace

// Per-word masks parked in LDS between classification and emission: the
// in-string mask for escape carry 0 outside a string, newlines, open, close,
// ':' ',', quotes and ':'; verify mode also keeps the backslashes.  (The colon
// and quote words need no carry, but writing them apart from the string word
// of the same record sends many record lines to memory twice.)
constexpr int kSlotS0 = 0, kSlotNL = 1, kSlotO = 2, kSlotC = 3, kSlotP = 4, kSlotQ = 5, kSlotK = 6, kSlotB = 7;

// Leveled bitmaps with the brackets spread over the wavefront.  Lane-per-word bracket loops run as long as the word
// with the most brackets (about 10x the mean on tweets); here every bracket of
// the wavefront's 64 words is one work item.  Per wavefront: the word's level
// words (up to kSpreadLevels x 64) and the work list (kBracketCap items of
// word lane << 6 | bit); a row with more brackets than that takes the
// per-word loop.  320 items (5 rounds of the wavefront) is as many as fit
// beside the rest of the block's LDS at 4 blocks per CU (160 KiB): on
// bracket-dense input most rows then stay on the spread path (128 items left
// more than half of them to the per-word loop).
constexpr int kBracketCap = 320;
constexpr int kSpreadLevels = 4;  // levels per pass over the brackets

// Inclusive prefix sum of an int over the wavefront (DPP, as wave_add_scan).
__device__ __forceinline__ int wave_scan_int(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
  return x;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  const unsigned lo = __shfl((unsigned)v, src), hi = __shfl((unsigned)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// LDS accesses of one wavefront made visible to its other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Level words [lo, lo + NLV) of this lane's word (O, C, P outside strings, d
// the depth before it, clamped to [-66, L + 66]; act false: no word), every
// lane of the wavefront taking part.  lv: this wavefront's NLV x 64 level
// words; items: its list.
template <int NLV>
__device__ __forceinline__ void levels_spread(uint64_t O, uint64_t C, uint64_t P, int d, int lo, int L,
                                              bool act, int lane, uint64_t (&lv)[NLV][64],
                                              uint16_t (&items)[kBracketCap], uint64_t out[NLV]) {
  const int dmin = d - popc64(C), dmax = d + popc64(O);
  const bool hit = act && !(dmax < lo || dmin >= lo + NLV) && (O | C | P) != 0;
  const uint64_t T = hit ? (O | C) : 0;
  const int n = popc64(T);
  const int incl = wave_scan_int(n);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  if (total > kBracketCap) {  // wave-uniform: the per-word loop
#pragma unroll
    for (int i = 0; i < NLV; i++) out[i] = 0;
    if (hit) leveled_word<NLV>(O, C, P, d, lo, L, out);
    return;
  }
  // own level words: zero, then the separators before the first bracket
#pragma unroll
  for (int i = 0; i < NLV; i++) lv[i][lane] = 0;
  if (hit) {
    const int k = d - lo;
    if (k >= 0 && k < NLV && d < L) lv[k][lane] = leading_seps(O, C, P);
  }
  // the work list: this word's brackets at [incl - n, incl)
  {
    int j = incl - n;
    uint64_t Tm = T;
    while (Tm) {
      items[j++] = (uint16_t)((lane << 6) | ctz64(Tm));
      Tm &= Tm - 1;
    }
  }
  wave_lds_sync();
  for (int u0 = 0; u0 < total; u0 += 64) {
    const int u = u0 + lane;
    const int e = u < total ? items[u] : 0;
    const int o = e >> 6, t = e & 63;
    const uint64_t Oo = shfl_u64(O, o), Co = shfl_u64(C, o), Po = shfl_u64(P, o);
    const int dd = __shfl(d, o);
    if (u < total) {
      int lvl, after;
      const uint64_t seps = bracket_step(Oo, Co, Po, dd, t, &lvl, &after);
      const int k = lvl - lo;
      if (k >= 0 && k < NLV && lvl < L) atomicOr(&lv[k][o], 1ull << t);
      const int k2 = after - lo;
      if (seps && k2 >= 0 && k2 < NLV && after < L) atomicOr(&lv[k2][o], seps);
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < NLV; i++) out[i] = lv[i][lane];
  wave_lds_sync();  // the next row overwrites these words
}

// The index words of word w -- string index, raw colon and quote bitmaps,
// verify-mode escape/quote words, leveled bitmaps -- from its masks (row r,
// lane tid of `wc`), the escape carry p, string state s and depth d before its
// first byte.
// act false: a lane past the last word (it writes nothing, but takes part in
// the wavefront's work when the brackets are spread).
// Fused tiles (in-tile chain evaluation, fz.on): the word's index words go to
// memory only when wr (the word lies at or before the tile's first newline or
// at or after its last), and the final string word and levels 0..2 are left in
// the word's LDS slots (S, O, C, P) with the levels' summaries in fz.sl.
struct FuseEmit {
  bool on;
  bool wr;
  uint64_t* sl;  // [kFuseMaxDepth][TILE_WORDS / 64]
  int nsum;
};

// LDS slot of leveled word l < 3 of a fused tile (the bracket masks are dead by then).
__device__ __forceinline__ constexpr int fuse_lb_slot(int l) { return l == 0 ? kSlotO : (l == 1 ? kSlotC : kSlotP); }

// A fused tile's words in LDS, as fused_eval.h reads them: word w of the tile
// is lane w % BLOCK of row w / BLOCK; after emission its slots hold the final
// string word (S0), newlines, the leveled words 0..2 (O, C, P), raw quotes and,
// after the key tests, the match word (K).
template <int ROWS, int NSLOT, int BLOCK>
struct FuseTile {
  const uint64_t (&wc)[ROWS][NSLOT][BLOCK];
  const uint64_t* sm;  // [kFuseMaxDepth][nsum]: level l's match words that are not zero
  const uint64_t* sl;  // [kFuseMaxDepth][nsum]: level l's leveled words that are not zero
  int nsum;
  __device__ __forceinline__ uint64_t at(int slot, int w) const { return wc[w / BLOCK][slot][w % BLOCK]; }
  __device__ __forceinline__ uint64_t S(int w) const { return at(kSlotS0, w); }
  __device__ __forceinline__ uint64_t NL(int w) const { return at(kSlotNL, w); }
  __device__ __forceinline__ uint64_t Q(int w) const { return at(kSlotQ, w); }
  __device__ __forceinline__ uint64_t M(int w) const { return at(kSlotK, w); }
  __device__ __forceinline__ uint64_t LB(int l, int w) const { return at(fuse_lb_slot(l), w); }
  __device__ __forceinline__ uint64_t SM(int l, int s) const { return sm[l * nsum + s]; }
  __device__ __forceinline__ uint64_t SL(int l, int s) const { return sl[l * nsum + s]; }
};

template <int NLV, bool VERIFY, int ROWS, int NSLOT, int BLOCK, int NLVS>
__device__ __forceinline__ void emit_word(const FileJob& J, int64_t w, uint64_t (&wc)[ROWS][NSLOT][BLOCK],
                                          const uint8_t (&lead)[ROWS][BLOCK], int r, int tid, int p, int s,
                                          int d, bool act, bool full, uint64_t (&lv)[NLVS][64],
                                          uint16_t (&items)[kBracketCap], const FuseEmit& fz) {
  const bool wr = act && (!fz.on || fz.wr);
  uint64_t valid = ~0ull;  // bytes of the word inside the file (full: the tile's words all are)
  if (!full) {
    const int64_t rem = J.N - w * 64;
    if (rem < 64) valid = rem > 0 ? (1ull << rem) - 1ull : 0;
  }
  // in-string mask from the one for escape carry 0 outside a string: carry 1
  // escapes the leading backslash run's first byte instead of its second, which
  // toggles only a quote right after the run, and so every bit from there on
  const int ld = lead[r][tid];
  uint64_t S = wc[r][kSlotS0][tid];
  if (p && (ld & 0x80)) S ^= ~0ull << (ld & 0x7F);
  S = (s ? ~S : S) & valid;
  if (VERIFY && act) {
    const uint64_t E = escape_word(wc[r][kSlotB][tid], (uint32_t)p);
    if (J.E) J.E[w] = E & valid;
    if (J.Q) J.Q[w] = wc[r][kSlotQ][tid] & ~E & valid;
  }
  const uint64_t O = wc[r][kSlotO][tid] & ~S, C = wc[r][kSlotC][tid] & ~S, P = wc[r][kSlotP][tid] & ~S;
  if (fz.on) wc[r][kSlotS0][tid] = S;  // only this lane reads its slots until the next barrier
  // A word moves the depth by at most 64, so a depth below -66 or above L + 66
  // puts every byte of the word below level 0 or at level L or above either
  // way: clamped there it leaves the level bits unchanged.
  const int dc = min(max(d, -66), J.L + 66);
  uint64_t lb0 = 0;  // level 0, stored with the record
  for (int lo = 0; lo < J.L; lo += NLVS) {
    uint64_t out[NLVS];
    levels_spread<NLVS>(O, C, P, dc, lo, J.L, act, tid & 63, lv, items, out);
#pragma unroll
    for (int i = 0; i < NLVS; i++) {
      if (lo + i == 0) lb0 = out[i];
      else if (wr && lo + i < J.L) gstore64(J.LB + (int64_t)(lo + i - 1) * J.Wp + w, out[i]);
    }
    // summaries: the 64 lanes of a wavefront hold 64 consecutive words that
    // start at a multiple of 64 (lanes past the file end are inactive, and
    // lane 0 holds the lowest word, so it is active whenever any lane is)
    if (J.LS || fz.on) {
#pragma unroll
      for (int i = 0; i < NLVS; i++) {
        if (lo + i < J.L) {
          const unsigned long long m = __ballot(out[i] != 0);
          if ((tid & 63) == 0 && act && J.LS) gstore64(J.LS + (int64_t)(lo + i) * J.Ws + (w >> 6), m);
          if (fz.on && lo + i < kFuseMaxDepth) {
            if ((tid & 63) == 0) fz.sl[(lo + i) * fz.nsum + ((r * BLOCK + tid) >> 6)] = m;
            wc[r][fuse_lb_slot(lo + i)][tid] = out[i];
          }
        }
      }
    }
  }
  // the word's record: string, colon, raw-quote and level-0 words (bytes past
  // the end of the file were loaded as zeros, which are neither ':' nor '"')
  if (wr) {
    uint64_t* rec = J.rec + w * kRecWords;
    gstore128(rec, S, wc[r][kSlotK][tid]);
    gstore128(rec + 2, wc[r][kSlotQ][tid], lb0);
  }
}

template <int BLOCK, int ROWS, int NLV, bool VERIFY, bool FUSE>
// Registers are held to what the tile's LDS allows per CU: 256 x 2 tiles (the
// default, about 35 KB of LDS) fit 4 blocks per CU and 256 x 3 tiles 3, so
// 128 and 168 registers.  The fourth wavefront per SIMD pays for a few spilled
// registers: 256 x 2 at 4 blocks per CU indexed 4 GiB of tweets in 1.68-1.78 ms
// against 1.73-1.88 ms for 256 x 3 at 3.
__global__ __launch_bounds__(BLOCK, (BLOCK == 256 && ROWS == 3 && !VERIFY)   ? 3
                                    : (BLOCK == 256 && ROWS == 2 && !VERIFY) ? 4
                                                                              : 1) void index_kernel(
    IndexLaunch launch) {
  static_assert(!(FUSE && VERIFY), "in-tile evaluation runs without the verify words");
  static_assert(!FUSE || NLV <= kFuseMaxDepth, "in-tile evaluation keeps at most three levels in LDS");
  constexpr int NW = BLOCK / 64;
  constexpr int TILE_WORDS = BLOCK * ROWS;
  static_assert(BLOCK % 64 == 0, "a wavefront's words must start at a multiple of 64 (level summaries)");
  static_assert(TILE_WORDS * 64 <= 65536, "lookback state fields assume tiles of at most 64 KiB (groups of 4 MiB)");

  // Per-tile scalars, two sets used by alternate tiles: thread 0 prepares the
  // block's next tile (its ticket, its file, the cleared fields) before the
  // barrier that ends the current tile, so a tile starts without a barrier.
  struct TileSlot {
    unsigned ticket;
    int w0;       // sequence index of the tile's first word that is not all '\\'
    int w0_flip;
    int fnl;      // sequence index of the tile's first word holding a '\n'
    int lnl;      // sequence index of the tile's last word holding a '\n' (in-tile evaluation)
    int ncand;    // key candidates whose bytes are still to be compared (in-tile evaluation)
    M4 w0_m;
  };
  __shared__ TileSlot s_slot[2];
  __shared__ FileJob s_J[2];       // the file of each set's tile
  __shared__ int s_esc[ROWS][NW];  // last escape code that is not 0 of each (row, wave)
  __shared__ M4 s_m[ROWS][NW];     // transfer function of each (row, wave)
  static_assert(ROWS * NW <= 16, "the (row, wave) prefixes are scanned within one DPP row");
  __shared__ Carry s_carry;
  __shared__ Carry s_exact;        // carry-in from the lookback (speculative tiles)
  __shared__ int s_mis;            // the speculative carry-in was wrong
  // the masks every word needs at emission, parked here (field-major, so the
  // lanes of a wavefront touch consecutive 8-byte slots) while the scans and
  // the lookback run, so they occupy no registers there; s_lead: length of the
  // word's leading backslash run (64: all backslashes) | 0x80 if a quote follows it
  constexpr int NSLOT = VERIFY ? 8 : 7;
  __shared__ uint64_t s_wc[ROWS][NSLOT][BLOCK];
  // per wavefront: level words and work list of the spread leveled bitmaps
  constexpr int NLVS = NLV < kSpreadLevels ? NLV : kSpreadLevels;
  __shared__ uint64_t s_lv[NW][NLVS][64];
  __shared__ uint16_t s_items[NW][kBracketCap];
  __shared__ uint8_t s_lead[ROWS][BLOCK];
  // in-tile evaluation: key bytes and summaries; after emission the tile's
  // newline positions reuse the bracket lists and the key candidates the
  // level scratch
  constexpr int NSUM = TILE_WORDS / 64;
  __shared__ uint8_t s_keys[FUSE ? kFuseKeyBytes : 1];
  __shared__ uint64_t s_fsum[FUSE ? 2 * kFuseMaxDepth * NSUM : 1];  // [match | level][l][NSUM]
  static_assert(!FUSE || TILE_WORDS / 3 <= NW * kBracketCap, "the newline positions fit the bracket lists");

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const unsigned long long lanes_below = (1ull << lane) - 1ull;
  if (launch.run_if && *launch.run_if == 0) return;
  const bool dl = launch.nl_scratch != nullptr;  // deferred newline bookkeeping
  if (FUSE)
    for (int k = tid; k < kFuseKeyBytes; k += BLOCK) s_keys[k] = launch.fuse_keys[k];

  // Thread 0 claims each tile one tile ahead: the ticket for the block's next
  // tile is taken when its current tile starts, so the atomic's round trip
  // overlaps the current tile instead of standing before the next one.  The
  // tiles lie in npart residue classes (tile % npart), each with its counter;
  // block b claims from class b % npart, and every class has a block because
  // npart <= gridDim.x.  Claims stay in increasing order per block, and a
  // class's tiles are claimed in increasing order, so the lowest unfinished
  // tile is always one a block is working on or its class's next (its class's
  // blocks have finished every earlier tile of the class, so one of them claims
  // it next): the lookback still cannot wait forever.
  // The 2-level variant takes two consecutive tiles of its class per atomic
  // when every block gets 64 tiles or more (half the claims on the counter:
  // 1.66 -> 1.57-1.62 ms on 4 GiB of tweets at L = 1, 2).  On fewer tiles
  // pairs leave blocks idle at the end (1 GiB of bracket-dense lines, 32 tiles
  // per block: 0.61 -> 0.68 ms; one-query passes over a mix of 1-512 MiB
  // files: 377 -> 399 ms), and the deeper variants measured slower with
  // pairs.  A block still holds its tiles in increasing order.
  const unsigned kClaim =
      NLV <= 2 && launch.total_tiles - launch.tile_begin >= 64 * (int64_t)gridDim.x ? 2u : 1u;
  unsigned npart = 1;
  while (npart * 2 <= (unsigned)kTicketParts && npart * 2 <= gridDim.x) npart *= 2;
  const unsigned part = blockIdx.x & (npart - 1);
  unsigned int* const my_ticket = launch.ticket + part * kTicketStride;
  unsigned next_ticket = 0;            // thread 0 only
  unsigned in_claim = 0;               // thread 0 only: tiles of the current claim handed out
  int cur_job = -1;                    // thread 0 only: the job of the last tile prepared
  int64_t cur_end = 0;                 // thread 0 only: its first tile past the job
  int slot_job[2] = {-1, -1};          // thread 0 only: the job each s_J set holds
  auto prepare = [&](int k) {          // thread 0: set k for the block's next tile
    const unsigned t = (unsigned)launch.tile_begin + next_ticket;
    TileSlot& q = s_slot[k];
    q.ticket = t;
    q.w0 = -1;
    q.w0_flip = 0;
    q.fnl = TILE_WORDS;
    q.lnl = -1;
    q.ncand = 0;
    if (t >= launch.total_tiles) return;
    if (++in_claim < kClaim) {
      next_ticket += npart;
    } else {
      in_claim = 0;
      next_ticket = atomicAdd(my_ticket, kClaim) * npart + part;
    }
    int lo = 0;
    if (launch.jobs) {
      if (cur_job >= 0 && (int64_t)t < cur_end) {
        lo = cur_job;  // tickets grow, so the tile is in the last job or a later one
      } else {
        int hi = (int)launch.njobs - 1;
        while (lo < hi) {  // last job with tile0 <= t
          const int mid = (lo + hi + 1) >> 1;
          if (launch.jobs[mid].tile0 <= (int64_t)t) lo = mid; else hi = mid - 1;
        }
      }
    }
    if (lo != slot_job[k]) {  // a tile of another file than the set's last one
      s_J[k] = launch.jobs ? launch.jobs[lo] : launch.one;
      slot_job[k] = lo;
    }
    if (lo != cur_job) {
      cur_job = lo;
      cur_end = s_J[k].tile0 + s_J[k].ntiles;
    }
  };
  if (tid == 0) {
    next_ticket = atomicAdd(my_ticket, kClaim) * npart + part;
    prepare(0);
  }
  __syncthreads();
  // This lane's input word of every row of a tile (bytes past the end of the
  // file read as zeros).
  uint4 raw[ROWS][4];
  auto load_rows = [&](const FileJob& JJ, int64_t w0) {
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int64_t off = (w0 + (int64_t)r * BLOCK + tid) * 64;
      if (off + 64 <= JJ.N) {
#pragma unroll
        for (int q = 0; q < 4; q++) raw[r][q] = gload16(JJ.F + off + 16 * q);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          uint32_t w[4] = {0, 0, 0, 0};
          for (int b = 0; b < 16; b++) {
            const int64_t o = off + 16 * q + b;
            if (o < JJ.N) w[b >> 2] |= (uint32_t)JJ.F[o] << (8 * (b & 3));
          }
          raw[r][q] = make_uint4(w[0], w[1], w[2], w[3]);
        }
      }
    }
  };
  for (int par = 0;; par ^= 1) {
    TileSlot& ts = s_slot[par];
    const unsigned ticket = ts.ticket;
    if (ticket >= launch.total_tiles) break;
    const FileJob J = s_J[par];
    const int64_t lt = (int64_t)ticket - J.tile0;
    const int64_t word0 = lt * (int64_t)TILE_WORDS;

    // ---- load and classify this lane's word of every row ----
    int code[ROWS];  // escape summary: 0 all backslashes, else 2 + carry out
    M4 wm0[ROWS];    // the word's transfer function for escape carry-in 0
    int flip[ROWS];  // the byte after the leading backslash run is a quote
    bool rowb[ROWS];  // the wavefront's words of the row hold a backslash (wave-uniform)
    {
      load_rows(J, word0);
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        uint32_t d[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          d[4 * q + 0] = raw[r][q].x;
          d[4 * q + 1] = raw[r][q].y;
          d[4 * q + 2] = raw[r][q].z;
          d[4 * q + 3] = raw[r][q].w;
        }
        const WordClass w = classify64(d);
        // Most rows hold no backslash at all: then no byte is escaped with
        // carry-in 0, every word passes carry 0 on, and its leading backslash
        // run is empty (so a quote after it is a quote at byte 0).
        rowb[r] = __ballot(w.B != 0) != 0;
        uint64_t Qe = w.Q;
        if (rowb[r]) {
          code[r] = (w.B == ~0ull) ? 0 : (2 + (int)escape_carry_out(w.B, 0));
          flip[r] = code[r] ? (int)((w.Q >> ctz64(~w.B)) & 1) : 0;
          Qe &= ~escape_word(w.B, 0u);
          s_lead[r][tid] = (uint8_t)((code[r] ? ctz64(~w.B) : 64) | (flip[r] << 7));
        } else {
          code[r] = 2;
          flip[r] = (int)(w.Q & 1);
          s_lead[r][tid] = (uint8_t)(flip[r] << 7);
        }
        const uint64_t S0 = prefix_xor(Qe);
        M4 m;
        m.q = popc64(Qe) & 1;
        m.d0 = popc64(w.O & ~S0) - popc64(w.C & ~S0);
        m.tot = popc64(w.O) - popc64(w.C);
        m.nl = popc64(w.NL);
        wm0[r] = m;
        s_wc[r][kSlotS0][tid] = S0;
        s_wc[r][kSlotNL][tid] = w.NL;
        s_wc[r][kSlotO][tid] = w.O;
        s_wc[r][kSlotC][tid] = w.C;
        s_wc[r][kSlotP][tid] = w.P;
        if (VERIFY) s_wc[r][kSlotB][tid] = w.B;
        s_wc[r][kSlotQ][tid] = w.Q;
        s_wc[r][kSlotK][tid] = w.K;
        if (!VERIFY) {
          const unsigned long long nlw = __ballot(w.NL != 0);
          if (nlw && lane == __builtin_ctzll(nlw)) atomicMin(&ts.fnl, r * BLOCK + tid);
          if (FUSE && nlw && lane == 63 - __builtin_clzll(nlw)) atomicMax(&ts.lnl, r * BLOCK + tid);
        }
      }
    }

    // ---- escape carry: "last word that is not all backslashes" scan ----
    int ent[ROWS];  // code of the nearest such word before this one in the tile (0: none)
    {
      unsigned long long nzm[ROWS];
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        if (!rowb[r]) {  // every word's code is 2 (carry 0)
          if (lane == 0) s_esc[r][wave] = 2;
          ent[r] = lane ? 2 : -1;
          continue;
        }
        nzm[r] = __ballot(code[r] != 0);
        const int last = nzm[r] ? __shfl(code[r], 63 - __clzll(nzm[r])) : 0;
        if (lane == 0) s_esc[r][wave] = last;
        const unsigned long long below = nzm[r] & lanes_below;
        const int e = __shfl(code[r], below ? (63 - __clzll(below)) : 0);
        ent[r] = below ? e : -1;  // -1: resolve from earlier waves / rows
      }
      __syncthreads();
      int run = 0;  // last code that is not 0 before the current (row, wave)
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
#pragma unroll
        for (int w = 0; w < NW; w++) {
          if (w == wave && ent[r] < 0) ent[r] = run;
          const int e = s_esc[r][w];
          if (e) run = e;
        }
      }
    }
    int tile_esc = 0;
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int w = 0; w < NW; w++)
        if (s_esc[r][w]) tile_esc = s_esc[r][w];

    // ---- per-word transfer functions (escape carry 0 where it comes from P) ----
    M4 wm[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      wm[r] = wm0[r];
      if (ent[r] == 3) {
        // escape carry-in 1 (the word before ends in an odd backslash run): rare
        // (the quote right after the leading backslash run, if any, toggles)
        const int ld = s_lead[r][tid];
        const uint64_t S0 = s_wc[r][kSlotS0][tid], O = s_wc[r][kSlotO][tid], C = s_wc[r][kSlotC][tid];
        const uint64_t S1 = (ld & 0x80) ? (S0 ^ (~0ull << (ld & 0x7F))) : S0;
        wm[r].q ^= ld >> 7;
        wm[r].d0 = popc64(O & ~S1) - popc64(C & ~S1);
      }
      if (ent[r] == 0 && code[r] != 0) {
        // the tile's first word that is not all backslashes
        ts.w0 = r * BLOCK + tid;
        ts.w0_flip = flip[r];
        ts.w0_m = wm[r];
      }
    }

    // ---- block scan of the transfer functions in row-major order ----
    M4 lex[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      M4 inc;
      lex[r] = scan_words<false>(wm[r], &inc);
      if (lane == 63) s_m[r][wave] = inc;
    }
    __syncthreads();
    // the (row, wave) totals' exclusive prefixes: every wavefront scans them
    // itself (within one 16-lane DPP row), so no barrier waits for one of them
    M4 texcl[ROWS];
    M4 mtile;
    {
      M4 v;
      const M4 ex = scan_words<true>(lane < ROWS * NW ? s_m[lane / NW][lane % NW] : m_id(), &v);
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        const int k = __builtin_amdgcn_readfirstlane(r * NW + wave);
        const M4 e{__builtin_amdgcn_readlane(ex.q, k), __builtin_amdgcn_readlane(ex.d0, k),
                   __builtin_amdgcn_readlane(ex.tot, k), __builtin_amdgcn_readlane(ex.nl, k)};
        texcl[r] = m_cat(e, lex[r]);
      }
      mtile = M4{__builtin_amdgcn_readlane(v.q, ROWS * NW - 1), __builtin_amdgcn_readlane(v.d0, ROWS * NW - 1),
                 __builtin_amdgcn_readlane(v.tot, ROWS * NW - 1), __builtin_amdgcn_readlane(v.nl, ROWS * NW - 1)};
    }

    // ---- carry-in: speculated at the first newline, or from the lookback ----
    // In NDJSON every raw '\n' lies outside strings at depth -1 (one value per
    // line).  A tile holding a newline whose first byte is neither '\\' nor '"'
    // (so the escape carry into it changes none of its outputs) takes the
    // carry-in that puts its first newline in that state, emits its indexes
    // at once, and resolves its exact carry-in with the lookback afterwards:
    // only the newline offsets need it.  A wrong guess (the input is not
    // NDJSON there) re-emits the tile with the exact carry-in.
    const bool spec = launch.speculate && !VERIFY && lt > 0 && ts.fnl < TILE_WORDS && s_lead[0][0] == 0;
    TileAgg a;
    a.esc = tile_esc;
    a.m[0] = mtile;
    a.m[1] = mtile;
    if (ts.w0 >= 0 && ts.w0_flip) a.m[1] = m_cat(m_flip(ts.w0_m), m_left_div(ts.w0_m, mtile));
    TileState* st0 = launch.states + J.tile0;
    if (wave == 0) {
      // every tile publishes its aggregate (tile 0's too: the last tile of
      // group 0 waits for the aggregates of all the group's tiles)
      if (lane == 0) publish_agg(st0 + lt, a);
      if (!spec) {
        Carry cin;
        if (lt == 0) {
          cin = Carry{0, 0, -1, 1};
          if (lane == 0 && J.nl_cap > 0) J.NL[0] = 0;
        } else {
          cin = lookback_wave(st0, lt, lane, a);
        }
        if (lane == 0) {
          publish_inc(st0 + lt, apply_agg(a, cin));
          s_carry = cin;
          if (J.tile_lines && !dl) J.tile_lines[lt] = lt == 0 ? 0 : cin.lines;
        }
      }
    }
    if (spec) {
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        if (r * BLOCK + tid == ts.fnl) {
          // state at the word's first byte that makes its first '\n' lie
          // outside a string at depth -1
          const int ld = s_lead[r][tid];
          const int pf = ent[r] ? (ent[r] - 2) : 0;
          uint64_t S = s_wc[r][kSlotS0][tid];
          if (pf && (ld & 0x80)) S ^= ~0ull << (ld & 0x7F);
          const int b = ctz64(s_wc[r][kSlotNL][tid]);
          const int sf = (int)((S >> b) & 1);
          if (sf) S = ~S;
          const uint64_t below = (1ull << b) - 1ull;
          const int64_t df = -1 - (int64_t)(popc64(s_wc[r][kSlotO][tid] & ~S & below) -
                                            popc64(s_wc[r][kSlotC][tid] & ~S & below));
          const M4 mx = texcl[r];
          const int sin = sf ^ mx.q;
          s_carry = Carry{0, sin, df - (sin ? (int64_t)(mx.tot - mx.d0) : (int64_t)mx.d0), 0};
        }
      }
    }
    __syncthreads();
    const Carry X = s_carry;
    const int w0 = ts.w0;
    const bool fix = (X.p == 1) && w0 >= 0 && ts.w0_flip;
    const M4 w0m = ts.w0_m;
    if (dl && tid == 0) {
      // deferred bookkeeping: a speculated tile publishes its inclusive prefix
      // now and never looks back (index_finish checks every carry-in against
      // its predecessor's carry-out after the launch)
      if (spec) publish_inc(st0 + lt, apply_agg(a, X));
      store_carry_in(st0 + lt, X, spec);
    }

    // ---- emit the indexes (lane t writes word t of each row: coalesced) ----
    // A tile with few enough lines evaluates the chain query of the launch on
    // every line inside it (FUSE); it then writes its index words only around
    // the lines that cross into its neighbours.
    const bool ft = FUSE && dl && launch.fuse_depth > 0 && mtile.nl <= fused_line_cap(TILE_WORDS);
    const int fnl = ts.fnl, lnl = ts.lnl;
    const bool full = (word0 + TILE_WORDS) * 64 <= J.N;  // every byte of the tile is in the file
    auto emit_tile = [&](const Carry& cx, bool fx) {
      // the carry-in depth held to 32 bits: a tile moves the depth by at most
      // its bytes, so a carry-in beyond +-2^30 leaves every word's depth beyond
      // the clamp in emit_word either way
      const int cxd = (int)(cx.d < -(1ll << 30) ? -(1ll << 30) : (cx.d > (1ll << 30) ? (1ll << 30) : cx.d));
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        const int sidx = r * BLOCK + tid;
        const int64_t w = word0 + sidx;
        M4 mx = texcl[r];
        if (fx && sidx > w0) mx = m_cat(m_flip(w0m), m_left_div(w0m, mx));
        const int p = ent[r] ? (ent[r] - 2) : cx.p;
        const int s = cx.s ^ mx.q;
        const int d = cxd + (cx.s ? mx.tot - mx.d0 : mx.d0);
        const FuseEmit fz{ft, sidx <= fnl || sidx >= lnl, s_fsum + kFuseMaxDepth * NSUM, NSUM};
        emit_word<NLV, VERIFY>(J, w, s_wc, s_lead, r, tid, p, s, d, w < J.W, full, s_lv[wave], s_items[wave], fz);
      }
    };
    emit_tile(X, fix);
    Carry XL = X;  // carry-in whose line count places the newline offsets
    if (spec && !dl) {
      if (wave == 0) {
        const Carry cin = lookback_wave(st0, lt, lane, a);
        if (lane == 0) {
          publish_inc(st0 + lt, apply_agg(a, cin));
          s_exact = cin;
          s_mis = (cin.s != X.s) || (cin.d != X.d);
          if (J.tile_lines) J.tile_lines[lt] = cin.lines;
        }
      }
      __syncthreads();
      XL = s_exact;
      // a tile whose first byte is neither '\\' nor '"' needs no escape fix
      if (s_mis) emit_tile(XL, false);
    }
    if (FUSE && ft) {
      // ---- in-tile chain evaluation ----
      // match summaries start empty; a matching ':' sets its word's bit
      for (int i = tid; i < kFuseMaxDepth * NSUM; i += BLOCK) s_fsum[i] = 0;
      __syncthreads();  // every word's string and leveled words are in LDS
      const int D = launch.fuse_depth;
      const bool tile0 = lt == 0;
      const int64_t pos0 = word0 * 64;
      const FuseTile<ROWS, NSLOT, BLOCK> view{s_wc, s_fsum, s_fsum + kFuseMaxDepth * NSUM, NSUM};
      uint16_t* nlpos = &s_items[0][0];
      uint64_t* cands = &s_lv[0][0][0];  // tile-local ':' position | ss << 16 | level << 32
      constexpr int CCAP = NW * NLVS * 64;
      auto key_of = [&](int l, int& klen) {
        int koff = 0;
#pragma unroll
        for (int k = 0; k < kFuseMaxDepth; k++)
          if (k < l) koff += launch.fuse_klen[k];
        klen = launch.fuse_klen[l];
        return (const uint8_t*)(s_keys + koff);
      };
      // (1) every ':' at a level < D whose key could match (the string,
      // newline and raw-quote words: this word's and the one before it, read
      // once; a key test that needs words further back walks the LDS); its
      // bytes are compared in (2).  The candidates of a wavefront are taken
      // one per lane at a time and appended with one atomic per step.
      const int kl0 = launch.fuse_klen[0], kl1 = launch.fuse_klen[1], kl2 = launch.fuse_klen[2];
      auto note_match = [&](int l, int pc) {
        const int wq = pc >> 6;
        atomicOr(reinterpret_cast<unsigned long long*>(&s_wc[wq / BLOCK][kSlotK][wq % BLOCK]), 1ull << (pc & 63));
        atomicOr(reinterpret_cast<unsigned long long*>(&s_fsum[l * NSUM + (wq >> 6)]), 1ull << (wq & 63));
      };
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        const int sidx = r * BLOCK + tid;
        uint64_t mw = 0;
        uint64_t c0 = 0, c1 = 0, c2 = 0;
        // words before the tile's first newline belong to a line that began in
        // an earlier tile, words after its last to one that runs on: neither
        // is evaluated here (except the file's first and last lines)
        const bool inside = (sidx >= fnl || tile0) && (sidx <= lnl || lt == J.ntiles - 1);
        if (word0 + sidx < J.W && inside) {
          const uint64_t kw = s_wc[r][kSlotK][tid];
          c0 = s_wc[r][fuse_lb_slot(0)][tid] & kw;
          if (D > 1) c1 = s_wc[r][fuse_lb_slot(1)][tid] & kw;
          if (D > 2) c2 = s_wc[r][fuse_lb_slot(2)][tid] & kw;
        }
        uint64_t S1 = 0, NL1 = 0, Q1 = 0, S0 = 0, NL0 = 0, Q0 = 0;
        if (__any((c0 | c1 | c2) != 0)) {
          S1 = s_wc[r][kSlotS0][tid];
          NL1 = s_wc[r][kSlotNL][tid];
          Q1 = s_wc[r][kSlotQ][tid];
          if (sidx >= 1) {
            S0 = view.S(sidx - 1);
            NL0 = view.NL(sidx - 1);
            Q0 = view.Q(sidx - 1);
          }
        }
        // the usual layout (a key's closing quote right before its ':') for
        // every ':' of the word at once; the other ':' (and the tile's first
        // word) take the walk of fused_key_locate
        uint64_t fast = 0, slow = c0 | c1 | c2;
        if (sidx >= 1) {
          const uint64_t cs[3] = {c0, c1, c2};
          const int kls[3] = {kl0, kl1, kl2};
          fused_key_split(cs, D, kls, S1, S0, Q1, Q0, &fast, &slow);
        }
        uint64_t call = fast | slow;
        for (;;) {
          const bool has = call != 0;
          if (!__any(has)) break;
          int ss = -1, pc = 0, l = 0;
          if (has) {
            const int b = ctz64(call);
            const uint64_t bit = 1ull << b;
            l = (c0 & bit) ? 0 : ((c1 & bit) ? 1 : 2);
            const int klen = l == 0 ? kl0 : (l == 1 ? kl1 : kl2);
            if (fast & bit) {
              ss = sidx * 64 + b - 2 - klen;
            } else {
              ss = sidx >= 1 ? fused_key_locate_near(S1, NL1, Q1, S0, NL0, Q0, sidx, b, tile0, klen) : -2;
              if (ss == -2) ss = fused_key_locate(view, sidx * 64 + b, tile0, klen);
            }
            pc = sidx * 64 + b;
            call &= call - 1;
          }
          const unsigned long long m = __ballot(ss >= 0);
          if (m) {
            const int first = __builtin_ctzll(m);
            int base = 0;
            if (lane == first) base = atomicAdd(&ts.ncand, __popcll(m));
            base = __shfl(base, first);
            if (ss >= 0) {
              const int slot = base + __popcll(m & lanes_below);
              if (slot < CCAP) {
                cands[slot] = (uint64_t)pc | ((uint64_t)ss << 16) | ((uint64_t)l << 32);
              } else {  // a tile with more candidates than the list holds compares here
                int kl;
                const uint8_t* key = key_of(l, kl);
                if (fused_key_bytes(J.F, pos0 + ss, key, kl)) {
                  mw |= 1ull << (pc & 63);
                  atomicOr(reinterpret_cast<unsigned long long*>(&s_fsum[l * NSUM + (sidx >> 6)]), 1ull << (sidx & 63));
                }
              }
            }
          }
        }
        s_wc[r][kSlotK][tid] = mw;  // the match word replaces the colon word (its lane's writes end here)
        // the tile's newline positions, in order
        uint64_t nlm = s_wc[r][kSlotNL][tid];
        int k = texcl[r].nl;
        while (nlm) {
          nlpos[k++] = (uint16_t)(sidx * 64 + ctz64(nlm));
          nlm &= nlm - 1;
        }
      }
      __syncthreads();
      // (2) the candidates' key bytes, one candidate per lane: one round trip
      // to memory for the whole tile
      const int ncand = min(ts.ncand, CCAP);
      for (int i = tid; i < ncand; i += BLOCK) {
        const uint64_t e = cands[i];
        const int pc = (int)(e & 0xFFFF), ss = (int)((e >> 16) & 0xFFFF), l = (int)(e >> 32);
        int kl;
        const uint8_t* key = key_of(l, kl);
        if (fused_key_bytes(J.F, pos0 + ss, key, kl)) note_match(l, pc);
      }
      __syncthreads();
      // (3) one lane per line, the lines spread over the wavefronts: slot 0 is
      // the file's first line (tile 0), slot k + 1 the line starting at the
      // tile's newline k; a line ending past the tile (or a tile without
      // newlines) is left to the interpreter.  The result's leading blanks are
      // skipped by the bookkeeping.
      const int c = mtile.nl;
      const bool last = lt == J.ntiles - 1;
      const int nlocal = last ? (int)(J.N - pos0) : TILE_WORDS * 64;
      int64_t* sp = launch.nl_scratch + (J.tile0 + lt) * launch.nl_scratch_cap + fused_span_base(TILE_WORDS);
      for (int base = 0; base <= c; base += BLOCK) {
        const int sl = base + lane * NW + wave;
        if (sl > c) continue;
        const bool valid = sl == 0 ? (tile0 && (c >= 1 || last)) : (sl < c || last);
        if (!valid) continue;
        const int a = sl == 0 ? 0 : nlpos[sl - 1];
        const int b = sl < c ? nlpos[sl] : nlocal;
        int64_t o[2];
        fused_chain_line(view, D, a, b, pos0, o);
        sp[2 * sl] = o[0];
        sp[2 * sl + 1] = o[1];
      }
    }
    // newline offsets: at their slots in the newline index, or (deferred) at
    // tile-local slots of the tile's scratch row
    int64_t* nl_out = dl ? launch.nl_scratch + (J.tile0 + lt) * launch.nl_scratch_cap : J.NL;
    const int64_t nl_lim = dl ? launch.nl_scratch_cap : J.nl_cap;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int sidx = r * BLOCK + tid;
      const int64_t w = word0 + sidx;
      if (w >= J.W) continue;
      uint64_t nlm = s_wc[r][kSlotNL][tid];
      int64_t idx = (dl ? 0 : XL.lines) + texcl[r].nl;
      while (nlm) {
        const int b = ctz64(nlm);
        if (idx < nl_lim) gstore64i(nl_out + idx, w * 64 + b);
        idx++;
        nlm &= nlm - 1;
      }
    }
    if (tid == 0) prepare(par ^ 1);
    __syncthreads();
  }
}

// Zeroes the lookback states and the tickets (a kernel, not hipMemsetAsync, so
// the index pipeline can be captured into a graph).
__global__ void index_reset_kernel(TileState* states, int64_t n, unsigned int* ticket, const int* run_if,
                                   unsigned long long* also) {
  if (run_if && *run_if == 0) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long* p = reinterpret_cast<unsigned long long*>(states);
  for (int64_t k = i; k < n * 8; k += stride) p[k] = 0;
  if (i < kTicketParts) ticket[i * kTicketStride] = 0;
  if (i == 0 && also) *also = 0;
}

// The carry-in a tile stored for the deferred bookkeeping.
__device__ __forceinline__ Carry load_carry_in(const TileState* st, bool* spec) {
  const unsigned long long w = st->pad[2];
  *spec = (w >> 60) & 1;
  return Carry{(int)((w >> 62) & 1), (int)((w >> 61) & 1), sext(w & 0xFFFFFFFFFFFFull, 48), 0};
}

// Deferred newline bookkeeping of one file, in four launches of at most a
// thread per tile (tiles in blocks of kTileBlock).
constexpr int kTileBlock = 256;

// Block-wide exclusive scan of one int per thread (kTileBlock threads);
// returns the block total.
__device__ int tile_block_scan(int v, int* excl) {
  __shared__ int s_w[kTileBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_scan_int(v);
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kTileBlock / 64; w++) {
    if (w < wave) before += s_w[w];
    total += s_w[w];
  }
  *excl = before + incl - v;
  __syncthreads();
  return total;
}

// Block-wide exclusive scan of one int64 per thread (kTileBlock threads;
// Hillis-Steele in LDS); returns the block total.  The block sums of step 2
// need 64 bits: 256 of them may hold more than 2^31 newlines.
__device__ int64_t tile_block_scan64(int64_t v, int64_t* excl) {
  __shared__ int64_t s_v[kTileBlock];
  const int t = threadIdx.x;
  s_v[t] = v;
  __syncthreads();
  for (int off = 1; off < kTileBlock; off <<= 1) {
    const int64_t add = t >= off ? s_v[t - off] : 0;
    __syncthreads();
    s_v[t] += add;
    __syncthreads();
  }
  *excl = s_v[t] - v;
  const int64_t total = s_v[kTileBlock - 1];
  __syncthreads();
  return total;
}

// Step 1: newline counts (from the published aggregates) summed per block.
// Every tile's stored carry-in is checked against its predecessor's carry-out
// (the predecessor's aggregate applied to its carry-in): from tile 0's exact
// one, by induction, this proves every speculated carry-in and every carry-in
// a lookback took from a speculated prefix.  *mismatch gets 1 when one differs
// (a speculation the input broke).
__global__ __launch_bounds__(kTileBlock) void index_tile_sums_kernel(IndexLaunch L, int64_t* block_sums,
                                                                     int* mismatch) {
  if (L.run_if && *L.run_if == 0) return;
  const FileJob& J = L.one;
  const TileState* st = L.states + J.tile0;
  const int64_t t = (int64_t)blockIdx.x * kTileBlock + threadIdx.x;
  int cnt = 0;
  if (t < J.ntiles) {
    cnt = (int)(st[t].agg0 & kField24);
    if (t > 0) {
      bool ps, cs;
      const Carry e = apply_agg(decode_agg(st[t - 1].agg0, st[t - 1].agg1), load_carry_in(st + t - 1, &ps));
      const Carry c = load_carry_in(st + t, &cs);
      if (c.s != e.s || c.d != e.d || (!cs && c.p != e.p)) atomicOr(mismatch, 1);
    }
  }
  int excl;
  const int total = tile_block_scan(cnt, &excl);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Step 2 (one block): exclusive scan of the block sums in place; the line
// count (1 + newlines) and the closing tile_lines entry.
__global__ __launch_bounds__(kTileBlock) void index_tile_base_kernel(IndexLaunch L, int64_t* block_sums,
                                                                     int64_t nb, int64_t* lines_out) {
  if (L.run_if && *L.run_if == 0) return;
  const FileJob& J = L.one;
  if (J.ntiles == 0) {  // empty file: one line starting at 0
    if (threadIdx.x == 0) {
      if (J.nl_cap > 0) J.NL[0] = 0;
      if (J.tile_lines) J.tile_lines[0] = 1;
      *lines_out = 1;
      if (L.fuse_depth > 0 && L.fuse_list_cap > 0) {  // its one line goes to the interpreter
        L.fuse_list[0] = 0;
        *L.fuse_count = 1;
      }
    }
    return;
  }
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += kTileBlock) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = i < nb ? block_sums[i] : 0;
    int64_t excl;
    const int64_t total = tile_block_scan64(v, &excl);
    if (i < nb) block_sums[i] = carry + excl;
    carry += total;
  }
  if (threadIdx.x == 0) {
    if (J.tile_lines) J.tile_lines[J.ntiles] = 1 + carry;
    *lines_out = 1 + carry;
  }
}

// In-tile evaluation: the spans of the lines tile t evaluated, from its
// scratch row to their lines, and the tile's other lines (those starting in it
// that end past it, or all of them when it holds more lines than it
// evaluates) onto the interpreter's list.
__device__ void fuse_place_lines(const IndexLaunch& L, const FileJob& J, int64_t t, int cnt, int64_t before) {
  const int tw = (int)L.nl_scratch_cap;
  const bool tile0 = t == 0, last = t == J.ntiles - 1;
  auto push = [&](int64_t first, int64_t n) {  // lines first .. first + n - 1
    const int64_t k = (int64_t)atomicAdd(L.fuse_count, (unsigned long long)n);
    for (int64_t i = 0; i < n && k + i < L.fuse_list_cap; i++) L.fuse_list[k + i] = first + i;
  };
  if (cnt > fused_line_cap(tw)) {
    if (tile0) push(0, 1);
    if (cnt > 0) push(1 + before, cnt);
    return;
  }
  const int64_t* sp = L.nl_scratch + (J.tile0 + t) * L.nl_scratch_cap + fused_span_base(tw);
  for (int sl = 0; sl <= cnt; sl++) {
    const int64_t line = sl == 0 ? 0 : 1 + before + (sl - 1);
    const bool valid = sl == 0 ? (tile0 && (cnt >= 1 || last)) : (sl < cnt || last);
    if (valid) {
      if (line < J.nl_cap) {
        L.fuse_spans[2 * line] = fused_skip_blanks(J.F, J.N, sp[2 * sl]);
        L.fuse_spans[2 * line + 1] = sp[2 * sl + 1];
      }
    } else if (sl > 0 || tile0) {
      push(line, 1);
    }
  }
}

// Step 3: tile_lines, and each tile's newline offsets from its scratch row to
// their slots (tile t's first newline takes slot 1 + the newlines before it).
// A tile with more newlines than its row holds (lines shorter than 64 bytes on
// average) sets *overflow for step 4.
// Copies the newline offsets of a wavefront's 64 tiles (this lane's tile:
// cnt offsets from src to dst[0 .. min(cnt, room))) with consecutive lanes
// on consecutive offsets: the wavefront walks its tiles' offsets as one run,
// so the reads within a row and the writes (consecutive across the tiles of
// a file) are coalesced.  Every lane of the wavefront takes part.
__device__ void place_tile_newlines(const int64_t* src, int cnt, int64_t* dst, int64_t room) {
  __shared__ int s_incl[kTileBlock];  // inclusive scan of each wavefront's counts
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int* inc = s_incl + wave * 64;
  const int incl = wave_scan_int(cnt);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  inc[lane] = incl;
  wave_lds_sync();
  for (int u0 = 0; u0 < total; u0 += 64) {
    const int u = u0 + lane;
    int lo = 0, hi = 63;  // the lane whose tile holds offset u: first with inc > u
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (inc[mid] > u) hi = mid; else lo = mid + 1;
    }
    const int64_t* s_ = (const int64_t*)shfl_u64((uint64_t)src, lo);
    int64_t* d_ = (int64_t*)shfl_u64((uint64_t)dst, lo);
    const int64_t rm = (int64_t)shfl_u64((uint64_t)room, lo);
    if (u < total) {
      const int k = u - (lo ? inc[lo - 1] : 0);
      if (k < rm) d_[k] = s_[k];
    }
  }
  wave_lds_sync();
}

__global__ __launch_bounds__(kTileBlock) void index_tile_lines_kernel(IndexLaunch L, const int64_t* block_base,
                                                                      int* overflow) {
  if (L.run_if && *L.run_if == 0) return;
  const FileJob& J = L.one;
  const int64_t t = (int64_t)blockIdx.x * kTileBlock + threadIdx.x;
  const int cnt = t < J.ntiles ? (int)(L.states[J.tile0 + t].agg0 & kField24) : 0;
  int excl;
  tile_block_scan(cnt, &excl);
  const int64_t before = block_base[blockIdx.x] + excl;  // newlines before tile t
  const bool in = t < J.ntiles;  // threads past the last tile take part in the placement only
  if (in && J.tile_lines) J.tile_lines[t] = t == 0 ? 0 : 1 + before;
  if (in && L.fuse_depth > 0) fuse_place_lines(L, J, t, cnt, before);
  const bool over = cnt > L.nl_scratch_cap;  // step 4 places this tile's offsets
  if (over) atomicOr(overflow, 1);
  place_tile_newlines(L.nl_scratch + (J.tile0 + t) * L.nl_scratch_cap, over ? 0 : cnt, J.NL + 1 + before,
                      J.nl_cap - 1 - before);
}

// Step 4 (only when *overflow): the offsets of tiles with more newlines than
// their scratch row, scanned again from the input, a wavefront per tile.
__global__ __launch_bounds__(256) void index_tile_overflow_kernel(IndexLaunch L, const int* overflow) {
  if ((L.run_if && *L.run_if == 0) || *overflow == 0) return;
  const FileJob& J = L.one;
  const int lane = threadIdx.x & 63;
  const int64_t cap = L.nl_scratch_cap;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < J.ntiles; t += nwaves) {
    const int64_t cnt = (int64_t)(L.states[J.tile0 + t].agg0 & kField24);
    if (cnt <= cap) continue;
    int64_t slot = t == 0 ? 1 : J.tile_lines[t];
    const int64_t wb = t * cap, we = min(wb + cap, J.W);
    for (int64_t w0 = wb; w0 < we; w0 += 64) {
      const int64_t w = w0 + lane;
      uint64_t m = 0;
      if (w < we)
        for (int b = 0; b < 64; b++) {
          const int64_t o = w * 64 + b;
          if (o < J.N && J.F[o] == '\n') m |= 1ull << b;
        }
      const int c = popc64(m);
      const int incl = wave_scan_int(c);
      int64_t idx = slot + incl - c;
      while (m) {
        const int b = ctz64(m);
        if (idx < J.nl_cap) J.NL[idx] = w * 64 + b;
        idx++;
        m &= m - 1;
      }
      slot += __shfl(incl, 63);
    }
  }
}

// The same four steps for a launch over many files (L.jobs, in tile order:
// file f owns tiles tile0 .. tile0 + ntiles - 1 of the launch).  The block
// sums and their scan run over all tiles; a file's newlines before one of its
// tiles are the scan there minus the scan at the file's first tile.

// The file holding tile t: the last with tile0 <= t (an empty file shares
// tile0 with the file after it, so it is never the last such file of a tile).
__device__ __forceinline__ int file_of_tile(const FileJob* jobs, int64_t njobs, int64_t t) {
  int lo = 0, hi = (int)njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Step 1: as index_tile_sums_kernel; a file's first tile has no predecessor
// to check against.
__global__ __launch_bounds__(kTileBlock) void batch_tile_sums_kernel(IndexLaunch L, int64_t* block_sums,
                                                                     int64_t* file_base, int* mismatch) {
  if (L.run_if && *L.run_if == 0) return;
  const int64_t t = (int64_t)blockIdx.x * kTileBlock + threadIdx.x;
  int cnt = 0;
  int f = -1;  // the file of a file's first tile
  if (t < L.total_tiles) {
    const TileState* st = L.states + t;
    cnt = (int)(st->agg0 & kField24);
    const int ft = file_of_tile(L.jobs, L.njobs, t);
    if (t > L.jobs[ft].tile0) {
      bool ps, cs;
      const Carry e = apply_agg(decode_agg(st[-1].agg0, st[-1].agg1), load_carry_in(st - 1, &ps));
      const Carry c = load_carry_in(st, &cs);
      if (c.s != e.s || c.d != e.d || (!cs && c.p != e.p)) atomicOr(mismatch, 1);
    } else {
      f = ft;
    }
  }
  int excl;
  const int total = tile_block_scan(cnt, &excl);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
  // a file's first tile: the newlines before it within its block (step 2 adds the block's base)
  if (f >= 0) file_base[f] = excl;
}

// Step 2 (one block): exclusive scan of the block sums in place, then per
// file the newlines before its first tile (file_base) and its line count.
__global__ __launch_bounds__(kTileBlock) void batch_tile_base_kernel(IndexLaunch L, int64_t* block_sums, int64_t nb,
                                                                     int64_t* file_base, int64_t* lines_out) {
  if (L.run_if && *L.run_if == 0) return;
  // exclusive scan of the block sums in place: each thread takes a run of
  // consecutive sums, one block scan joins the runs
  const int64_t per = (nb + kTileBlock - 1) / kTileBlock;
  const int64_t i0 = (int64_t)threadIdx.x * per, i1 = min(i0 + per, nb);
  int64_t part = 0;
  for (int64_t i = i0; i < i1; i++) part += block_sums[i];
  int64_t run;
  const int64_t total = tile_block_scan64(part, &run);
  for (int64_t i = i0; i < i1; i++) {
    const int64_t v = block_sums[i];
    block_sums[i] = run;
    run += v;
  }
  __syncthreads();
  // per file, a thread each: the newlines before its first tile (step 1 left
  // the count within the tile's block in file_base) ...
  for (int64_t f = threadIdx.x; f < L.njobs; f += kTileBlock) {
    const FileJob& J = L.jobs[f];
    if (J.ntiles == 0) {  // empty file: one line starting at 0
      if (J.nl_cap > 0) J.NL[0] = 0;
      if (J.tile_lines) J.tile_lines[0] = 1;
      file_base[f] = 0;
      lines_out[f] = 1;
    } else {
      file_base[f] += block_sums[J.tile0 / kTileBlock];
    }
  }
  __syncthreads();
  // ... then its line count: its last tile ends where the next file with
  // tiles starts (files are consecutive in tile order), or at the total
  for (int64_t f = threadIdx.x; f < L.njobs; f += kTileBlock) {
    const FileJob& J = L.jobs[f];
    if (J.ntiles == 0) continue;
    int64_t end = total;
    if (J.tile0 + J.ntiles < L.total_tiles) {
      int64_t g = f + 1;
      while (L.jobs[g].ntiles == 0) g++;
      end = file_base[g];
    }
    const int64_t lines = 1 + end - file_base[f];
    if (J.tile_lines) J.tile_lines[J.ntiles] = lines;
    lines_out[f] = lines;
  }
}

// Step 3: as index_tile_lines_kernel.  A tile that overflows its scratch row
// leaves the newlines before it in the row's first slot for step 4.
__global__ __launch_bounds__(kTileBlock) void batch_tile_lines_kernel(IndexLaunch L, const int64_t* block_base,
                                                                      const int64_t* file_base, int* overflow) {
  if (L.run_if && *L.run_if == 0) return;
  const int64_t t = (int64_t)blockIdx.x * kTileBlock + threadIdx.x;
  const int cnt = t < L.total_tiles ? (int)(L.states[t].agg0 & kField24) : 0;
  int excl;
  tile_block_scan(cnt, &excl);
  int ccnt = 0;  // offsets to copy (threads past the last tile take part in the copy only)
  int64_t* dst = nullptr;
  int64_t room = 0;
  int64_t* row = L.nl_scratch;
  if (t < L.total_tiles) {
    const int f = file_of_tile(L.jobs, L.njobs, t);
    const FileJob& J = L.jobs[f];
    const int64_t lt = t - J.tile0;
    const int64_t before = block_base[blockIdx.x] + excl - file_base[f];  // the file's newlines before tile lt
    if (J.tile_lines) J.tile_lines[lt] = lt == 0 ? 0 : 1 + before;
    row = L.nl_scratch + t * L.nl_scratch_cap;
    if (cnt > L.nl_scratch_cap) {
      row[0] = before;
      atomicOr(overflow, 1);
    } else {
      ccnt = cnt;
      dst = J.NL + 1 + before;
      room = J.nl_cap - 1 - before;
    }
  }
  place_tile_newlines(row, ccnt, dst, room);
}

// Step 4 (only when *overflow): as index_tile_overflow_kernel.
__global__ __launch_bounds__(256) void batch_tile_overflow_kernel(IndexLaunch L, const int* overflow) {
  if ((L.run_if && *L.run_if == 0) || *overflow == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t cap = L.nl_scratch_cap;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < L.total_tiles; t += nwaves) {
    const int64_t cnt = (int64_t)(L.states[t].agg0 & kField24);
    if (cnt <= cap) continue;
    const FileJob& J = L.jobs[file_of_tile(L.jobs, L.njobs, t)];
    const int64_t lt = t - J.tile0;
    int64_t slot = 1 + L.nl_scratch[t * cap];
    const int64_t wb = lt * cap, we = min(wb + cap, J.W);
    for (int64_t w0 = wb; w0 < we; w0 += 64) {
      const int64_t w = w0 + lane;
      uint64_t m = 0;
      if (w < we)
        for (int b = 0; b < 64; b++) {
          const int64_t o = w * 64 + b;
          if (o < J.N && J.F[o] == '\n') m |= 1ull << b;
        }
      const int c = popc64(m);
      const int incl = wave_scan_int(c);
      int64_t idx = slot + incl - c;
      while (m) {
        const int b = ctz64(m);
        if (idx < J.nl_cap) J.NL[idx] = w * 64 + b;
        idx++;
        m &= m - 1;
      }
      slot += __shfl(incl, 63);
    }
  }
}

// Reads the line count of each file from its last tile's inclusive prefix.
__global__ void index_lines_kernel(const FileJob* jobs, FileJob one, int64_t njobs, const TileState* states,
                                   int64_t* lines_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= njobs) return;
  const FileJob& J = jobs ? jobs[i] : one;
  if (J.ntiles == 0) {  // empty file: one line starting at 0
    if (J.nl_cap > 0) J.NL[0] = 0;
    if (J.tile_lines) J.tile_lines[0] = 1;
    lines_out[i] = 1;
    return;
  }
  unsigned long long v = states[J.tile0 + J.ntiles - 1].inc1;
  lines_out[i] = (int64_t)(v & ~kValid);
  if (J.tile_lines) J.tile_lines[J.ntiles] = lines_out[i];
}

// ---------------------------------------------------------------------------
// Host-side launch helpers
// ---------------------------------------------------------------------------

struct IndexConfig {
  int block;
  int wpt;
  int nlv;
};

// Blocks of `fn` the device holds at once (occupancy x CUs).
static int resident_blocks(const void* fn, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess) return 1;
  return cus * (per_cu > 0 ? per_cu : 1) > 0 ? cus * (per_cu > 0 ? per_cu : 1) : 1;
}

// The grid is persistent: at most as many blocks as are resident at once, each
// claiming tiles until none are left, so no block is launched per tile.
template <int BLOCK, int WPT, int NLV, bool VERIFY, bool FUSE>
static hipError_t launch_one(const IndexLaunch& L, int grid, hipStream_t stream) {
  static int resident = 0;  // per kernel variant (same value on every device of a process)
  if (resident == 0)
    resident = resident_blocks(reinterpret_cast<const void*>(&index_kernel<BLOCK, WPT, NLV, VERIFY, FUSE>), BLOCK);
  if (grid > resident) grid = resident;
  hipLaunchKernelGGL((index_kernel<BLOCK, WPT, NLV, VERIFY, FUSE>), dim3(grid), dim3(BLOCK), 0, stream, L);
  return hipGetLastError();
}

template <int BLOCK, int WPT, int NLV>
static hipError_t launch_cfg(const IndexLaunch& L, bool verify, int grid, hipStream_t stream) {
  if (verify) return launch_one<BLOCK, WPT, NLV, true, false>(L, grid, stream);
  return launch_one<BLOCK, WPT, NLV, false, false>(L, grid, stream);
}

// In-tile chain evaluation keeps levels 0..L-1 (L <= 3) in LDS: variants with
// 2 and 3 level words.
template <int BLOCK, int WPT>
static hipError_t launch_nlv(const IndexLaunch& L, int nlv, bool verify, int grid, hipStream_t s) {
  if (L.fuse_depth > 0) {
    if (verify || L.one.L > kFuseMaxDepth || !L.nl_scratch) return hipErrorInvalidValue;
    if (L.one.L <= 2) return launch_one<BLOCK, WPT, 2, false, true>(L, grid, s);
    return launch_one<BLOCK, WPT, 3, false, true>(L, grid, s);
  }
  switch (nlv) {
    case 2: return launch_cfg<BLOCK, WPT, 2>(L, verify, grid, s);
    case 4: return launch_cfg<BLOCK, WPT, 4>(L, verify, grid, s);
    default: return launch_cfg<BLOCK, WPT, 8>(L, verify, grid, s);
  }
}

int64_t index_tile_bytes(int block, int wpt) { return (int64_t)block * wpt * 64; }
// smallest supported tile: 128 x 2 words = 16 KiB (same as 256 x 1)

bool index_config_supported(int block, int wpt) {
  return (block == 256 && (wpt >= 1 && wpt <= 4)) || (block == 512 && (wpt == 1 || wpt == 2)) ||
         (block == 128 && (wpt >= 2 && wpt <= 4));
}

hipError_t index_launch(const IndexLaunch& L, int block, int wpt, int nlv, bool verify, int grid,
                        hipStream_t stream) {
  if (block == 256 && wpt == 1) return launch_nlv<256, 1>(L, nlv, verify, grid, stream);
  if (block == 256 && wpt == 2) return launch_nlv<256, 2>(L, nlv, verify, grid, stream);
  if (block == 256 && wpt == 3) return launch_nlv<256, 3>(L, nlv, verify, grid, stream);
  if (block == 256 && wpt == 4) return launch_nlv<256, 4>(L, nlv, verify, grid, stream);
  if (block == 512 && wpt == 1) return launch_nlv<512, 1>(L, nlv, verify, grid, stream);
  if (block == 512 && wpt == 2) return launch_nlv<512, 2>(L, nlv, verify, grid, stream);
  if (block == 128 && wpt == 2) return launch_nlv<128, 2>(L, nlv, verify, grid, stream);
  if (block == 128 && wpt == 3) return launch_nlv<128, 3>(L, nlv, verify, grid, stream);
  if (block == 128 && wpt == 4) return launch_nlv<128, 4>(L, nlv, verify, grid, stream);
  return hipErrorInvalidValue;
}

hipError_t index_reset(TileState* states, int64_t ntiles, unsigned int* ticket, hipStream_t stream,
                       const int* run_if, unsigned long long* also) {
  int64_t n8 = ntiles * 8;
  int grid = (int)((n8 + 255) / 256);
  if (grid < 1) grid = 1;
  // a conditional reset (the fallback pass) nearly always returns at once,
  // and a launch's cost then is its workgroups' dispatch
  if (grid > (run_if ? 256 : 4096)) grid = run_if ? 256 : 4096;
  hipLaunchKernelGGL(index_reset_kernel, dim3(grid), dim3(256), 0, stream, states, ntiles, ticket, run_if, also);
  return hipGetLastError();
}

// Steps 1-4 over the files of a batched launch; block_sums holds one int64
// per kTileBlock tiles of the launch, file_base and lines_out one per file.
hipError_t index_finish_batch(const IndexLaunch& L, int64_t* block_sums, int64_t* file_base, int64_t* lines_out,
                              int* mismatch, int* overflow, hipStream_t stream) {
  const int64_t nt = L.total_tiles;
  const int64_t nb = (nt + kTileBlock - 1) / kTileBlock;
  const unsigned grid = (unsigned)(nb > 0 ? nb : 1);
  if (nt > 0)
    hipLaunchKernelGGL(batch_tile_sums_kernel, dim3(grid), dim3(kTileBlock), 0, stream, L, block_sums, file_base,
                       mismatch);
  hipLaunchKernelGGL(batch_tile_base_kernel, dim3(1), dim3(kTileBlock), 0, stream, L, block_sums, nb, file_base,
                     lines_out);
  if (nt > 0) {
    hipLaunchKernelGGL(batch_tile_lines_kernel, dim3(grid), dim3(kTileBlock), 0, stream, L,
                       (const int64_t*)block_sums, (const int64_t*)file_base, overflow);
    const unsigned ogrid = (unsigned)std::min<int64_t>((nt + 3) / 4, 256);  // nearly always returns at once
    hipLaunchKernelGGL(batch_tile_overflow_kernel, dim3(ogrid), dim3(256), 0, stream, L, (const int*)overflow);
  }
  return hipGetLastError();
}

// Steps 1-4 of the deferred newline bookkeeping; block_sums holds one int64
// per kTileBlock tiles.
hipError_t index_finish(const IndexLaunch& L, int64_t* block_sums, int64_t* lines_out, int* mismatch,
                        int* overflow, hipStream_t stream) {
  const int64_t nt = L.one.ntiles;
  const int64_t nb = (nt + kTileBlock - 1) / kTileBlock;
  const unsigned grid = (unsigned)(nb > 0 ? nb : 1);
  if (nt > 0)
    hipLaunchKernelGGL(index_tile_sums_kernel, dim3(grid), dim3(kTileBlock), 0, stream, L, block_sums, mismatch);
  hipLaunchKernelGGL(index_tile_base_kernel, dim3(1), dim3(kTileBlock), 0, stream, L, block_sums, nb, lines_out);
  if (nt > 0) {
    hipLaunchKernelGGL(index_tile_lines_kernel, dim3(grid), dim3(kTileBlock), 0, stream, L,
                       (const int64_t*)block_sums, overflow);
    const unsigned ogrid = (unsigned)std::min<int64_t>((nt + 3) / 4, 256);  // nearly always returns at once
    hipLaunchKernelGGL(index_tile_overflow_kernel, dim3(ogrid), dim3(256), 0, stream, L, (const int*)overflow);
  }
  return hipGetLastError();
}

hipError_t index_lines(const IndexLaunch& L, int64_t* lines_out, hipStream_t stream) {
  int grid = (int)((L.njobs + 255) / 256);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(index_lines_kernel, dim3(grid), dim3(256), 0, stream, L.jobs, L.one, L.njobs, L.states,
                     lines_out);
  return hipGetLastError();
}

}  // namespace gj
