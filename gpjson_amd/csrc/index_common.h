// index_common.h -- host/device shared definitions of the structural index.
#pragma once
#include <stdint.h>

namespace gj {

// Levels the interpreter supports (reference: 16, J/stages/LeveledBitmapsIndex.java:16).
constexpr int kMaxLevels = 64;

// Per-tile lookback state: self-validating 64-bit words (bit 63 = written), so a
// reader never needs a fence between a flag and its payload.  The tile's
// aggregate, its inclusive prefix and (for the last tile of each group of 64
// tiles) the group's aggregate are each two words; a reader takes a pair once
// both words are valid.  One 64-byte line per tile.
struct alignas(64) TileState {
  unsigned long long agg0, agg1;  // the tile's aggregate (packed counts, see the index kernel)
  unsigned long long inc0;        // valid | p(62) | s(61) | depth (signed 48 bits)
  unsigned long long inc1;        // valid | lines (63 bits)
  unsigned long long pad[4];      // [0..1]: group aggregate (last tile of a group)
};

#if defined(__HIPCC__)
#define GJ_REC_HD __host__ __device__ __forceinline__
#else
#define GJ_REC_HD static inline
#endif

// The string, colon, raw-quote and level-0 words of a 64-byte input word form
// one 32-byte record, so an interpreter step at level 0 touches one cache line
// for all four (four records per 128-byte line).  Levels 1 .. L-1 of the
// leveled bitmaps are level-major arrays of Wp words each; most of a member
// chain's walk is at level 0, and a step at a deeper level adds one line.
constexpr int kRecWords = 4;
constexpr int kRecLevel0 = 3;  // word of the record holding leveled bitmap 0
GJ_REC_HD uint64_t rec_s(const uint64_t* R, int64_t w) { return R[w * kRecWords]; }      // in-string bits
GJ_REC_HD uint64_t rec_colon(const uint64_t* R, int64_t w) { return R[w * kRecWords + 1]; }  // ':' bytes
GJ_REC_HD uint64_t rec_quote(const uint64_t* R, int64_t w) { return R[w * kRecWords + 2]; }  // '"' bytes

// Level summaries: bit k of summary word s of level l is set iff leveled word
// 64*s + k of level l is not zero, so a scan over a level can skip 64 words
// (4 KB of input) with one load where the level has no structure.  They follow
// levels 1 .. L-1 in the same allocation, Ws words per level (L >= 1 levels).
GJ_REC_HD int64_t summary_stride(int64_t Wp) { return (Wp + 63) / 64; }
// Leveled words of levels 1 .. L-1 in the allocation (level 0 is in the records).
GJ_REC_HD int64_t deep_level_words(int64_t Wp, int L) { return Wp * (L > 1 ? L - 1 : 0); }
// Words of a leveled-bitmap allocation (levels 1 .. L-1, then the summaries).
GJ_REC_HD int64_t leveled_alloc_words(int64_t Wp, int L) {
  return deep_level_words(Wp, L) + summary_stride(Wp) * (L > 1 ? L : 1);
}
// Start of the summaries in a leveled-bitmap allocation.
GJ_REC_HD uint64_t* summaries_of(uint64_t* LB, int64_t Wp, int L) { return LB + deep_level_words(Wp, L); }

// One file of an index launch.  The plain path launches with one file.
struct FileJob {
  const uint8_t* F;     // input bytes
  int64_t N;            // bytes
  int64_t W;            // words = ceil(N/64)
  int64_t Wp;           // level stride (W rounded up to 8)
  uint64_t* rec;        // records (string, colon, raw-quote, level-0 words), kRecWords * W words
  uint64_t* LB;         // leveled bitmaps of levels 1 .. L-1, (L - 1) * Wp words
  uint64_t* LS;         // level summaries, L * Ws words (may be null)
  int64_t Ws;           // summary stride
  int64_t* NL;          // newline index, capacity nl_cap
  int64_t nl_cap;
  uint64_t* E;          // escape index (verify mode, may be null)
  uint64_t* Q;          // quote index (verify mode, may be null)
  int64_t* tile_lines;  // optional: first line id starting in each tile, [ntiles + 1]
  int64_t tile0;        // first global tile id of this file
  int64_t ntiles;
  int L;                // levels to write
  int pad_;
};

// In-tile evaluation limits: chains up to 3 keys (the tile's LDS holds three
// level words per input word), keys up to 120 bytes together.
constexpr int kFuseMaxDepth = 3;
constexpr int kFuseKeyBytes = 120;

// Tile tickets of an index launch: kTicketParts counters, kTicketStride
// 32-bit words (256 bytes) apart, in kTicketBytes at IndexLaunch.ticket.
// Block b of a grid of at least kTicketParts blocks claims the tiles
// p, p + P, p + 2P, ... (p = b % P, P = kTicketParts) from counter p.  With
// P = 8 the launch ran 4 % faster alone (one address's atomics take turns),
// but its progress then needs a running block of every class, and a class
// can wait for CUs that other kernels hold (a server builds indexes beside
// its query launches); a test of that mix hung once, with no cause shown, so
// P stays 1: one counter, and any running block keeps the launch going.
constexpr int kTicketParts = 1;
constexpr int kTicketStride = 64;
constexpr int kTicketBytes = kTicketParts * kTicketStride * 4;

struct IndexLaunch {
  const FileJob* jobs;     // device array, or null: the single job `one`
  FileJob one;
  int64_t njobs;
  int64_t total_tiles;     // tiles [tile_begin, total_tiles) are this launch's
  int64_t tile_begin;      // earlier tiles were indexed by earlier launches
  TileState* states;       // total_tiles entries, zeroed before the first launch
  unsigned int* ticket;    // kTicketBytes: kTicketParts tile tickets, zeroed before launch
  int speculate;           // tiles holding a newline may speculate their carry-in
  // Deferred newline bookkeeping (one file): each tile writes its newline
  // offsets to nl_scratch[tile * nl_scratch_cap + k] and its carry-in to its
  // state; index_finish then builds tile_lines, the line count and the
  // newline index, and checks every speculated carry-in.  Null:
  // the tiles' lookback gives offsets and line counts directly.
  int64_t* nl_scratch;
  int64_t nl_scratch_cap;  // slots per tile (the tile's words)
  const int* run_if;       // when set, the launch does nothing unless *run_if != 0
  // In-tile evaluation of one member chain (deferred bookkeeping, one file;
  // fuse_depth = 0: off).  A tile holding at most fused_line_cap(tile words)
  // newlines evaluates the chain (keys k_0 .. k_{D-1}, concatenated in
  // fuse_keys, lengths fuse_klen) for every line that starts and ends inside
  // it and leaves the spans in its newline scratch row; it writes its records
  // and leveled words only where a line crosses into a neighbouring tile (the
  // words up to its first newline and from its last one).  The bookkeeping
  // then places those spans in fuse_spans (int64[lines][2]) and lists every
  // other line in fuse_list (capacity fuse_list_cap, count *fuse_count) for
  // the interpreter.
  int fuse_depth;
  int fuse_klen[kFuseMaxDepth];
  uint8_t fuse_keys[kFuseKeyBytes];
  int64_t* fuse_spans;
  int64_t* fuse_list;
  int64_t fuse_list_cap;
  unsigned long long* fuse_count;  // zeroed with the ticket
};

}  // namespace gj
