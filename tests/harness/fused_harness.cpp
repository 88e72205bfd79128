// CPU build of the in-tile member-chain evaluation (gpjson_amd/csrc/fused_eval.h):
// the file is cut into tiles of tw 64-byte words as the index kernel cuts it,
// each tile's match words, summaries and lines are built the way the kernel
// builds them, and every line that the kernel would evaluate inside its tile
// gets its spans here.  The lines it leaves to the interpreter (they cross a
// tile boundary, or their tile holds more lines than the fused cap) are listed
// by the same rule the newline bookkeeping uses.
#include <cstddef>
#include <cstdint>
#include <vector>

#include "fused_eval.h"

namespace {

struct TileView {
  const uint64_t *pS, *pNL, *pQ, *LBg, *Mw;
  int64_t w0, W, Wstride;
  int L;
  const std::vector<uint64_t>* sm;  // [D][tw / 64]
  const std::vector<uint64_t>* sl;
  int nsum;
  uint64_t at(const uint64_t* a, int w) const { return (w0 + w < W) ? a[w0 + w] : 0; }
  uint64_t S(int w) const { return at(pS, w); }
  uint64_t NL(int w) const { return at(pNL, w); }
  uint64_t Q(int w) const { return at(pQ, w); }
  uint64_t LB(int l, int w) const { return (w0 + w < W) ? LBg[(int64_t)l * Wstride + w0 + w] : 0; }
  uint64_t M(int w) const { return Mw[w]; }
  uint64_t SM(int l, int s) const { return (*sm)[(size_t)l * nsum + s]; }
  uint64_t SL(int l, int s) const { return (*sl)[(size_t)l * nsum + s]; }
};

}  // namespace

extern "C" {
// out: int64[lines][2], filled for the lines evaluated in a tile (done[line] =
// 1); crossing: the lines left to the interpreter, in tile order.  Returns the
// number of crossing lines, -1 when a line is claimed twice, or -2 when the
// register fast path of the key test disagrees with the full walk.  *near
// gets the number of key tests the fast path decided.
int64_t fe_run(const uint8_t* F, int64_t N, const uint64_t* S, const uint64_t* LB, int L, const uint64_t* NLm,
               const uint64_t* Q, const uint64_t* K, int tw, int D, const uint8_t* keys, const int* klens,
               int64_t lines, int64_t* out, uint8_t* done, int64_t* crossing, int64_t* near) {
  const int64_t W = (N + 63) / 64;
  const int64_t tile_bytes = (int64_t)tw * 64;
  const int64_t ntiles = (N + tile_bytes - 1) / tile_bytes;
  const int cap = gj::fused_line_cap(tw);
  const int nsum = (tw + 63) / 64;
  std::vector<int> koff(D + 1, 0);
  for (int l = 0; l < D; l++) koff[l + 1] = koff[l] + klens[l];
  int64_t before = 0, ncross = 0, near_decided = 0;
  if (ntiles == 0) {  // the empty file: its one line goes to the interpreter
    crossing[0] = 0;
    return 1;
  }
  std::vector<uint64_t> M(tw), sm((size_t)D * nsum), sl((size_t)D * nsum);
  for (int64_t t = 0; t < ntiles; t++) {
    const int64_t w0 = t * tw;
    const int64_t pos0 = t * tile_bytes;
    const bool tile0 = t == 0, last = t == ntiles - 1;
    // newlines of the tile, in order
    std::vector<int> nlpos;
    for (int w = 0; w < tw; w++) {
      uint64_t m = (w0 + w < W) ? NLm[w0 + w] : 0;
      while (m) {
        nlpos.push_back(w * 64 + gj::ctz64(m));
        m &= m - 1;
      }
    }
    const int c = (int)nlpos.size();
    const bool fused = c <= cap;
    if (!fused) {
      if (tile0) crossing[ncross++] = 0;
      for (int k = 0; k < c; k++) crossing[ncross++] = 1 + before + k;
      before += c;
      continue;
    }
    TileView v{S, NLm, Q, LB, M.data(), w0, W, W, L, &sm, &sl, nsum};
    // match words and summaries, the way the kernel finds them: the usual
    // layout for the whole word at once, then the register walk, then the
    // full walk; each shortcut is checked against the full walk
    int kl3[3] = {0, 0, 0};
    for (int l = 0; l < D && l < 3; l++) kl3[l] = klens[l];
    for (int w = 0; w < tw; w++) {
      uint64_t mw = 0;
      const uint64_t kw = (w0 + w < W) ? K[w0 + w] : 0;
      uint64_t c[3] = {0, 0, 0};
      for (int l = 0; l < D && l < L && l < 3; l++) c[l] = v.LB(l, w) & kw;
      uint64_t fast = 0, slow = c[0] | c[1] | c[2];
      if (w >= 1) gj::fused_key_split(c, D, kl3, v.S(w), v.S(w - 1), v.Q(w), v.Q(w - 1), &fast, &slow);
      for (int l = 0; l < D && l < L; l++) {
        uint64_t cand = v.LB(l, w) & kw;
        while (cand) {
          const int b = gj::ctz64(cand);
          const uint64_t bit = 1ull << b;
          const uint8_t* key = keys + koff[l];
          const int full = gj::fused_key_locate(v, w * 64 + b, tile0, klens[l]);
          const bool want = full >= 0 && gj::fused_key_bytes(F, pos0 + full, key, klens[l]);
          int ss;
          if (fast & bit) {
            // the key bytes decide what the walk's line-start check would
            ss = w * 64 + b - 2 - klens[l];
            near_decided++;
          } else if (slow & bit) {
            ss = -2;
            if (w >= 1) {
              ss = gj::fused_key_locate_near(v.S(w), v.NL(w), v.Q(w), v.S(w - 1), v.NL(w - 1), v.Q(w - 1), w, b,
                                             tile0, klens[l]);
              if (ss != -2 && ss != full) return -2;
            }
            if (ss == -2) ss = full;
          } else {
            ss = -1;
          }
          const bool got = ss >= 0 && gj::fused_key_bytes(F, pos0 + ss, key, klens[l]);
          if (got != want) return -2;  // a shortcut changed the key test's outcome
          if (got) mw |= bit;
          cand &= cand - 1;
        }
      }
      M[w] = mw;
    }
    for (int l = 0; l < D; l++)
      for (int s = 0; s < nsum; s++) {
        uint64_t a = 0, b = 0;
        for (int k = 0; k < 64 && s * 64 + k < tw; k++) {
          const int w = s * 64 + k;
          if (v.LB(l, w)) b |= 1ull << k;
          if (M[w] & v.LB(l, w)) a |= 1ull << k;
        }
        sm[(size_t)l * nsum + s] = a;
        sl[(size_t)l * nsum + s] = b;
      }
    const int nlocal = (int)((last ? N : pos0 + tile_bytes) - pos0);
    // slots: 0 = the file's first line (tile 0), k + 1 = the line starting at newline k
    for (int s = 0; s <= c; s++) {
      const bool valid = (s == 0) ? (tile0 && (c >= 1 || last)) : (s < c || last);
      int64_t line;
      if (s == 0) line = 0;
      else line = 1 + before + (s - 1);
      if (!valid) {
        if (s == 0 ? tile0 : true) crossing[ncross++] = line;
        continue;
      }
      const int a = s == 0 ? 0 : nlpos[s - 1];
      const int b = s < c ? nlpos[s] : nlocal;
      if (line >= lines || done[line]) return -1;
      gj::fused_chain_line(v, D, a, b, pos0, out + line * 2);
      out[line * 2] = gj::fused_skip_blanks(F, N, out[line * 2]);
      done[line] = 1;
    }
    before += c;
  }
  *near = near_decided;
  return ncross;
}
}
