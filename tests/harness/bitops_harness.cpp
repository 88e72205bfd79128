// CPU build of gpjson_amd/csrc/bitops.h so the word-level bit arithmetic of the
// index kernel can be checked exhaustively without a GPU.
#include "bitops.h"

extern "C" {
uint64_t bh_escape_word(uint64_t B, uint32_t p) { return gj::escape_word(B, p); }
uint32_t bh_escape_carry_out(uint64_t B, uint32_t p) { return gj::escape_carry_out(B, p); }
uint64_t bh_prefix_xor(uint64_t q) { return gj::prefix_xor(q); }
void bh_classify64(const uint8_t* bytes, uint64_t* out) {
  uint32_t d[16];
  for (int i = 0; i < 16; i++)
    d[i] = (uint32_t)bytes[4 * i] | ((uint32_t)bytes[4 * i + 1] << 8) | ((uint32_t)bytes[4 * i + 2] << 16) |
           ((uint32_t)bytes[4 * i + 3] << 24);
  gj::WordClass w = gj::classify64(d);
  out[0] = w.B; out[1] = w.Q; out[2] = w.NL; out[3] = w.O; out[4] = w.C; out[5] = w.P; out[6] = w.K;
}
int64_t bh_leveled_word(uint64_t O, uint64_t C, uint64_t P, int64_t d, int lo, int L, uint64_t* out) {
  return gj::leveled_word<8>(O, C, P, d, lo, L, out);
}
// leveled bits assembled bracket by bracket (bracket_step, leading_seps), in
// the order the lanes of a wavefront would add them
void bh_leveled_by_brackets(uint64_t O, uint64_t C, uint64_t P, int64_t d, int lo, int L, uint64_t* out) {
  for (int k = 0; k < 8; k++) out[k] = 0;
  auto put = [&](int64_t level, uint64_t bits) {
    const int64_t k = level - lo;
    if (k >= 0 && k < 8 && level < L) out[k] |= bits;
  };
  put(d, gj::leading_seps(O, C, P));
  uint64_t T = O | C;
  while (T) {
    const int t = __builtin_ctzll(T);
    int64_t lvl, after;
    const uint64_t seps = gj::bracket_step(O, C, P, d, t, &lvl, &after);
    put(lvl, 1ull << t);
    put(after, seps);
    T &= T - 1;
  }
}
}
