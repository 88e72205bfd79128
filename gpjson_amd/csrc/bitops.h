// bitops.h -- word-level (64 input bytes per 64-bit word) bit arithmetic of the
// structural index.  Header-only, compiled both by hipcc for gfx950 and by g++
// for the CPU harness that checks these functions exhaustively.
//
// Bit i%64 of word i/64 stands for input byte i (LSB first), as in the
// reference's bitmaps (K/create_combined_escape_newline_index.cu:34-36,
// K/create_string_index.cu:16-30; K/ = gpjson/src/main/resources/com/koenv/
// gpjson/kernels/).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GJ_HD __host__ __device__ __forceinline__
#else
#define GJ_HD static inline
#endif

namespace gj {

constexpr uint64_t kEven = 0x5555555555555555ull;  // bits at even positions
constexpr uint64_t kOdd = 0xAAAAAAAAAAAAAAAAull;

GJ_HD int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popcll(x);
#else
  return __builtin_popcountll(x);
#endif
}

GJ_HD int ctz64(uint64_t x) {  // x != 0
#if defined(__HIP_DEVICE_COMPILE__)
  return __ffsll((unsigned long long)x) - 1;
#else
  return __builtin_ctzll(x);
#endif
}

GJ_HD int clz64(uint64_t x) {  // x != 0
#if defined(__HIP_DEVICE_COMPILE__)
  return __clzll(x);
#else
  return __builtin_clzll(x);
#endif
}

// Escape bitmap of one word for escape carry-in p (reference definition
// T/stages/StringIndexTest.java:109-127: E[i] = c_i, c_{i+1} = F[i]=='\\' ?
// 1-c_i : 0).  Backslash runs are resolved with one add per run parity class:
// a run starting at bit s (not continued from the previous word) marks the
// positions s+1, s+3, ... up to and including the byte after the run; a run
// continued from the previous word with carry 1 marks s, s+2, ... instead.
GJ_HD uint64_t escape_word(uint64_t B, uint32_t p) {
  uint64_t starts = B & ~((B << 1) | (uint64_t)p);
  uint64_t es = starts & kEven, os = starts & kOdd;
  uint64_t E = (((B + es) ^ B) & kOdd) | (((B + os) ^ B) & kEven);
  if (p) E |= ((B + 1) ^ B) & kEven;
  return E;
}

// Escape carry into the next word: an all-backslash word passes the carry
// through (64 is even); otherwise it is the parity of the trailing run.
GJ_HD uint32_t escape_carry_out(uint64_t B, uint32_t p) {
  if (B == ~0ull) return p;
  uint64_t nb = ~B;
  int run = clz64(nb);  // leading (high) ones of B
  return (uint32_t)(run & 1);
}

// Inclusive prefix XOR (bit i = XOR of bits 0..i), done on 32-bit halves.
GJ_HD uint64_t prefix_xor(uint64_t q) {
  uint32_t lo = (uint32_t)q, hi = (uint32_t)(q >> 32);
  lo ^= lo << 1; lo ^= lo << 2; lo ^= lo << 4; lo ^= lo << 8; lo ^= lo << 16;
  hi ^= hi << 1; hi ^= hi << 2; hi ^= hi << 4; hi ^= hi << 8; hi ^= hi << 16;
  hi ^= (uint32_t)((int32_t)lo >> 31);
  return ((uint64_t)hi << 32) | lo;
}

// Bytes of one word classified into 64-bit masks.
struct WordClass {
  uint64_t B;   // '\\'
  uint64_t Q;   // '"' (escaped or not)
  uint64_t NL;  // '\n'
  uint64_t O;   // '{' '['
  uint64_t C;   // '}' ']'
  uint64_t P;   // ':' ','
  uint64_t K;   // ':'
};

#if defined(__HIP_DEVICE_COMPILE__)
GJ_HD uint32_t perm_b32(uint32_t s0, uint32_t s1, uint32_t sel) {
  return __builtin_amdgcn_perm(s0, s1, sel);
}
#else
// v_perm_b32: result byte i = byte sel_i of the 8-byte value {s0:s1} (0..3 from s1).
GJ_HD uint32_t perm_b32(uint32_t s0, uint32_t s1, uint32_t sel) {
  uint64_t v = ((uint64_t)s0 << 32) | s1;
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) {
    uint32_t s = (sel >> (8 * i)) & 0xFF;
    r |= (uint32_t)((v >> (8 * s)) & 0xFF) << (8 * i);
  }
  return r;
}
#endif

// 4x4 byte transpose: out[k] = byte k of a0, a1, a2, a3 (in that order).
GJ_HD void transpose4x4_bytes(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t out[4]) {
  const uint32_t u0 = perm_b32(a1, a0, 0x05010400u);  // a0.0 a1.0 a0.1 a1.1
  const uint32_t u1 = perm_b32(a1, a0, 0x07030602u);  // a0.2 a1.2 a0.3 a1.3
  const uint32_t u2 = perm_b32(a3, a2, 0x05010400u);
  const uint32_t u3 = perm_b32(a3, a2, 0x07030602u);
  out[0] = perm_b32(u2, u0, 0x05040100u);
  out[1] = perm_b32(u2, u0, 0x07060302u);
  out[2] = perm_b32(u3, u1, 0x05040100u);
  out[3] = perm_b32(u3, u1, 0x07060302u);
}

// One butterfly stage of the bit-plane extraction: registers a and b differ
// in one bit of the register index, and the bits of a at positions with bit S
// of the position set trade places with the bits of b at positions with it
// clear (m selects positions with bit S clear, SH = 2^S).  Afterwards that
// register-index bit and position bit have exchanged their meaning.
template <int SH>
GJ_HD void plane_swap(uint32_t& a, uint32_t& b, uint32_t m) {
  const uint32_t na = (a & m) | ((b << SH) & ~m);
  const uint32_t nb = ((a >> SH) & m) | (b & ~m);
  a = na;
  b = nb;
}

// A mask constant the optimizer cannot see through (device code: an SGPR).
GJ_HD uint32_t opaque_mask(uint32_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
  // an SGPR whose value the optimizer does not see: it would otherwise move
  // the stages' masks across stages and split bitfield inserts into and/or
  asm("" : "=s"(m) : "0"(m));
#endif
  return m;
}

// Bit planes of 32 bytes (d[0..7], byte b of d[i] = position 4i+b): bit j of
// pl[k] is bit k of byte j.  A bit of byte j (j = j4..j0) starts in register
// j >> 2 at position 8*(j & 3) + k.  Two 4x4 byte transposes put byte
// j = t + 8s in slot s of register t, so the position reads (j4 j3 k2 k1 k0)
// and the register index (j2 j1 j0); three butterfly stages then exchange k0,
// k1, k2 with j0, j1, j2, which leaves bit k of byte j in register k at
// position j.  (48 shifts and bitfield inserts and 16 byte permutes per 32
// bytes, against 84 + 16 for an 8x8 bit transpose of each 8-byte group.)
GJ_HD void bit_planes32(const uint32_t d[8], uint32_t pl[8]) {
  uint32_t u[8];
  transpose4x4_bytes(d[0], d[2], d[4], d[6], u);
  transpose4x4_bytes(d[1], d[3], d[5], d[7], u + 4);
  const uint32_t m1 = opaque_mask(0x55555555u), m2 = opaque_mask(0x33333333u), m4 = opaque_mask(0x0F0F0F0Fu);
#pragma unroll
  for (int t = 0; t < 8; t += 2) plane_swap<1>(u[t], u[t + 1], m1);
#pragma unroll
  for (int t = 0; t < 8; t++)
    if (!(t & 2)) plane_swap<2>(u[t], u[t + 2], m2);
#pragma unroll
  for (int t = 0; t < 4; t++) plane_swap<4>(u[t], u[t + 4], m4);
#pragma unroll
  for (int k = 0; k < 8; k++) pl[k] = u[k];
}

// Classifies 32 bytes into 32-bit masks from their bit planes: each class is a
// conjunction over the 8 bits of a byte, split into high-nibble and low-nibble
// terms shared between classes.  Bytes from 0x80 up (bit 7 set) match nothing.
//   '\\' 0x5C  '"' 0x22  '\n' 0x0A  '{' 0x7B '[' 0x5B  '}' 0x7D ']' 0x5D  ':' 0x3A  ',' 0x2C
GJ_HD void classify32(const uint32_t d[8], uint32_t& B, uint32_t& Q, uint32_t& NL, uint32_t& O,
                      uint32_t& C, uint32_t& P, uint32_t& K) {
  uint32_t p[8];
  bit_planes32(d, p);
  const uint32_t a = ~p[7] & ~p[6];
  const uint32_t h0 = a & ~p[5] & ~p[4];          // 0x0_
  const uint32_t h2 = a & p[5] & ~p[4];           // 0x2_
  const uint32_t h3 = a & p[5] & p[4];            // 0x3_
  const uint32_t h57 = ~p[7] & p[6] & p[4];       // 0x5_ or 0x7_
  const uint32_t h5 = h57 & ~p[5];                // 0x5_
  const uint32_t l1 = p[3] & ~p[2] & p[1];        // 0x_A or 0x_B
  const uint32_t l2 = p[3] & p[2] & ~p[1];        // 0x_C or 0x_D
  const uint32_t lA = l1 & ~p[0], lB = l1 & p[0], lC = l2 & ~p[0], lD = l2 & p[0];
  const uint32_t l_2 = ~p[3] & ~p[2] & p[1] & ~p[0];  // 0x_2
  B = h5 & lC;
  Q = h2 & l_2;
  NL = h0 & lA;
  O = h57 & lB;
  C = h57 & lD;
  K = h3 & lA;
  P = K | (h2 & lC);
}

// Classifies one 64-byte word given as 16 little-endian dwords.
GJ_HD WordClass classify64(const uint32_t d[16]) {
  uint32_t b0, q0, n0, o0, c0, p0, k0, b1, q1, n1, o1, c1, p1, k1;
  classify32(d, b0, q0, n0, o0, c0, p0, k0);
  classify32(d + 8, b1, q1, n1, o1, c1, p1, k1);
  WordClass w;
  w.B = ((uint64_t)b1 << 32) | b0;
  w.Q = ((uint64_t)q1 << 32) | q0;
  w.NL = ((uint64_t)n1 << 32) | n0;
  w.O = ((uint64_t)o1 << 32) | o0;
  w.C = ((uint64_t)c1 << 32) | c0;
  w.P = ((uint64_t)p1 << 32) | p0;
  w.K = ((uint64_t)k1 << 32) | k0;
  return w;
}

// Leveled-bitmap bits of one word for levels [lo, lo+NLV): depth d (the depth
// before the word's first byte, -1 at file start) follows the reference rule
// (J/sequential/Sequential.java:84-100): '{' '[' increment then mark, '}' ']'
// mark then decrement, ':' ',' mark at the current depth.  O, C, P are already
// restricted to bytes outside strings.  Returns the depth after the word.
// D: the depth's integer type (the kernel passes a depth clamped to 32 bits).
template <int NLV, class D>
GJ_HD D leveled_word(uint64_t O, uint64_t C, uint64_t P, D d, int lo, int L, uint64_t out[NLV]) {
#pragma unroll
  for (int k = 0; k < NLV; k++) out[k] = 0;
  uint64_t T = O | C;
  uint64_t rem = ~0ull;  // positions not yet assigned
  while (T) {
    int t = ctz64(T);
    uint64_t bit = 1ull << t;
    uint64_t before = rem & (bit - 1);  // separators strictly before t
    uint64_t seps = P & before;
    int is_open = (O & bit) != 0;
    D lvl_open = d + 1;
    // separators before t and a closing bracket at t sit at depth d; an
    // opening bracket at t sits at depth d + 1.
    uint64_t at_d = seps | (is_open ? 0 : bit);
    uint64_t at_d1 = is_open ? bit : 0;
#pragma unroll
    for (int k = 0; k < NLV; k++) {
      D lv = lo + k;
      if (lv < L) {
        out[k] |= (lv == d) ? at_d : 0;
        out[k] |= (lv == lvl_open) ? at_d1 : 0;
      }
    }
    d += is_open ? 1 : -1;
    rem &= ~((bit - 1) | bit);
    T &= T - 1;
  }
  uint64_t seps = P & rem;
#pragma unroll
  for (int k = 0; k < NLV; k++) {
    D lv = lo + k;
    if (lv < L) out[k] |= (lv == d) ? seps : 0;
  }
  return d;
}

// The same bits, one bracket at a time, with no dependence between brackets
// (so the brackets of many words can be spread over the lanes of a wavefront).
// The bracket at bit t sits at level d + (opens at or before t) - (closes
// before t); the separators after it, up to the next bracket, sit at the depth
// after it (that level for an open, one less for a close).  Returns those
// separators.
template <class D>
GJ_HD uint64_t bracket_step(uint64_t O, uint64_t C, uint64_t P, D d, int t, D* lvl, D* after) {
  const uint64_t bit = 1ull << t;
  const uint64_t le = bit | (bit - 1);
  const D l = d + popc64(O & le) - popc64(C & (bit - 1));
  *lvl = l;
  *after = (O & bit) ? l : l - 1;
  const uint64_t rest = (O | C) & ~le;
  const uint64_t seg = rest ? (((rest & (0 - rest)) - 1) & ~le) : ~le;
  return P & seg;
}

// Separators before the word's first bracket (all of them if it has none):
// they sit at depth d.
GJ_HD uint64_t leading_seps(uint64_t O, uint64_t C, uint64_t P) {
  const uint64_t T = O | C;
  return T ? (P & ((T & (0 - T)) - 1)) : P;
}

}  // namespace gj
