/*
 * gpjson_oracle.c -- CPU restatement of GPJSON's NDJSON index + query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library (gpjson_amd) never links or calls it.
 *
 * Every function restates the reference's sequential semantics (file:line
 * below are relative to the reference repository; J/ = gpjson/src/main/java/
 * com/koenv/gpjson/, K/ = gpjson/src/main/resources/com/koenv/gpjson/kernels/,
 * T/ = gpjson/src/test/java/com/koenv/gpjson/).
 *
 * Definitions this oracle fixes where the reference is undefined or limited
 * (the GPU path implements the same):
 *   - all offsets are 64-bit (reference: int, files < 2 GiB);
 *   - bitmap bits at positions >= n are 0 (reference string kernel can set
 *     them when the file ends inside a string);
 *   - newline_index[0] = 0 is written explicitly (reference relies on zeroed
 *     cudaMalloc memory);
 *   - up to GJO_MAX_LEVELS (64) levels (reference: 16);
 *   - result spans are -1 for every query before it runs (reference resets the
 *     buffer once for all queries, J/functions/QueryFunction.java:197);
 *   - MOVE_TO_INDEX on a value whose first non-blank byte is not '[' ends the
 *     line's evaluation (reference: device assert, K/find_value.cu:151-154);
 *   - blank skipping stops at n (reference reads past the end);
 *   - MOVE_UP out of a level whose end was never found (its end is -1: the
 *     level-end search of MOVE_DOWN skips the byte after the cursor, so
 *     `[0:k]` over an empty array misses the ']') ends the line's evaluation
 *     (reference: the cursor becomes -1 and the walk restarts at byte 0 of the
 *     file, reading bitmap word -1 and f[-1]; K/find_value.cu:81-88).
 *
 * Parity pins (tests/test_oracle.py): the reference's fixtures and closed forms
 * (T/stages/StringIndexTest.java:109-168, NewlineIndexTest.java:81-108,
 * LeveledBitmapsIndexTest.java:94-155), the known-answer queries of
 * T/function/QueryFunctionTest.java:86-198, brute-force JSON-structure
 * invariants on generated lines, and MT == single-thread on random inputs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define GJO_MAX_LEVELS 64

enum { OP_END = 0, OP_STORE = 1, OP_UP = 2, OP_DOWN = 3, OP_KEY = 4, OP_INDEX = 5, OP_STREQ = 6 };

static inline int64_t words_of(int64_t n) { return (n + 63) / 64; }
static inline int is_blank(uint8_t c) { return c == '\n' || c == '\r' || c == '\t' || c == ' '; }
static inline int get_bit(const uint64_t *m, int64_t i) { return (int)((m[i >> 6] >> (i & 63)) & 1u); }

/* J/sequential/Sequential.java:14-34 -- [0] ++ positions of every '\n'.
 * Returns the line count (1 + number of '\n'); writes at most cap entries. */
int64_t gjo_newline_index(const uint8_t *f, int64_t n, int64_t *out, int64_t cap) {
    int64_t k = 0;
    if (cap > 0) out[0] = 0;
    k = 1;
    for (int64_t i = 0; i < n; i++) {
        if (f[i] != '\n') continue;
        if (k < cap) out[k] = i;
        k++;
    }
    return k;
}

/* T/stages/StringIndexTest.java:109-127 -- escape bitmap: E[i] = c_i,
 * c_0 = 0, c_{i+1} = F[i]=='\\' ? 1-c_i : 0. */
void gjo_escape_index(const uint8_t *f, int64_t n, uint64_t *out) {
    memset(out, 0, (size_t)words_of(n) * 8);
    int c = 0;
    for (int64_t i = 0; i < n; i++) {
        if (c) out[i >> 6] |= 1ull << (i & 63);
        c = (f[i] == '\\') ? (c ^ 1) : 0;
    }
}

/* T/stages/StringIndexTest.java:129-147 -- unescaped quotes: Q[i] = F[i]=='"' && !E[i]. */
void gjo_quote_index(const uint8_t *f, int64_t n, uint64_t *out) {
    memset(out, 0, (size_t)words_of(n) * 8);
    int c = 0;
    for (int64_t i = 0; i < n; i++) {
        if (f[i] == '"' && !c) out[i >> 6] |= 1ull << (i & 63);
        c = (f[i] == '\\') ? (c ^ 1) : 0;
    }
}

/* J/sequential/Sequential.java:36-64 -- in-string bitmap; the opening quote is
 * inside the string, the closing quote is not. */
void gjo_string_index(const uint8_t *f, int64_t n, uint64_t *out) {
    memset(out, 0, (size_t)words_of(n) * 8);
    int escaped = 0, in_string = 0;
    for (int64_t i = 0; i < n; i++) {
        if (f[i] == '"' && escaped != 1) in_string = !in_string;
        if (in_string) out[i >> 6] |= 1ull << (i & 63);
        escaped = (f[i] == '\\') ? (escaped ^ 1) : 0;
    }
}

/* J/sequential/Sequential.java:66-128 (== K/create_leveled_bitmaps.cu:26-54):
 * depth starts at -1 for the whole file; '{' '[' increment then mark, '}' ']'
 * mark then decrement, ':' ',' mark; only outside strings and for 0 <= level < L.
 * Layout: out[level * W + i/64]. */
static void leveled_range(const uint8_t *f, int64_t lo, int64_t hi, const uint64_t *S, int L,
                          int64_t W, int64_t level, uint64_t *out) {
    for (int64_t i = lo; i < hi; i++) {
        if (get_bit(S, i)) continue;
        uint8_t v = f[i];
        if (v == '{' || v == '[') {
            level++;
            if (level >= 0 && level < L) out[level * W + (i >> 6)] |= 1ull << (i & 63);
        } else if (v == '}' || v == ']') {
            if (level >= 0 && level < L) out[level * W + (i >> 6)] |= 1ull << (i & 63);
            level--;
        } else if ((v == ':' || v == ',') && level >= 0 && level < L) {
            out[level * W + (i >> 6)] |= 1ull << (i & 63);
        }
    }
}

void gjo_leveled_index(const uint8_t *f, int64_t n, const uint64_t *S, int L, uint64_t *out) {
    int64_t W = words_of(n);
    memset(out, 0, (size_t)(W * L) * 8);
    leveled_range(f, 0, n, S, L, W, -1, out);
}

/* IR varint, J/jsonpath/IRByteOutputBuffer.java:40-49 (LEB128, read as in
 * K/find_value.cu:111-118). */
static int32_t read_varint(const uint8_t *ir, int64_t ir_len, int64_t *pc) {
    int32_t v = 0;
    int shift = 0;
    while (*pc < ir_len) {
        uint8_t b = ir[(*pc)++];
        v |= (int32_t)(b & 0x7F) << shift;
        if (!(b & 0x80)) break;
        shift += 7;
        if (shift > 28) break;
    }
    return v;
}

static inline int lb_bit(const uint64_t *LB, int L, int64_t W, int level, int64_t j) {
    if (level < 0 || level >= L) return 0;
    return (int)((LB[(int64_t)level * W + (j >> 6)] >> (j & 63)) & 1u);
}

/* Per-line interpreter, K/find_value.cu:21-263 == J/sequential/Sequential.java:146-337.
 * Evaluates lines [l0, l1).  out holds lines*R*2 int64 spans; the caller
 * pre-fills it with -1. */
static void find_value_lines(const uint8_t *f, int64_t n, const int64_t *nl, int64_t lines,
                             const uint64_t *S, const uint64_t *LB, int L,
                             const uint8_t *ir, int64_t ir_len, int R, int64_t *out,
                             int64_t l0, int64_t l1) {
    int64_t W = words_of(n);
    int64_t le[GJO_MAX_LEVELS + 1];
    int64_t ci[GJO_MAX_LEVELS + 1];
    for (int64_t i = l0; i < l1; i++) {
        int64_t start = nl[i];
        int64_t end = (i + 1 < lines) ? nl[i + 1] : n;
        int lvl = 0;
        le[0] = end;
        for (int k = 1; k <= GJO_MAX_LEVELS; k++) le[k] = -1;
        for (int k = 0; k <= GJO_MAX_LEVELS; k++) ci[k] = -1;
        int64_t pc = 0;
        int r = 0;
        int op = OP_END;
        int32_t want_len = 0, want_idx = 0;
        const uint8_t *want = NULL;
        int exec = 1;
        for (int64_t j = start; lvl >= 0 && j < le[lvl] && j < n; j++) {
            while (exec) {
                if (pc >= ir_len) goto line_done;
                op = ir[pc++];
                switch (op) {
                case OP_END:
                    goto line_done;
                case OP_STORE: {
                    while (j < n && is_blank(f[j])) j++;
                    if (r < R) {
                        int64_t *o = out + ((int64_t)i * R + r) * 2;
                        o[0] = j;
                        o[1] = le[lvl];
                    }
                    r++;
                    break;
                }
                case OP_UP:
                    j = le[lvl];
                    le[lvl] = -1;
                    lvl--;
                    if (lvl < 0 || j < 0) goto line_done;
                    break;
                case OP_DOWN:
                    lvl++;
                    if (lvl > GJO_MAX_LEVELS) goto line_done;
                    if (le[lvl] == -1) {
                        for (int64_t m = j + 1; m < le[lvl - 1]; m++) {
                            if (lb_bit(LB, L, W, lvl - 1, m)) { le[lvl] = m; break; }
                        }
                    }
                    break;
                case OP_KEY:
                    want_len = read_varint(ir, ir_len, &pc);
                    want = ir + pc;
                    pc += want_len;
                    exec = 0;
                    break;
                case OP_INDEX:
                    want_idx = read_varint(ir, ir_len, &pc);
                    if (ci[lvl] == -1) {
                        while (j < n && f[j] != '[') {
                            if (!is_blank(f[j])) goto line_done; /* not an array */
                            j++;
                        }
                        if (j >= n) goto line_done;
                        ci[lvl] = 0;
                    }
                    if (want_idx == 0) { j++; exec = 1; } else { exec = 0; }
                    break;
                case OP_STREQ: {
                    want_len = read_varint(ir, ir_len, &pc);
                    want = ir + pc;
                    pc += want_len;
                    while (j < n && is_blank(f[j])) j++;
                    int64_t actual = le[lvl] - j;
                    if (actual != want_len) goto line_done;
                    for (int64_t k = 0; k < want_len; k++)
                        if (want[k] != f[j + k]) goto line_done;
                    break;
                }
                default:
                    goto line_done;
                }
            }
            if (!(j < n)) break;
            int structural = lb_bit(LB, L, W, lvl, j);
            if (op == OP_KEY) {
                if (structural && f[j] == ':') {
                    int64_t se = -1;
                    for (int64_t k = j - 1; k > start; k--) {
                        if (get_bit(S, k)) { se = k; break; }
                    }
                    if (se < 0) continue;
                    int64_t ss = se - want_len;
                    if (ss < start || f[ss] != '"') continue;
                    exec = 1;
                    for (int64_t k = 0; k < want_len; k++) {
                        if (want[k] != f[ss + k + 1]) { exec = 0; break; }
                    }
                }
            } else if (op == OP_INDEX) {
                if (structural && f[j] == ',') {
                    ci[lvl]++;
                    if (want_idx == ci[lvl]) exec = 1;
                }
            }
        }
    line_done:;
    }
}

/* Whole-file evaluation of one query: out = lines*R*2 spans, pre-filled -1. */
void gjo_find_value(const uint8_t *f, int64_t n, const int64_t *nl, int64_t lines,
                    const uint64_t *S, const uint64_t *LB, int L,
                    const uint8_t *ir, int64_t ir_len, int R, int64_t *out) {
    for (int64_t k = 0; k < lines * R * 2; k++) out[k] = -1;
    find_value_lines(f, n, nl, lines, S, LB, L, ir, ir_len, R, out, 0, lines);
}

/* ------------------------------------------------------------------------- */
/* Multithreaded CPU evaluator (the reported CPU baseline).  Same outputs as  */
/* the functions above; chunks of 64-byte multiples with a serial combine of  */
/* per-chunk escape / quote-parity / depth / newline summaries.              */
/* ------------------------------------------------------------------------- */

typedef struct {
    const uint8_t *f;
    int64_t n, lo, hi;
    /* phase A results */
    int p_out[2];      /* escape carry out for carry in 0/1 */
    int q_par[2];      /* unescaped quote parity for carry in 0/1 */
    int64_t nl_count;
    /* carries */
    int p_in, s_in;
    int64_t d_in, nl_base;
    int64_t d_delta;
    /* shared outputs */
    uint64_t *S, *LB;
    int64_t *NL;
    int L;
    int phase;
} mt_chunk;

static void chunk_phase_a(mt_chunk *c) {
    const uint8_t *f = c->f;
    int64_t cnt = 0;
    for (int pin = 0; pin < 2; pin++) {
        int e = pin, q = 0;
        for (int64_t i = c->lo; i < c->hi; i++) {
            uint8_t v = f[i];
            if (v == '"' && !e) q ^= 1;
            e = (v == '\\') ? (e ^ 1) : 0;
            if (pin == 0 && v == '\n') cnt++;
        }
        c->p_out[pin] = e;
        c->q_par[pin] = q;
    }
    c->nl_count = cnt;
}

static void chunk_phase_b(mt_chunk *c) {
    const uint8_t *f = c->f;
    int e = c->p_in, in_str = c->s_in;
    int64_t k = c->nl_base, d = 0;
    for (int64_t i = c->lo; i < c->hi; i++) {
        uint8_t v = f[i];
        if (v == '"' && !e) in_str = !in_str;
        if (in_str) c->S[i >> 6] |= 1ull << (i & 63);
        else if (v == '{' || v == '[') d++;
        else if (v == '}' || v == ']') d--;
        e = (v == '\\') ? (e ^ 1) : 0;
        if (v == '\n') c->NL[k++] = i;
    }
    c->d_delta = d;
}

static void chunk_phase_c(mt_chunk *c) {
    int64_t W = words_of(c->n);
    leveled_range(c->f, c->lo, c->hi, c->S, c->L, W, c->d_in, c->LB);
}

static void *chunk_worker(void *arg) {
    mt_chunk *c = (mt_chunk *)arg;
    if (c->phase == 0) chunk_phase_a(c);
    else if (c->phase == 1) chunk_phase_b(c);
    else chunk_phase_c(c);
    return NULL;
}

typedef struct {
    const uint8_t *f;
    int64_t n;
    const int64_t *nl;
    int64_t lines;
    const uint64_t *S, *LB;
    int L;
    const uint8_t *ir;
    int64_t ir_len;
    int R;
    int64_t *out;
    int64_t l0, l1;
} mt_lines;

static void *lines_worker(void *arg) {
    mt_lines *w = (mt_lines *)arg;
    for (int64_t k = w->l0 * w->R * 2; k < w->l1 * w->R * 2; k++) w->out[k] = -1;
    find_value_lines(w->f, w->n, w->nl, w->lines, w->S, w->LB, w->L, w->ir, w->ir_len, w->R,
                     w->out, w->l0, w->l1);
    return NULL;
}

static void run_chunks(mt_chunk *cs, int T, int phase) {
    pthread_t th[256];
    for (int t = 0; t < T; t++) {
        cs[t].phase = phase;
        pthread_create(&th[t], NULL, chunk_worker, &cs[t]);
    }
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
}

/*
 * Builds the newline, string and leveled indexes with nthreads threads.
 * nl_out must hold nl_cap entries; returns the number of lines (if it exceeds
 * nl_cap, nothing is written past nl_cap and the caller retries), or -1.
 */
int64_t gjo_index_mt(const uint8_t *f, int64_t n, int L, int nthreads,
                     int64_t *nl_out, int64_t nl_cap, uint64_t *S, uint64_t *LB) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    int64_t W = words_of(n);
    int64_t words_per = (W + nthreads - 1) / nthreads;
    if (words_per < 1) words_per = 1;
    int T = (int)((W + words_per - 1) / words_per);
    if (T < 1) T = 1;
    mt_chunk *cs = (mt_chunk *)calloc((size_t)T, sizeof(mt_chunk));
    if (!cs) return -1;
    for (int t = 0; t < T; t++) {
        cs[t].f = f; cs[t].n = n; cs[t].L = L;
        cs[t].lo = (int64_t)t * words_per * 64;
        cs[t].hi = cs[t].lo + words_per * 64;
        if (cs[t].hi > n) cs[t].hi = n;
        if (cs[t].lo > n) cs[t].lo = n;
        cs[t].S = S; cs[t].LB = LB; cs[t].NL = nl_out;
    }
    run_chunks(cs, T, 0);
    int p = 0, s = 0;
    int64_t lines = 1;
    for (int t = 0; t < T; t++) {
        cs[t].p_in = p; cs[t].s_in = s; cs[t].nl_base = lines;
        s ^= cs[t].q_par[p];
        p = cs[t].p_out[p];
        lines += cs[t].nl_count;
    }
    if (lines > nl_cap) { free(cs); return lines; }
    memset(S, 0, (size_t)W * 8);
    memset(LB, 0, (size_t)(W * L) * 8);
    nl_out[0] = 0;
    run_chunks(cs, T, 1);
    int64_t d = -1;
    for (int t = 0; t < T; t++) { cs[t].d_in = d; d += cs[t].d_delta; }
    run_chunks(cs, T, 2);
    free(cs);
    return lines;
}

/* Per-line query evaluation with nthreads threads; out pre-filled with -1 per query. */
void gjo_find_value_mt(const uint8_t *f, int64_t n, const int64_t *nl, int64_t lines,
                       const uint64_t *S, const uint64_t *LB, int L,
                       const uint8_t *ir, int64_t ir_len, int R, int64_t *out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    mt_lines ws[256];
    int64_t per = (lines + nthreads - 1) / nthreads;
    int T = 0;
    for (int t = 0; t < nthreads; t++) {
        int64_t l0 = (int64_t)t * per, l1 = l0 + per;
        if (l1 > lines) l1 = lines;
        if (l0 >= l1) break;
        ws[t] = (mt_lines){f, n, nl, lines, S, LB, L, ir, ir_len, R, out, l0, l1};
        pthread_create(&th[t], NULL, lines_worker, &ws[t]);
        T++;
    }
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
}
