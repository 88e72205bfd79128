/*
 * gen.c -- seeded synthetic NDJSON generators for the benchmark and test inputs.
 *
 * kind 0 ("tweets"): one social-media-post-shaped object per line.  Line length
 *   is lognormal (mean ~2.5 KB, sigma 0.5), ~30 top-level keys, nesting up to
 *   depth 6, a "user" object on every line holding a 2-char ASCII "lang" and
 *   "stats": {"followers": int32, heavy-tailed (discrete Pareto, exponent 1.1)},
 *   "id" int64 uniform in [1e17, 1e18], "place" null on 70% of lines and
 *   otherwise an object with "country_code".  String values: 2% carry \" or \\
 *   escapes, 1% \uXXXX, 5% multi-byte UTF-8.
 * kind 1 ("adversarial"): lines of 4-64 KB with long strings holding backslash
 *   runs of length 1-9 (odd runs before escaped quotes, even runs before
 *   closing quotes) placed across 64-byte, 4-KiB and 32/64-KiB boundaries,
 *   objects nested to depth 32, a 20-member chain "chain.c1...c19", and
 *   "arr": an array of 1000 small objects.
 *
 * Output is exactly `size` bytes of newline-terminated lines; the input is cut
 * into independent 8 MiB chunks generated in parallel, each seeded from
 * (seed, chunk index), so the bytes depend only on (kind, seed, size).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t s[2];
} rng_t;

static uint64_t splitmix(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t x = seed;
  r->s[0] = splitmix(&x);
  r->s[1] = splitmix(&x);
}

static uint64_t rng_u64(rng_t *r) { /* xorshift128+ */
  uint64_t s1 = r->s[0];
  const uint64_t s0 = r->s[1];
  r->s[0] = s0;
  s1 ^= s1 << 23;
  r->s[1] = s1 ^ s0 ^ (s1 >> 17) ^ (s0 >> 26);
  return r->s[1] + s0;
}

static double rng_unit(rng_t *r) { return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static uint32_t rng_below(rng_t *r, uint32_t n) { return (uint32_t)(rng_unit(r) * n); }

static double rng_normal(rng_t *r) {
  double u1 = rng_unit(r), u2 = rng_unit(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct {
  char *p;
  size_t n, cap;
} sbuf;

static void put(sbuf *b, const char *s, size_t n) {
  if (b->n + n > b->cap) {
    size_t c = b->cap ? b->cap * 2 : 4096;
    while (c < b->n + n) c *= 2;
    b->p = (char *)realloc(b->p, c);
    b->cap = c;
  }
  memcpy(b->p + b->n, s, n);
  b->n += n;
}
static void puts_(sbuf *b, const char *s) { put(b, s, strlen(s)); }
static void putc_(sbuf *b, char c) { put(b, &c, 1); }
static void put_u64(sbuf *b, uint64_t v) {
  char t[24];
  int k = 0;
  do { t[k++] = (char)('0' + v % 10); v /= 10; } while (v);
  while (k) putc_(b, t[--k]);
}

static const char *WORDS[] = {"the", "data", "gpu", "json", "query", "index", "stream", "model", "light",
                              "river", "north", "cloud", "green", "fast", "city", "night", "music", "game",
                              "news", "today", "love", "world", "time", "people", "happy", "great"};
#define NWORDS (sizeof(WORDS) / sizeof(WORDS[0]))
static const char *UTF8[] = {"\xc3\xa9", "\xc3\xbc", "\xe6\x97\xa5\xe6\x9c\xac", "\xd0\xbf\xd1\x80\xd0\xb8",
                             "\xf0\x9f\x98\x80", "\xe2\x82\xac"};
static const char *LANGS[] = {"en", "nl", "ja", "es", "pt", "fr", "de", "ar", "ko", "tr", "it", "ru"};
static const char *CC[] = {"US", "NL", "JP", "BR", "GB", "FR", "DE", "ES", "IN", "MX"};

/* A JSON string value of about `len` content bytes, with the tweet-kind
 * escape / unicode / UTF-8 mix. */
static void tweet_string(sbuf *b, rng_t *r, int len) {
  putc_(b, '"');
  double u = rng_unit(r);
  int mode = u < 0.02 ? 1 : (u < 0.03 ? 2 : (u < 0.08 ? 3 : 0));
  int w = 0;
  while (w < len) {
    const char *wd = WORDS[rng_below(r, NWORDS)];
    size_t l = strlen(wd);
    put(b, wd, l);
    w += (int)l;
    if (w < len) { putc_(b, ' '); w++; }
    if (mode == 1 && rng_below(r, 4) == 0) { puts_(b, rng_below(r, 2) ? "\\\"" : "\\\\"); w += 2; }
    if (mode == 2 && rng_below(r, 4) == 0) {
      char t[8];
      static const char hex[] = "0123456789abcdef";
      t[0] = '\\'; t[1] = 'u';
      for (int i = 0; i < 4; i++) t[2 + i] = hex[rng_below(r, 16)];
      put(b, t, 6);
      w += 6;
    }
    if (mode == 3 && rng_below(r, 3) == 0) { const char *m = UTF8[rng_below(r, 6)]; puts_(b, m); w += (int)strlen(m); }
  }
  putc_(b, '"');
}

static void key(sbuf *b, const char *k) {
  putc_(b, '"');
  puts_(b, k);
  puts_(b, "\":");
}

static uint32_t pareto_followers(rng_t *r) {
  double u = rng_unit(r);
  if (u < 1e-12) u = 1e-12;
  double x = pow(u, -1.0 / 0.1); /* tail exponent s - 1 = 0.1 */
  if (x > 2147483647.0) x = 2147483647.0;
  return (uint32_t)x - 1u;
}

static void tweet_line(sbuf *b, rng_t *r, size_t target) {
  size_t start = b->n;
  puts_(b, "{");
  key(b, "created_at"); puts_(b, "\"Mon Oct 17 11:57:27 +0000 2016\",");
  uint64_t id = 100000000000000000ull + rng_u64(r) % 900000000000000001ull;
  key(b, "id"); put_u64(b, id); putc_(b, ',');
  key(b, "id_str"); putc_(b, '"'); put_u64(b, id); puts_(b, "\",");
  size_t text_at = b->n; /* the text field is sized last */
  key(b, "source"); tweet_string(b, r, 20 + (int)rng_below(r, 30)); putc_(b, ',');
  key(b, "truncated"); puts_(b, rng_below(r, 2) ? "true," : "false,");
  key(b, "in_reply_to_status_id"); puts_(b, "null,");
  key(b, "in_reply_to_user_id"); puts_(b, "null,");
  key(b, "in_reply_to_screen_name"); puts_(b, "null,");
  key(b, "user"); putc_(b, '{');
  key(b, "id"); put_u64(b, rng_u64(r) % 4000000000ull); putc_(b, ',');
  key(b, "name"); tweet_string(b, r, 8 + (int)rng_below(r, 12)); putc_(b, ',');
  key(b, "screen_name"); tweet_string(b, r, 6 + (int)rng_below(r, 8)); putc_(b, ',');
  key(b, "location"); tweet_string(b, r, (int)rng_below(r, 20)); putc_(b, ',');
  key(b, "description"); tweet_string(b, r, 20 + (int)rng_below(r, 120)); putc_(b, ',');
  key(b, "protected"); puts_(b, "false,");
  key(b, "verified"); puts_(b, rng_below(r, 20) ? "false," : "true,");
  key(b, "stats"); putc_(b, '{');
  key(b, "followers"); put_u64(b, pareto_followers(r)); putc_(b, ',');
  key(b, "friends"); put_u64(b, rng_below(r, 5000)); putc_(b, ',');
  key(b, "statuses"); put_u64(b, rng_below(r, 100000)); puts_(b, "},");
  key(b, "created_at"); puts_(b, "\"Tue Mar 01 08:21:33 +0000 2011\",");
  key(b, "lang"); putc_(b, '"'); puts_(b, LANGS[rng_below(r, 12)]); puts_(b, "\",");
  key(b, "profile"); putc_(b, '{');
  key(b, "colors"); putc_(b, '{');
  key(b, "background"); puts_(b, "\"C0DEED\",");
  key(b, "link"); puts_(b, "\"1DA1F2\"},");
  key(b, "image"); tweet_string(b, r, 30 + (int)rng_below(r, 20)); puts_(b, "},");
  key(b, "time_zone"); puts_(b, "null");
  puts_(b, "},");
  key(b, "geo"); puts_(b, "null,");
  key(b, "coordinates"); puts_(b, "null,");
  key(b, "place");
  if (rng_unit(r) < 0.7) {
    puts_(b, "null,");
  } else {
    putc_(b, '{');
    key(b, "id"); tweet_string(b, r, 16); putc_(b, ',');
    key(b, "full_name"); tweet_string(b, r, 10 + (int)rng_below(r, 20)); putc_(b, ',');
    key(b, "country_code"); putc_(b, '"'); puts_(b, CC[rng_below(r, 10)]); puts_(b, "\",");
    key(b, "bounding_box"); putc_(b, '{');
    key(b, "type"); puts_(b, "\"Polygon\",");
    key(b, "coordinates"); puts_(b, "[[[-0.5,51.2],[0.3,51.2],[0.3,51.7],[-0.5,51.7]]]}},");
  }
  key(b, "contributors"); puts_(b, "null,");
  key(b, "is_quote_status"); puts_(b, "false,");
  key(b, "retweet_count"); put_u64(b, rng_below(r, 10000)); putc_(b, ',');
  key(b, "favorite_count"); put_u64(b, rng_below(r, 10000)); putc_(b, ',');
  key(b, "entities"); putc_(b, '{');
  key(b, "hashtags"); putc_(b, '[');
  int nh = (int)rng_below(r, 4);
  for (int h = 0; h < nh; h++) {
    if (h) putc_(b, ',');
    putc_(b, '{');
    key(b, "text"); tweet_string(b, r, 4 + (int)rng_below(r, 10)); putc_(b, ',');
    key(b, "indices"); puts_(b, "[");
    put_u64(b, rng_below(r, 100)); putc_(b, ','); put_u64(b, rng_below(r, 140)); puts_(b, "]}");
  }
  puts_(b, "],");
  key(b, "urls"); puts_(b, "[],");
  key(b, "user_mentions"); putc_(b, '[');
  int nm = (int)rng_below(r, 3);
  for (int m = 0; m < nm; m++) {
    if (m) putc_(b, ',');
    putc_(b, '{');
    key(b, "screen_name"); tweet_string(b, r, 8); putc_(b, ',');
    key(b, "meta"); putc_(b, '{');
    key(b, "ext"); putc_(b, '{');
    key(b, "k"); puts_(b, "[1,2,3]}}}");
  }
  puts_(b, "]},");
  key(b, "favorited"); puts_(b, "false,");
  key(b, "retweeted"); puts_(b, "false,");
  key(b, "filter_level"); puts_(b, "\"low\",");
  key(b, "possibly_sensitive"); puts_(b, "false,");
  key(b, "extra"); putc_(b, '[');
  int ne = (int)rng_below(r, 6);
  for (int e = 0; e < ne; e++) {
    if (e) putc_(b, ',');
    tweet_string(b, r, 10 + (int)rng_below(r, 40));
  }
  puts_(b, "],");
  key(b, "lang"); putc_(b, '"'); puts_(b, LANGS[rng_below(r, 12)]); puts_(b, "\",");
  key(b, "timestamp_ms"); puts_(b, "\"1476705447658\"");
  puts_(b, "}\n");
  /* insert the text field after id_str, sized to reach the target length */
  size_t have = b->n - start;
  int text_len = (int)(target > have + 16 ? target - have - 10 : 10);
  sbuf t = {0, 0, 0};
  key(&t, "text");
  tweet_string(&t, r, text_len);
  putc_(&t, ',');
  size_t tail = b->n - text_at;
  put(b, t.p, t.n); /* grow */
  memmove(b->p + text_at + t.n, b->p + text_at, tail);
  memcpy(b->p + text_at, t.p, t.n);
  free(t.p);
}

/* String content with backslash runs of length 1..9: odd runs precede an
 * escaped quote, even runs are escaped backslashes; runs are aligned to end
 * right at or just across multiples of 64 when `align` is set. */
static void adversarial_string(sbuf *b, rng_t *r, int len, size_t line_base) {
  putc_(b, '"');
  int w = 0;
  while (w < len) {
    int kind = (int)rng_below(r, 4);
    if (kind == 0) {
      /* odd run + quote, straddling a 64-byte boundary */
      int run = 1 + 2 * (int)rng_below(r, 5); /* 1,3,5,7,9 */
      size_t pos = b->n - line_base;
      int to_edge = (int)(64 - (pos % 64));
      int lead = to_edge - run / 2;
      if (lead < 0) lead += 64;
      for (int i = 0; i < lead && w < len; i++, w++) putc_(b, (char)('a' + rng_below(r, 26)));
      for (int i = 0; i < run; i++) putc_(b, '\\');
      putc_(b, '"');
      w += run + 1;
    } else if (kind == 1) {
      int run = 2 * (1 + (int)rng_below(r, 4)); /* 2,4,6,8 */
      size_t pos = b->n - line_base;
      int to_edge = (int)(64 - (pos % 64));
      int lead = to_edge - run / 2 - (int)rng_below(r, 2);
      if (lead < 0) lead += 64;
      for (int i = 0; i < lead && w < len; i++, w++) putc_(b, (char)('A' + rng_below(r, 26)));
      for (int i = 0; i < run; i++) putc_(b, '\\');
      w += run;
    } else {
      int n = 4 + (int)rng_below(r, 60);
      for (int i = 0; i < n; i++) putc_(b, (char)("abcdefghij:,{}[] xyz"[rng_below(r, 20)]));
      w += n;
    }
  }
  /* an even run right before the closing quote */
  int tail = 2 * (int)rng_below(r, 5);
  for (int i = 0; i < tail; i++) putc_(b, '\\');
  putc_(b, '"');
}

static void adversarial_line(sbuf *b, rng_t *r, size_t target) {
  size_t base = b->n;
  putc_(b, '{');
  key(b, "id"); put_u64(b, rng_u64(r) % 1000000000ull); putc_(b, ',');
  key(b, "s1"); adversarial_string(b, r, 200 + (int)rng_below(r, 400), base); putc_(b, ',');
  /* depth-20 member chain */
  key(b, "chain"); putc_(b, '{');
  for (int d = 1; d < 20; d++) {
    char k[8];
    k[0] = 'c';
    if (d < 10) { k[1] = (char)('0' + d); k[2] = 0; }
    else { k[1] = (char)('0' + d / 10); k[2] = (char)('0' + d % 10); k[3] = 0; }
    key(b, "pad"); put_u64(b, rng_below(r, 100)); putc_(b, ',');
    key(b, k);
    if (d < 19) putc_(b, '{');
  }
  tweet_string(b, r, 4);
  for (int d = 1; d < 20; d++) putc_(b, '}');
  putc_(b, ',');
  /* nested to depth 32 */
  key(b, "deep");
  for (int d = 0; d < 31; d++) { putc_(b, rng_below(r, 2) ? '{' : '{'); key(b, "d"); }
  puts_(b, "[1,\"x\\\\\",{\"q\":\"\\\"\"}]");
  for (int d = 0; d < 31; d++) putc_(b, '}');
  putc_(b, ',');
  /* array of 1000 small objects */
  key(b, "arr"); putc_(b, '[');
  for (int i = 0; i < 1000; i++) {
    if (i) putc_(b, ',');
    puts_(b, "{\"i\":");
    put_u64(b, (uint64_t)i);
    puts_(b, ",\"v\":\"");
    int k = (int)rng_below(r, 3);
    for (int j = 0; j < k; j++) putc_(b, (char)('a' + rng_below(r, 26)));
    puts_(b, "\"}");
  }
  puts_(b, "],");
  key(b, "s2");
  size_t have = b->n - base;
  int rest = (int)(target > have + 8 ? target - have - 4 : 16);
  adversarial_string(b, r, rest, base);
  puts_(b, "}\n");
}

typedef struct {
  int kind;
  uint64_t seed;
  uint64_t chunk;
  char *dst;
  size_t len;
} job_t;

static void fill_chunk(job_t *j) {
  rng_t r;
  rng_seed(&r, j->seed * 0x100000001B3ull + j->chunk * 0x9E3779B97F4A7C15ull + (uint64_t)j->kind);
  sbuf b = {0, 0, 0};
  size_t max_line = j->kind == 0 ? 40000 : 70000;
  for (;;) {
    size_t remain = j->len - b.n;
    if (remain < max_line + 64) break;
    size_t target;
    if (j->kind == 0) {
      double mu = log(2500.0) - 0.125;
      double x = exp(mu + 0.5 * rng_normal(&r));
      if (x < 1200) x = 1200;
      if (x > 30000) x = 30000;
      target = (size_t)x;
    } else {
      target = 4096 + rng_below(&r, 60 * 1024);
    }
    if (j->kind == 0) tweet_line(&b, &r, target);
    else adversarial_line(&b, &r, target);
  }
  size_t remain = j->len - b.n;
  if (remain >= 9) {
    /* pad line of exactly `remain` bytes */
    put(&b, "{\"p\":\"", 6);
    for (size_t i = 0; i < remain - 9; i++) putc_(&b, 'x');
    put(&b, "\"}\n", 3);
  } else if (remain > 0) {
    /* blanks before the previous line's newline */
    if (b.n > 0) {
      b.n--;
      for (size_t i = 0; i < remain; i++) putc_(&b, ' ');
      putc_(&b, '\n');
    } else {
      for (size_t i = 0; i < remain; i++) putc_(&b, ' ');
    }
  }
  memcpy(j->dst, b.p, j->len);
  free(b.p);
}

static void *worker(void *arg) {
  job_t *jobs = (job_t *)arg;
  for (; jobs->dst; jobs++) fill_chunk(jobs);
  return NULL;
}

/* Fills dst[0..size) with generator `kind` (0 tweets, 1 adversarial). */
int gjgen_fill(int kind, uint64_t seed, char *dst, size_t size, int nthreads) {
  const size_t CH = (size_t)8 << 20;
  size_t nch = (size + CH - 1) / CH;
  if (nch == 0) return 0;
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > nch) nthreads = (int)nch;
  job_t *all = (job_t *)calloc(nch + (size_t)nthreads, sizeof(job_t));
  if (!all) return -1;
  /* thread t gets chunks t, t+T, ... laid out contiguously with a NULL terminator */
  size_t k = 0;
  size_t *starts = (size_t *)calloc((size_t)nthreads, sizeof(size_t));
  for (int t = 0; t < nthreads; t++) {
    starts[t] = k;
    for (size_t c = (size_t)t; c < nch; c += (size_t)nthreads) {
      all[k].kind = kind;
      all[k].seed = seed;
      all[k].chunk = c;
      all[k].dst = dst + c * CH;
      all[k].len = (c + 1) * CH <= size ? CH : size - c * CH;
      k++;
    }
    all[k++].dst = NULL;
  }
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, &all[starts[t]]);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(starts);
  free(all);
  return 0;
}
