// CPU model of the index kernel's carry-in protocol across tiles
// (gpjson_amd/csrc/index_kernel.hip: index_kernel's ticket claim and carry-in,
// lookback_wave, publish_agg / publish_group / publish_inc), run with threads
// as the persistent grid's blocks.  It restates the walk lane by lane (a
// wavefront's 64 lanes as a loop, ballots as masks) over a non-commutative
// monoid (affine maps mod 2^61 - 1), with random delays at every step, and
// checks that every launch ends and every tile's carry-in equals the
// sequential prefix.  A published pair of state words is one event here (the
// kernel's reader takes a pair only once both words are valid).  Launch
// patterns:
//   - one launch over all tiles (the plain and batched index builds);
//   - several launches over consecutive tile ranges [tile_begin, end) with the
//     tickets reset before each, the tile states kept (the staging pipeline);
// tiles are claimed as the kernel claims them: npart residue classes (the
// largest power of two up to the kernel's kTicketParts -- set here by
// lb_set_parts, 1 by default as in the kernel -- and the grid), block b
// claiming the tiles of class b % npart in order from that class's counter;
// and tile kinds:
//   - tiles that look back before emitting;
//   - speculating tiles that publish their inclusive prefix at once (deferred
//     bookkeeping) or only after their own lookback (in-kernel bookkeeping).
#include <atomic>
#include <chrono>
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kP = (1ull << 61) - 1;

uint64_t mulmod(uint64_t a, uint64_t b) {
  const unsigned __int128 r = (unsigned __int128)a * b;
  uint64_t lo = (uint64_t)(r & kP), hi = (uint64_t)(r >> 61);
  uint64_t s = lo + hi;
  if (s >= kP) s -= kP;
  return s;
}
uint64_t addmod(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  if (s >= kP) s -= kP;
  return s;
}

// Transfer function of a tile or a run of tiles: x -> a x + b.
struct Agg {
  uint64_t a = 1, b = 0;
};
Agg agg_id() { return Agg{}; }
// x then y
Agg agg_cat(const Agg& x, const Agg& y) { return Agg{mulmod(y.a, x.a), addmod(mulmod(y.a, x.b), y.b)}; }
uint64_t apply(const Agg& f, uint64_t c) { return addmod(mulmod(f.a, c), f.b); }

struct State {
  std::atomic<bool> agg_ok{false}, inc_ok{false}, grp_ok{false};
  Agg agg, grp;
  uint64_t inc = 0;
};

struct Model {
  int64_t ntiles;
  std::vector<Agg> tile_agg;   // the data: each tile's transfer function
  std::vector<uint8_t> spec;   // the tile speculates its carry-in
  std::vector<State> st;
  std::vector<uint64_t> carry_in;
  std::vector<uint64_t> prefix;  // the sequential carry-in of each tile
  bool deferred;               // speculating tiles publish their inclusive prefix at once
  uint64_t start = 12345;      // carry before tile 0
  std::atomic<bool> abort{false};
  std::atomic<unsigned> ticket[8];    // a counter per residue class of tiles
  int delay_pct;               // chance of a delay at each step

  explicit Model(int64_t n) : ntiles(n), tile_agg(n), spec(n), st(n), carry_in(n), prefix(n) {}
};

void maybe_delay(Model& m, std::mt19937_64& rng) {
  if ((int)(rng() % 100) < m.delay_pct) {
    if (rng() & 1) std::this_thread::yield();
    else std::this_thread::sleep_for(std::chrono::microseconds(rng() % 40));
  }
}

// The composition of lanes [0, 64) in tile order (lane l holds the unit l
// steps back from the cursor): x_63 then ... then x_0.
Agg reduce_units(const Agg (&x)[64]) {
  Agg r = agg_id();
  for (int l = 63; l >= 0; l--) r = agg_cat(r, x[l]);
  return r;
}

// lookback_wave of the kernel, lane by lane.
bool lookback(Model& m, int64_t lt, const Agg& own, std::mt19937_64& rng, uint64_t* out) {
  Agg acc = agg_id();
  int64_t c = lt - 1;
  bool group_todo = (lt & 63) == 63;
  for (;;) {
    if (m.abort.load()) return false;
    if (c < 0) {
      *out = apply(acc, m.start);
      return true;
    }
    const int64_t g = c >> 6;
    const int p = (int)(c - (g << 6)) + 1;
    int status[64];
    bool agg_ok[64];
    Agg a[64];
    uint64_t cc[64];
    for (int lane = 0; lane < 64; lane++) {
      status[lane] = 0;
      agg_ok[lane] = false;
      a[lane] = agg_id();
      cc[lane] = 0;
      State* s = nullptr;
      bool is_group = false;
      if (lane < p) {
        s = &m.st[c - lane];
      } else if (g - 1 - (lane - p) >= 0) {
        s = &m.st[(g - 1 - (lane - p)) * 64 + 63];
        is_group = true;
      } else {
        status[lane] = 2;
        cc[lane] = m.start;
      }
      if (s) {
        const bool inc_ok = s->inc_ok.load(std::memory_order_acquire);
        const bool aok = is_group ? s->grp_ok.load(std::memory_order_acquire) : s->agg_ok.load(std::memory_order_acquire);
        agg_ok[lane] = aok;
        if (aok) a[lane] = is_group ? s->grp : s->agg;
        if (inc_ok) {
          status[lane] = 2;
          cc[lane] = s->inc;
        } else if (aok) {
          status[lane] = 1;
        }
      }
    }
    maybe_delay(m, rng);
    if (group_todo) {
      bool all = true;
      for (int lane = 0; lane < 63; lane++) all = all && agg_ok[lane];
      if (!all) {
        std::this_thread::yield();
        continue;
      }
      Agg x[64];
      for (int lane = 0; lane < 64; lane++) x[lane] = lane < 63 ? a[lane] : agg_id();
      const Agg grp = agg_cat(reduce_units(x), own);
      m.st[lt].grp = grp;
      m.st[lt].grp_ok.store(true, std::memory_order_release);
      group_todo = false;
    }
    int a_inc = 64, a_nr = 64;
    for (int lane = 63; lane >= 0; lane--) {
      if (status[lane] == 2) a_inc = lane;
      if (status[lane] == 0) a_nr = lane;
    }
    if (a_inc < a_nr) {
      Agg x[64];
      for (int lane = 0; lane < 64; lane++) x[lane] = lane >= a_inc ? agg_id() : a[lane];
      *out = apply(agg_cat(reduce_units(x), acc), cc[a_inc]);
      return true;
    }
    if (a_nr == 0) {
      std::this_thread::yield();
      continue;
    }
    Agg x[64];
    for (int lane = 0; lane < 64; lane++) x[lane] = lane >= a_nr ? agg_id() : a[lane];
    acc = agg_cat(reduce_units(x), acc);
    if (a_nr < p) c -= a_nr;
    else if (a_nr < 64) c = (g - 1 - (a_nr - p)) * 64 + 63;
    else c = (g - (64 - p)) * 64 - 1;
  }
}

void publish_inc(Model& m, int64_t lt, uint64_t v) {
  m.st[lt].inc = v;
  m.st[lt].inc_ok.store(true, std::memory_order_release);
}

// One block of the persistent grid: claims tiles of [begin, end) until none
// are left (the kernel's loop body, minus the data work).  As in the kernel,
// the ticket for the block's next tile is claimed when its current tile
// starts, so a block always holds one claimed tile it has not started, and
// block b of a grid of `grid` claims the tiles of class b % npart.
unsigned g_max_parts = 1;
unsigned g_claim = 1;  // consecutive tiles of the class per atomic

void block_main(Model& m, int64_t begin, int64_t end, int b, int grid, uint64_t seed) {
  std::mt19937_64 rng(seed);
  unsigned npart = 1;
  while (npart * 2 <= g_max_parts && npart * 2 <= (unsigned)grid) npart *= 2;
  const unsigned part = (unsigned)b & (npart - 1);
  unsigned in_claim = 0;
  auto claim = [&] { return begin + (int64_t)m.ticket[part].fetch_add(g_claim) * npart + part; };
  int64_t next = claim();
  for (;;) {
    if (m.abort.load()) return;
    const int64_t lt = next;
    if (lt >= end) return;
    if (++in_claim < g_claim) {
      next += npart;
    } else {
      in_claim = 0;
      next = claim();
    }
    maybe_delay(m, rng);
    const Agg own = m.tile_agg[lt];
    // every tile publishes its aggregate first
    m.st[lt].agg = own;
    m.st[lt].agg_ok.store(true, std::memory_order_release);
    maybe_delay(m, rng);
    const bool spec = m.spec[lt] && lt > 0;
    uint64_t cin = 0;
    if (!spec) {
      if (lt == 0) cin = m.start;
      else if (!lookback(m, lt, own, rng, &cin)) return;
      publish_inc(m, lt, apply(own, cin));
      m.carry_in[lt] = cin;
      maybe_delay(m, rng);  // emission
    } else {
      // the speculated carry-in: the true one (input whose speculation holds)
      const uint64_t guess = m.prefix[lt];
      if (m.deferred) {
        publish_inc(m, lt, apply(own, guess));
        m.carry_in[lt] = guess;
        maybe_delay(m, rng);  // emission; the bookkeeping checks the guess later
      } else {
        maybe_delay(m, rng);  // emission with the guess, then the exact carry-in
        if (!lookback(m, lt, own, rng, &cin)) return;
        publish_inc(m, lt, apply(own, cin));
        m.carry_in[lt] = cin;
      }
    }
  }
}

}  // namespace

extern "C" {
// Largest number of ticket classes (a power of two, at most 8).
void lb_set_parts(int parts) { g_max_parts = (unsigned)parts; }
// Consecutive tiles of a class per claim (the kernel: 2 in its 2-level variant).
void lb_set_claim(int claim) { g_claim = (unsigned)claim; }

// Runs the launches [bounds[k], bounds[k+1]) one after another (the tickets
// reset before each, states kept), each on `grid` threads.  spec: tiles that
// speculate (1) or look back (0); deferred: speculating tiles publish their
// prefix at once.  Returns 0 when every launch ended and every carry-in is
// the sequential prefix, 1 on a wrong carry-in, 2 when a launch did not end
// within timeout_ms (the walk would hang).
int lb_run(int64_t ntiles, const int64_t* bounds, int nlaunch, int grid, const uint8_t* spec, int deferred,
           int delay_pct, uint64_t seed, int timeout_ms) {
  Model m(ntiles);
  std::mt19937_64 rng(seed);
  for (int64_t t = 0; t < ntiles; t++) {
    m.tile_agg[t] = Agg{1 + rng() % (kP - 1), rng() % kP};
    m.spec[t] = spec[t];
  }
  uint64_t pre = m.start;
  for (int64_t t = 0; t < ntiles; t++) {
    m.prefix[t] = pre;
    pre = apply(m.tile_agg[t], pre);
  }
  m.deferred = deferred != 0;
  m.delay_pct = delay_pct;
  for (int k = 0; k < nlaunch; k++) {
    for (auto& t : m.ticket) t.store(0);
    std::vector<std::thread> th;
    for (int b = 0; b < grid; b++)
      th.emplace_back(block_main, std::ref(m), bounds[k], bounds[k + 1], b, grid, seed * 7919 + b + k * 104729);
    // watchdog: a launch that does not end is reported, and its waiters released
    std::atomic<bool> done{false};
    std::thread dog([&] {
      const auto t0 = std::chrono::steady_clock::now();
      while (!done.load()) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(timeout_ms)) {
          m.abort.store(true);
          return;
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
      }
    });
    for (auto& t : th) t.join();
    done.store(true);
    dog.join();
    if (m.abort.load()) return 2;
  }
  for (int64_t t = 0; t < ntiles; t++)
    if (m.carry_in[t] != m.prefix[t]) return 1;
  return 0;
}
}
