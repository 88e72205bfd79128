// server.cpp -- query serving with resident structural indexes.
//
// The reference re-reads and re-indexes the file on every query() call
// (J/functions/QueryFunction.java:149-197, J/ = gpjson/src/main/java/com/
// koenv/gpjson/); it has no serving mode.  This server keeps files and their
// indexes resident in HBM, keyed by file:
//   - registered files live in host memory (the server's copy); a file is
//     resident when its bytes and its index (built for `levels` levels, which
//     serves every query up to that depth bit-exactly) are on the device;
//   - residency is bounded by a stated HBM budget (input bytes + index bytes);
//     admitting a file evicts least-recently-used resident files that no
//     in-flight batch uses (their device copies are freed; the host copy stays);
//   - an admission thread stages and indexes files on its own stream while
//     queries run on the serving stream;
//   - the batching thread coalesces the pending requests whose file is
//     resident into ONE interpreter launch (a backlog larger than the free
//     slots can take is shared among them, kMinSplit requests or more per
//     launch; continuous batching: requests that
//     arrive while batches are on the GPU join the next batch; up to kSlots
//     batches are in flight, each with its own span buffers and its own
//     stream, so one batch's span copy to the host overlaps the next batch's
//     interpreter launch on the device);
//   - a finished request's spans stay in its batch's pinned buffer until the
//     client collects them (one copy, made by the waiting thread outside the
//     server lock); when every slot holds uncollected results and requests are
//     pending, the oldest slot's results are moved to per-request host memory.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "jsonpath.h"
#include "runtime.h"

using namespace gj;

namespace {

using Clock = std::chrono::steady_clock;

struct ServedFile {
  std::string key;
  std::vector<uint8_t> host;  // the server's host copy
  int64_t n = 0;
  // residency
  bool resident = false;
  bool admitting = false;
  gpjson_buf* buf = nullptr;
  gpjson_index* idx = nullptr;
  int64_t device_bytes = 0;
  int inflight = 0;  // batches using it
  Clock::time_point last_use{};
};

constexpr int kSlots = 4;
// A backlog is split among the free slots only into batches of at least this
// many requests.
constexpr size_t kMinSplit = 16;

struct Request {
  uint64_t ticket = 0;
  std::string key;
  std::string path;
  CompiledPath cp;
  Clock::time_point submitted{};
  Clock::time_point completed{};
  int state = 0;  // 0 pending, 1 in flight, 2 done, 3 failed
  std::string error;
  int64_t lines = 0;
  int R = 0;
  std::vector<int64_t> spans;  // results moved out of a batch buffer
  int slot = -1;               // batch slot whose pinned buffer holds the results
  int64_t off = 0;             // element offset in that buffer
  bool collecting = false;     // a waiter is copying from the slot, or a client holds a view of it
  int64_t file_bytes = 0;
};

struct Batch {
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<ServedFile*> files;
  DevBuf spans, jobs;
  void* host_spans = nullptr;  // pinned
  size_t host_cap = 0;
  void* host_jobs = nullptr;   // pinned staging of the job table + IR bytes
  size_t host_jobs_cap = 0;
  hipStream_t stream = nullptr;  // this slot's launches and copies
  hipEvent_t done = nullptr;
  bool active = false;           // launched, not yet complete
  int outstanding = 0;           // completed requests whose spans are still here
  uint64_t seq = 0;              // launch order (oldest is spilled first)
  std::vector<int64_t> offsets;  // element offset of each request's spans
};

}  // namespace

struct gpjson_server {
  gpjson_ctx* qctx = nullptr;  // serving stream (owned)
  gpjson_ctx* actx = nullptr;  // admission stream (owned)
  int64_t budget = 0;
  int levels = 0;
  std::mutex mu;
  std::condition_variable cv;       // wakes the batching thread
  std::condition_variable done_cv;  // wakes waiters
  std::condition_variable admit_cv;
  std::map<std::string, std::unique_ptr<ServedFile>> files;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  std::deque<std::shared_ptr<Request>> pending;
  std::deque<ServedFile*> admit_queue;
  uint64_t next_ticket = 1;
  int64_t resident_bytes = 0;  // inputs + indexes of resident (and admitting) files
  // the admission context's index scratch (grow-only, sized by the largest
  // file it built), counted against the budget too.  The batch slots' span
  // and job buffers are not: they hold results in flight, not resident data,
  // and counting them could leave a file that fits the budget unadmittable.
  int64_t admit_scratch = 0;
  bool stop = false;
  std::thread batcher, admitter;
  Batch slots[kSlots];
  uint64_t batch_seq = 0;
  // stats
  int64_t launches = 0, batched_requests = 0, admissions = 0, evictions = 0, max_batch = 0, spills = 0;
  std::string fatal;
};

namespace {

int64_t index_bytes(int64_t n, int levels) {
  const int64_t W = (n + 63) / 64, Wp = (W + 7) / 8 * 8;
  // records (string, colon, raw-quote and level-0 words), leveled words, newline and tile tables
  return kRecWords * W * 8 + leveled_alloc_words(Wp, levels) * 8 + (n / 256 + 4096) * 8 + (n / 16384 + 2) * 8 + 64;
}

// Frees LRU resident files (not used by in-flight batches, not being admitted)
// until `need` more bytes fit beside the resident files and the admission
// scratch.  Caller holds the lock.
bool make_room(gpjson_server* s, int64_t need, std::vector<ServedFile*>* to_free) {
  while (s->resident_bytes + s->admit_scratch + need > s->budget) {
    ServedFile* victim = nullptr;
    for (auto& kv : s->files) {
      ServedFile* f = kv.second.get();
      if (!f->resident || f->inflight > 0) continue;
      if (!victim || f->last_use < victim->last_use) victim = f;
    }
    if (!victim) return false;
    victim->resident = false;
    s->resident_bytes -= victim->device_bytes;
    s->evictions++;
    to_free->push_back(victim);
  }
  return true;
}

void release_device(ServedFile* f) {
  if (f->idx) gpjson_index_free(f->idx);
  if (f->buf) gpjson_buf_free(f->buf);
  f->idx = nullptr;
  f->buf = nullptr;
  f->device_bytes = 0;
}

void admitter_main(gpjson_server* s) {
  for (;;) {
    ServedFile* f = nullptr;
    std::vector<ServedFile*> to_free;
    bool ok = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->admit_cv.wait(lk, [&] { return s->stop || !s->admit_queue.empty(); });
      if (s->stop) return;
      f = s->admit_queue.front();
      // the admission context keeps scratch for the largest file it built
      const int64_t grow = std::max<int64_t>(0, index_scratch_bytes(f->n, s->actx->itune) - s->admit_scratch);
      const int64_t need = f->n + 64 + index_bytes(f->n, s->levels);
      ok = make_room(s, need + grow, &to_free);
      if (ok) {
        s->admit_queue.pop_front();
        s->resident_bytes += need;
        s->admit_scratch += grow;
        f->device_bytes = need;
      }
    }
    // device memory of evicted files is released outside the lock
    for (ServedFile* v : to_free) release_device(v);
    if (!ok) {
      // everything resident is used by in-flight batches: retry shortly
      std::this_thread::sleep_for(std::chrono::microseconds(200));
      continue;
    }
    int rc = gpjson_buf_from_host(s->actx, f->host.data(), f->n, &f->buf);
    if (rc == GPJSON_OK) rc = gpjson_index_build(s->actx, f->buf, s->levels, 0, &f->idx);
    std::unique_lock<std::mutex> lk(s->mu);
    f->admitting = false;
    if (rc != GPJSON_OK) {
      s->fatal = std::string("admission of ") + f->key + " failed: " + gpjson_last_error();
      s->resident_bytes -= f->device_bytes;
      release_device(f);
    } else {
      // charge what the index really holds (its newline array grows for files
      // of lines shorter than the first estimate)
      const int64_t actual = f->n + 64 + index_device_bytes(f->idx);
      s->resident_bytes += actual - f->device_bytes;
      f->device_bytes = actual;
      f->resident = true;
      f->last_use = Clock::now();
      s->admissions++;
      // back under the budget if the estimate was short (the new file is the
      // most recently used, so others go first)
      std::vector<ServedFile*> over;
      f->inflight++;  // never the new file itself
      make_room(s, 0, &over);
      f->inflight--;
      if (!over.empty()) {
        lk.unlock();
        for (ServedFile* v : over) release_device(v);
        lk.lock();
      }
    }
    s->cv.notify_all();
  }
}

// Completes a finished batch: its requests' spans are now in the slot's pinned
// buffer, where waiters read them.  Lock held.
void finish_batch(gpjson_server* s, int k) {
  Batch& b = s->slots[k];
  const auto now = Clock::now();
  for (size_t i = 0; i < b.reqs.size(); i++) {
    Request& r = *b.reqs[i];
    r.slot = k;
    r.off = b.offsets[i];
    r.completed = now;
    r.state = 2;
  }
  for (ServedFile* f : b.files) {
    f->inflight--;
    f->last_use = now;
  }
  b.outstanding = (int)b.reqs.size();
  b.files.clear();
  b.active = false;
  if (b.outstanding == 0) b.reqs.clear();
  s->done_cv.notify_all();
  s->admit_cv.notify_all();
}

// Moves the uncollected results of the oldest completed slot into per-request
// host memory so the slot can take a new batch.  Requests that a waiter is
// copying right now keep the slot busy until that copy ends.  Lock held.
void spill_oldest(gpjson_server* s) {
  int k = -1;
  for (int j = 0; j < kSlots; j++) {
    const Batch& b = s->slots[j];
    if (!b.active && b.outstanding > 0 && (k < 0 || b.seq < s->slots[k].seq)) k = j;
  }
  if (k < 0) return;
  Batch& b = s->slots[k];
  for (auto& rp : b.reqs) {
    Request& r = *rp;
    if (r.slot != k || r.collecting) continue;
    const int64_t cnt = r.lines * r.R * 2;
    r.spans.assign((int64_t*)b.host_spans + r.off, (int64_t*)b.host_spans + r.off + cnt);
    r.slot = -1;
    b.outstanding--;
  }
  if (b.outstanding == 0) b.reqs.clear();
  s->spills++;
}

// Launches one interpreter over every request in `reqs` (files resident).  Lock
// NOT held (only the batching thread touches slots).
int launch_batch(gpjson_server* s, Batch& b) {
  gpjson_ctx* ctx = s->qctx;  // launch shapes and CU count
  hipStream_t stream = b.stream;
  const int n = (int)b.reqs.size();
  std::vector<QueryJob> jobs(n);
  int64_t items = 0, elems = 0, ir_total = 0;
  int maxL = 0;
  b.offsets.assign(n, 0);
  for (int i = 0; i < n; i++) ir_total += (int64_t)b.reqs[i]->cp.ir.size();
  // device layout: [work counters][jobs][IR bytes]
  const size_t jobs_off = kQueryCounterWords * 8;
  const size_t jobs_bytes = sizeof(QueryJob) * n;
  const size_t table_bytes = jobs_off + jobs_bytes + (size_t)ir_total + 16;
  int rc = b.jobs.ensure(table_bytes);
  if (rc) return rc;
  if (b.host_jobs_cap < table_bytes) {
    if (b.host_jobs) (void)hipHostFree(b.host_jobs);
    b.host_jobs = nullptr;
    b.host_jobs_cap = std::max<size_t>(table_bytes * 2, 1 << 16);
    if (hipHostMalloc(&b.host_jobs, b.host_jobs_cap, hipHostMallocDefault) != hipSuccess) {
      b.host_jobs_cap = 0;
      return fail(GPJSON_ERR_NO_MEMORY, "pinned job table");
    }
  }
  uint8_t* host = (uint8_t*)b.host_jobs;
  std::memset(host, 0, table_bytes);
  int64_t irb = (int64_t)(jobs_off + jobs_bytes);
  for (int i = 0; i < n; i++) {
    Request& r = *b.reqs[i];
    const gpjson_index* idx = b.files[i]->idx;
    r.lines = idx->lines;
    r.R = std::max(r.cp.num_results, 1);
    b.offsets[i] = elems;
    QueryJob& j = jobs[i];
    j.F = idx->buf->d;
    j.N = idx->n;
    j.NL = idx->NL;
    j.lines = idx->lines;
    j.rec = idx->rec;
    j.LB = idx->LB;
    j.LS = idx->LS;
    j.Ws = idx->Ws;
    j.Wp = idx->Wp;
    j.L = idx->L;
    j.R = r.R;
    j.ir = (const uint8_t*)b.jobs.p + irb;
    j.ir_len = (int64_t)r.cp.ir.size();
    j.item0 = items;
    std::memcpy(host + irb, r.cp.ir.data(), r.cp.ir.size());
    irb += (int64_t)r.cp.ir.size();
    items += idx->lines;
    elems += idx->lines * r.R * 2;
    maxL = std::max(maxL, idx->L);
  }
  rc = b.spans.ensure((size_t)std::max<int64_t>(elems, 1) * 8);
  if (rc) return rc;
  for (int i = 0; i < n; i++) jobs[i].out = (int64_t*)b.spans.p + b.offsets[i];
  std::memcpy(host + jobs_off, jobs.data(), jobs_bytes);
  if (b.host_cap < (size_t)elems * 8) {
    if (b.host_spans) (void)hipHostFree(b.host_spans);
    b.host_spans = nullptr;
    b.host_cap = std::max<size_t>((size_t)elems * 8 * 2, 1 << 20);
    if (hipHostMalloc(&b.host_spans, b.host_cap, hipHostMallocDefault) != hipSuccess) {
      b.host_cap = 0;
      return fail(GPJSON_ERR_NO_MEMORY, "pinned span buffer");
    }
  }
  // the pinned staging stays untouched until this slot's batch completes
  GJ_HIP(hipMemcpyAsync(b.jobs.p, host, table_bytes, hipMemcpyHostToDevice, stream));
  QueryLaunch L{};
  L.jobs = (const QueryJob*)((uint8_t*)b.jobs.p + jobs_off);
  L.njobs = n;
  L.total_items = items;
  L.counter = (unsigned long long*)b.jobs.p;
  L.lines_per_wave = ctx->qtune.lines_per_wave;
  L.ir_base = (const uint8_t*)b.jobs.p + jobs_off + jobs_bytes;
  L.ir_bytes = ir_total;
  GJ_HIP(zero_u64(L.counter, kQueryCounterWords, stream));
  const int waves_per_block = ctx->qtune.block / 64;
  const int grid = std::max(1, ctx->num_cus * ctx->qtune.waves_per_cu / waves_per_block);
  GJ_HIP(query_launch(L, ctx->qtune.block, ctx->qtune.group, maxL, grid, stream));
  GJ_HIP(hipMemcpyAsync(b.host_spans, b.spans.p, (size_t)elems * 8, hipMemcpyDeviceToHost, stream));
  GJ_HIP(hipEventRecord(b.done, stream));
  return GPJSON_OK;
}

void batcher_main(gpjson_server* s) {
  (void)hipSetDevice(s->qctx->device);
  for (;;) {
    int free_slot = -1;
    std::vector<std::shared_ptr<Request>> take;
    std::vector<ServedFile*> take_files;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) return;
        // complete finished batches
        for (int k = 0; k < kSlots; k++)
          if (s->slots[k].active && hipEventQuery(s->slots[k].done) == hipSuccess) finish_batch(s, k);
        free_slot = -1;
        for (int k = 0; k < kSlots; k++)
          if (!s->slots[k].active && s->slots[k].outstanding == 0) { free_slot = k; break; }
        if (free_slot < 0 && !s->pending.empty()) {
          bool any_active = false;
          for (const Batch& b : s->slots) any_active |= b.active;
          // slow collectors must not stall the GPU: free the oldest slot
          if (!any_active) spill_oldest(s);
          for (int k = 0; k < kSlots; k++)
            if (!s->slots[k].active && s->slots[k].outstanding == 0) { free_slot = k; break; }
        }
        // requests whose file is resident can run now; others wait for admission.
        // A large backlog is shared among the free slots, so that one batch's
        // span copy to the host overlaps the next batch's launch instead of
        // one batch holding everything
        if (free_slot >= 0) {
          size_t free_slots = 0;
          for (int k = 0; k < kSlots; k++) free_slots += !s->slots[k].active && s->slots[k].outstanding == 0;
          const size_t cap = std::max<size_t>(kMinSplit, (s->pending.size() + free_slots - 1) / free_slots);
          for (auto it = s->pending.begin(); it != s->pending.end();) {
            ServedFile* f = s->files[(*it)->key].get();
            if (f->resident && take.size() >= cap) break;
            if (f->resident) {
              take.push_back(*it);
              take_files.push_back(f);
              f->inflight++;
              f->last_use = Clock::now();
              it = s->pending.erase(it);
            } else {
              if (!f->admitting) {
                f->admitting = true;
                s->admit_queue.push_back(f);
                s->admit_cv.notify_all();
              }
              ++it;
            }
          }
        }
        if (!take.empty()) break;
        bool any_active = false;
        for (const Batch& b : s->slots) any_active |= b.active;
        if (any_active) {
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(20));
          lk.lock();
        } else {
          s->cv.wait_for(lk, std::chrono::milliseconds(2));
        }
      }
      Batch& b = s->slots[free_slot];
      b.reqs = take;
      b.files = take_files;
      b.active = true;
      b.seq = ++s->batch_seq;
      for (auto& r : take) r->state = 1;
      s->launches++;
      s->batched_requests += (int64_t)take.size();
      s->max_batch = std::max<int64_t>(s->max_batch, (int64_t)take.size());
    }
    int rc = launch_batch(s, s->slots[free_slot]);
    if (rc != GPJSON_OK) {
      std::lock_guard<std::mutex> lk(s->mu);
      Batch& b = s->slots[free_slot];
      for (auto& r : b.reqs) {
        r->state = 3;
        r->error = gpjson_last_error();
      }
      for (ServedFile* f : b.files) f->inflight--;
      b.reqs.clear();
      b.files.clear();
      b.active = false;
      s->done_cv.notify_all();
    }
  }
}

}  // namespace

extern "C" {

int gpjson_server_create(int device, int64_t hbm_budget_bytes, int levels, gpjson_server** out) {
  if (!out || hbm_budget_bytes <= 0 || levels < 1 || levels > GPJSON_MAX_LEVELS)
    return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  std::unique_ptr<gpjson_server> s(new gpjson_server());
  int rc = gpjson_ctx_create(device, &s->qctx);
  if (rc) return rc;
  rc = gpjson_ctx_create(device, &s->actx);
  if (rc) {
    gpjson_ctx_destroy(s->qctx);
    return rc;
  }
  s->budget = hbm_budget_bytes;
  s->levels = levels;
  GJ_HIP(hipSetDevice(device));
  for (Batch& b : s->slots) {
    GJ_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    GJ_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));  // default priority
  }
  gpjson_server* raw = s.get();
  s->batcher = std::thread(batcher_main, raw);
  s->admitter = std::thread(admitter_main, raw);
  *out = s.release();
  return GPJSON_OK;
}

int gpjson_server_add(gpjson_server* s, const char* key, const void* data, int64_t n) {
  if (!s || !key || n < 0 || (n > 0 && !data)) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  std::unique_ptr<ServedFile> f(new ServedFile());
  f->key = key;
  f->n = n;
  f->host.assign((const uint8_t*)data, (const uint8_t*)data + n);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->files.count(key)) return fail(GPJSON_ERR_INVALID_ARG, std::string("file already registered: ") + key);
  if (f->n + 64 + index_bytes(f->n, s->levels) + index_scratch_bytes(f->n, s->actx->itune) > s->budget)
    return fail(GPJSON_ERR_NO_MEMORY, std::string("file larger than the HBM budget: ") + key);
  s->files[key] = std::move(f);
  return GPJSON_OK;
}

int gpjson_server_admit(gpjson_server* s, const char* key) {
  if (!s || !key) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->files.find(key);
  if (it == s->files.end()) return fail(GPJSON_ERR_INVALID_ARG, std::string("unknown file: ") + key);
  ServedFile* f = it->second.get();
  if (!f->resident && !f->admitting) {
    f->admitting = true;
    s->admit_queue.push_back(f);
    s->admit_cv.notify_all();
  }
  s->cv.wait(lk, [&] { return f->resident || !s->fatal.empty() || (!f->admitting && !f->resident); });
  if (!f->resident) return fail(GPJSON_ERR_NO_MEMORY, s->fatal.empty() ? "admission failed" : s->fatal);
  return GPJSON_OK;
}

int gpjson_server_submit(gpjson_server* s, const char* key, const char* path, uint64_t* ticket) {
  if (!s || !key || !path || !ticket) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  auto r = std::make_shared<Request>();
  std::string err;
  PathStatus st = compile_path(path, &r->cp, &err);
  if (st == PathStatus::kUnsupported) return fail(GPJSON_ERR_UNSUPPORTED, err);
  if (st == PathStatus::kParseError) return fail(GPJSON_ERR_PARSE, err);
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->files.find(key);
  if (it == s->files.end()) return fail(GPJSON_ERR_INVALID_ARG, std::string("unknown file: ") + key);
  if (r->cp.max_depth > s->levels)
    return fail(GPJSON_ERR_CAPACITY, "path deeper than the server's index levels");
  r->ticket = s->next_ticket++;
  r->key = key;
  r->path = path;
  r->file_bytes = it->second->n;
  r->submitted = Clock::now();
  s->requests[r->ticket] = r;
  s->pending.push_back(r);
  *ticket = r->ticket;
  s->cv.notify_all();
  return GPJSON_OK;
}

// Waits for a request (timeout_ms < 0: forever).  On success fills lines, R,
// latency (submit to completion, microseconds) and, when spans is not NULL,
// copies int64[lines][R][2] spans (cap_elems capacity).  The request is
// forgotten after a successful call with spans != NULL.
int gpjson_server_wait(gpjson_server* s, uint64_t ticket, int64_t timeout_ms, int64_t* lines, int* R,
                       double* latency_us, int64_t* spans, int64_t cap_elems) {
  if (!s) return fail(GPJSON_ERR_INVALID_ARG, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return fail(GPJSON_ERR_INVALID_ARG, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->state >= 2 || !s->fatal.empty(); };
  if (timeout_ms < 0) s->done_cv.wait(lk, ready);
  else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return fail(GPJSON_ERR_CAPACITY, "timeout");
  if (r->state == 3) return fail(GPJSON_ERR_HIP, r->error);
  if (r->state != 2) return fail(GPJSON_ERR_NO_MEMORY, s->fatal);
  if (lines) *lines = r->lines;
  if (R) *R = r->R;
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->completed - r->submitted).count();
  if (spans) {
    const int64_t cnt = r->lines * r->R * 2;
    if (cap_elems < cnt) return fail(GPJSON_ERR_CAPACITY, "span buffer too small");
    if (r->collecting) return fail(GPJSON_ERR_INVALID_ARG, "ticket is being collected by another thread");
    s->requests.erase(ticket);
    if (r->slot < 0) {
      std::memcpy(spans, r->spans.data(), (size_t)cnt * 8);
      return GPJSON_OK;
    }
    // the slot is not reused while this request is outstanding: copy unlocked
    Batch& b = s->slots[r->slot];
    r->collecting = true;
    lk.unlock();
    std::memcpy(spans, (const int64_t*)b.host_spans + r->off, (size_t)cnt * 8);
    lk.lock();
    r->collecting = false;
    r->slot = -1;
    if (--b.outstanding == 0) {
      b.reqs.clear();
      s->cv.notify_all();
    }
  }
  return GPJSON_OK;
}

// Zero-copy access: waits like gpjson_server_wait, then gives a pointer to the
// request's spans where the server holds them (its batch's pinned buffer, or
// host memory after a spill).  The pointer stays valid, and that batch buffer
// stays reserved, until gpjson_server_release.
int gpjson_server_result(gpjson_server* s, uint64_t ticket, int64_t timeout_ms, const int64_t** spans,
                         int64_t* lines, int* R, double* latency_us) {
  if (!s || !spans) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return fail(GPJSON_ERR_INVALID_ARG, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->state >= 2 || !s->fatal.empty(); };
  if (timeout_ms < 0) s->done_cv.wait(lk, ready);
  else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return fail(GPJSON_ERR_CAPACITY, "timeout");
  if (r->state == 3) return fail(GPJSON_ERR_HIP, r->error);
  if (r->state != 2) return fail(GPJSON_ERR_NO_MEMORY, s->fatal);
  if (r->collecting) return fail(GPJSON_ERR_INVALID_ARG, "ticket is already held or being collected");
  r->collecting = true;  // keeps the slot reserved (spills skip it)
  *spans = r->slot >= 0 ? (const int64_t*)s->slots[r->slot].host_spans + r->off : r->spans.data();
  if (lines) *lines = r->lines;
  if (R) *R = r->R;
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->completed - r->submitted).count();
  return GPJSON_OK;
}

// Ends a zero-copy access: the ticket is forgotten and its batch buffer freed
// for reuse once every request of that batch is collected or released.
int gpjson_server_release(gpjson_server* s, uint64_t ticket) {
  if (!s) return fail(GPJSON_ERR_INVALID_ARG, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return fail(GPJSON_ERR_INVALID_ARG, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  if (!r->collecting || r->state != 2) return fail(GPJSON_ERR_INVALID_ARG, "ticket is not held");
  s->requests.erase(it);
  r->collecting = false;
  if (r->slot >= 0) {
    Batch& b = s->slots[r->slot];
    r->slot = -1;
    if (--b.outstanding == 0) {
      b.reqs.clear();
      s->cv.notify_all();
    }
  }
  return GPJSON_OK;
}

// Counters, in this order: resident bytes, resident files, interpreter
// launches, requests served in them, largest batch, admissions, evictions,
// spills (slots whose uncollected results were moved to host memory).  Fills
// min(n, 8) values.
int gpjson_server_stats(gpjson_server* s, int64_t* out, int n) {
  if (!s || !out || n < 0) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  int64_t res_files = 0;
  for (auto& kv : s->files) res_files += kv.second->resident ? 1 : 0;
  const int64_t v[8] = {s->resident_bytes, res_files, s->launches, s->batched_requests, s->max_batch,
                        s->admissions, s->evictions, s->spills};
  for (int i = 0; i < n && i < 8; i++) out[i] = v[i];
  return GPJSON_OK;
}

void gpjson_server_destroy(gpjson_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
    s->admit_cv.notify_all();
    s->done_cv.notify_all();
  }
  if (s->batcher.joinable()) s->batcher.join();
  if (s->admitter.joinable()) s->admitter.join();
  (void)hipSetDevice(s->qctx->device);
  (void)hipStreamSynchronize(s->qctx->stream);
  (void)hipStreamSynchronize(s->actx->stream);
  for (Batch& b : s->slots)
    if (b.stream) (void)hipStreamSynchronize(b.stream);
  for (Batch& b : s->slots) {
    b.spans.release();
    b.jobs.release();
    if (b.host_spans) (void)hipHostFree(b.host_spans);
    if (b.host_jobs) (void)hipHostFree(b.host_jobs);
    if (b.done) (void)hipEventDestroy(b.done);
    if (b.stream) (void)hipStreamDestroy(b.stream);
  }
  for (auto& kv : s->files) release_device(kv.second.get());
  gpjson_ctx_destroy(s->actx);
  gpjson_ctx_destroy(s->qctx);
  delete s;
}

}  // extern "C"
