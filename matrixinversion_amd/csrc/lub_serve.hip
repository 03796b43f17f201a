// lub_serve.hip -- continuous-batching server over the ragged path
// (include/lubatch.h, lu_server_*).
//
// The reference has no serving layer: its run.py compiles, launches and times
// one batch per process (parallel_pivot/run.py:198-234).  Here requests of
// any size mix arrive one at a time and are inverted in HBM-resident batches:
//
//   submit    a request (count matrices, packed row-major, sizes ns[]) is
//             admitted if its block fits the HBM budget: the block (its
//             matrices, then its pivot vectors) is carved from a ring arena in
//             submission order and filled by an async copy on the copy stream;
//             the request's slot entries (matrix offset, n, pivot offset) are
//             written to the pinned staging ring of its pivot rule.
//             Otherwise LUB_EBUSY: the caller keeps the request and retries
//             after progress/release.
//   progress  retires batches whose completion event has fired, then, for
//             each pivot rule whose stream is idle, launches every queued
//             request of that rule (contiguous in its slot ring) as one ragged
//             batch: the batch's slot entries go up in one copy per array, the
//             compute stream waits for the copies, runs the ragged plan and
//             records the batch's completion event.
//   query     state, device result pointers and end-to-end latency (host clock,
//             submit to the progress call that saw the batch complete).
//   release   the caller is done with the results; arena space is reclaimed
//             in submission order.
//
// One host thread drives a server; nothing here blocks except destroy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <deque>
#include <new>
#include <vector>

#include "lubatch.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kRules = 5;                 // LUB_PIVOT_NONE .. LUB_PIVOT_GEPP
constexpr int64_t kAlign = 256;           // request blocks start on 256-byte boundaries

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

enum State : int { kQueued = 0, kRunning = 1, kDone = 2 };

struct Request {
  uint64_t ticket = 0;
  int rule = 0;
  int flags = 0;
  int count = 0;
  int64_t block = 0;       // arena byte offset of the block
  int64_t block_end = 0;   // monotonic arena position after the block
  int64_t floats = 0;      // matrix floats (sum n^2)
  int64_t slot = 0;        // monotonic position of the first slot in its rule's ring
  int state = kQueued;
  bool released = false;
  Clock::time_point t_submit, t_done;
};

struct Batch {
  hipEvent_t done = nullptr;
  int64_t slot_end = 0;           // monotonic ring position after the batch
  std::vector<uint64_t> tickets;
};

struct Rule {
  hipStream_t stream = nullptr;
  lub_ragged_plan* plan = nullptr;
  int64_t* d_offs = nullptr;      // device slot ring
  int* d_ns = nullptr;
  int64_t* d_poffs = nullptr;
  int64_t* h_offs = nullptr;      // pinned staging ring
  int* h_ns = nullptr;
  int64_t* h_poffs = nullptr;
  int64_t head = 0, tail = 0;     // monotonic slot positions
  std::deque<uint64_t> queued;    // tickets in submission order
  std::deque<Batch> inflight;
};

}  // namespace

struct lub_server {
  int device = 0;
  char* arena = nullptr;
  int64_t arena_bytes = 0;
  int64_t head = 0, tail = 0;     // monotonic arena positions
  int64_t max_batch = 0;
  int64_t ring = 0;               // slots per rule ring
  int max_inflight = 1;
  hipStream_t copy = nullptr;
  Rule rules[kRules];
  std::deque<Request> reqs;       // reqs[i].ticket == first_ticket + i
  uint64_t first_ticket = 1;
  uint64_t next_ticket = 1;
  std::vector<hipEvent_t> event_pool;
};

namespace {

Request* find(lub_server* s, uint64_t t) {
  if (t < s->first_ticket || t >= s->next_ticket) return nullptr;
  return &s->reqs[t - s->first_ticket];
}

hipEvent_t take_event(lub_server* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

int rule_init(lub_server* s, Rule& r) {
  if (r.stream) return LUB_OK;
  if (hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking) != hipSuccess) return LUB_EHIP;
  int rc = lu_ragged_plan_create(&r.plan, s->max_batch);
  if (rc != LUB_OK) return rc;
  const size_t n = static_cast<size_t>(s->ring);
  if (hipMalloc(&r.d_offs, n * 8) != hipSuccess || hipMalloc(&r.d_ns, n * 4) != hipSuccess ||
      hipMalloc(&r.d_poffs, n * 8) != hipSuccess)
    return LUB_EHIP;
  if (hipHostMalloc(&r.h_offs, n * 8, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&r.h_ns, n * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&r.h_poffs, n * 8, hipHostMallocDefault) != hipSuccess)
    return LUB_EHIP;
  return LUB_OK;
}

void rule_free(Rule& r) {
  if (r.stream) (void)hipStreamSynchronize(r.stream);
  for (Batch& b : r.inflight)
    if (b.done) (void)hipEventDestroy(b.done);
  if (r.plan) lu_ragged_plan_destroy(r.plan);
  if (r.d_offs) (void)hipFree(r.d_offs);
  if (r.d_ns) (void)hipFree(r.d_ns);
  if (r.d_poffs) (void)hipFree(r.d_poffs);
  if (r.h_offs) (void)hipHostFree(r.h_offs);
  if (r.h_ns) (void)hipHostFree(r.h_ns);
  if (r.h_poffs) (void)hipHostFree(r.h_poffs);
  if (r.stream) (void)hipStreamDestroy(r.stream);
  r = Rule{};
}

// Arena space for `bytes` at the head, skipping the ring's end if the block
// would straddle it.  Returns the byte offset or -1 when it does not fit now.
int64_t arena_alloc(lub_server* s, int64_t bytes, int64_t* new_head) {
  int64_t pos = s->head;
  const int64_t at = pos % s->arena_bytes;
  if (at + bytes > s->arena_bytes) pos += s->arena_bytes - at;  // skip to the ring's start
  if (pos + bytes - s->tail > s->arena_bytes) return -1;
  *new_head = pos + bytes;
  return pos % s->arena_bytes;
}

// Reclaim arena space of released requests in submission order and drop
// them from the table.
void reclaim(lub_server* s) {
  while (!s->reqs.empty() && s->reqs.front().released) {
    s->tail = s->reqs.front().block_end;
    s->reqs.pop_front();
    ++s->first_ticket;
  }
  // empty: restart at the arena's beginning (otherwise a block that has to
  // skip the ring's end could wait for space that is already free)
  if (s->reqs.empty()) s->head = s->tail = 0;
}

int retire(lub_server* s, int* completed) {
  const Clock::time_point now = Clock::now();
  for (Rule& r : s->rules) {
    while (!r.inflight.empty()) {
      Batch& b = r.inflight.front();
      const hipError_t q = hipEventQuery(b.done);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) return LUB_EHIP;
      for (uint64_t t : b.tickets) {
        Request* rq = find(s, t);
        rq->state = kDone;
        rq->t_done = now;
      }
      if (completed) *completed += static_cast<int>(b.tickets.size());
      r.tail = b.slot_end;
      s->event_pool.push_back(b.done);
      r.inflight.pop_front();
    }
  }
  return LUB_OK;
}

int launch(lub_server* s, int rule, int* launched) {
  Rule& r = s->rules[rule];
  if (r.queued.empty() || static_cast<int>(r.inflight.size()) >= s->max_inflight) return LUB_OK;
  Batch b;
  const Request* first = find(s, r.queued.front());
  const int64_t begin = first->slot;
  int64_t end = begin;
  const int flags = first->flags;
  while (!r.queued.empty()) {
    const Request* rq = find(s, r.queued.front());
    // a batch is one contiguous run of ring slots with one set of flags
    if (rq->slot != end || rq->flags != flags || end - begin + rq->count > s->max_batch) break;
    if ((begin % s->ring) + (end - begin) + rq->count > s->ring) break;
    end += rq->count;
    b.tickets.push_back(rq->ticket);
    r.queued.pop_front();
  }
  if (b.tickets.empty()) return LUB_OK;
  hipEvent_t copied = take_event(s);
  b.done = take_event(s);
  if (!copied || !b.done) return LUB_EHIP;
  // the batch's slot entries in one upload per array (staged by submit)
  const int64_t at = begin % s->ring;
  const size_t cnt = static_cast<size_t>(end - begin);
  if (hipMemcpyAsync(r.d_offs + at, r.h_offs + at, cnt * 8, hipMemcpyHostToDevice, s->copy) != hipSuccess ||
      hipMemcpyAsync(r.d_ns + at, r.h_ns + at, cnt * 4, hipMemcpyHostToDevice, s->copy) != hipSuccess ||
      hipMemcpyAsync(r.d_poffs + at, r.h_poffs + at, cnt * 8, hipMemcpyHostToDevice, s->copy) != hipSuccess)
    return LUB_EHIP;
  if (hipEventRecord(copied, s->copy) != hipSuccess) return LUB_EHIP;
  if (hipStreamWaitEvent(r.stream, copied, 0) != hipSuccess) return LUB_EHIP;
  float* A = reinterpret_cast<float*>(s->arena);
  int* piv = rule == LUB_PIVOT_NONE ? nullptr : reinterpret_cast<int*>(s->arena);
  int rc = lu_batched_inverse_ragged_plan(r.plan, A, r.d_offs + at, r.d_ns + at, piv, piv ? r.d_poffs + at : nullptr,
                                          end - begin, flags, r.stream);
  if (rc != LUB_OK) return rc;
  if (hipEventRecord(b.done, r.stream) != hipSuccess) return LUB_EHIP;
  s->event_pool.push_back(copied);  // reusable once recorded work is ordered behind it
  for (uint64_t t : b.tickets) find(s, t)->state = kRunning;
  b.slot_end = end;
  r.inflight.push_back(std::move(b));
  if (launched) ++*launched;
  return LUB_OK;
}

}  // namespace

extern "C" {

int lu_server_create(lub_server** out, int64_t budget_bytes, int64_t max_batch, int64_t ring_slots) {
  if (out == nullptr) return LUB_ENULL;
  if (budget_bytes < kAlign || max_batch < 1 || ring_slots < max_batch) return LUB_EINVAL_ARG;
  auto* s = new (std::nothrow) lub_server;
  if (s == nullptr) return LUB_EHIP;
  s->arena_bytes = budget_bytes / kAlign * kAlign;
  s->max_batch = max_batch;
  s->ring = ring_slots;
  if (hipGetDevice(&s->device) != hipSuccess || hipMalloc(&s->arena, s->arena_bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess) {
    lu_server_destroy(s);
    return LUB_EHIP;
  }
  *out = s;
  return LUB_OK;
}

int lu_server_destroy(lub_server* s) {
  if (s == nullptr) return LUB_ENULL;
  if (s->copy) (void)hipStreamSynchronize(s->copy);
  for (Rule& r : s->rules) rule_free(r);
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->arena) (void)hipFree(s->arena);
  delete s;
  return LUB_OK;
}

int lu_server_submit(lub_server* s, const float* src, const int* ns, int count, int flags, uint64_t* ticket) {
  if (s == nullptr || src == nullptr || ns == nullptr || ticket == nullptr) return LUB_ENULL;
  if (count < 1 || count > s->max_batch) return LUB_EINVAL_BATCH;
  const int rule = flags & LUB_PIVOT_MASK;
  if ((flags & ~(LUB_PIVOT_MASK | LUB_LAYOUT_MASK)) != 0 || rule > LUB_PIVOT_GEPP ||
      (flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_MASK)
    return LUB_EFLAGS;
  int64_t floats = 0, pivs = 0;
  for (int i = 0; i < count; ++i) {
    if (ns[i] < 1 || ns[i] > LUB_MAX_N) return LUB_EINVAL_N;
    floats += int64_t(ns[i]) * ns[i];
    pivs += ns[i];
  }
  Rule& r = s->rules[rule];
  int rc = rule_init(s, r);
  if (rc != LUB_OK) return rc;
  // slot ring: contiguous run of `count` slots, skipping the ring's end
  if (r.head == r.tail) r.head = r.tail = 0;  // empty ring: restart at its beginning
  int64_t slot = r.head;
  if ((slot % s->ring) + count > s->ring) slot += s->ring - slot % s->ring;
  if (slot + count - r.tail > s->ring) return LUB_EBUSY;
  // arena block: matrices, then pivots
  const int64_t piv_at = round_up(floats * 4, 16);
  const int64_t bytes = round_up(piv_at + (rule == LUB_PIVOT_NONE ? 0 : pivs * 4), kAlign);
  if (bytes > s->arena_bytes) return LUB_EINVAL_ARG;
  int64_t new_head = 0;
  const int64_t block = arena_alloc(s, bytes, &new_head);
  if (block < 0) return LUB_EBUSY;

  if (hipMemcpyAsync(s->arena + block, src, size_t(floats) * 4, hipMemcpyDefault, s->copy) != hipSuccess)
    return LUB_EHIP;
  const int64_t at = slot % s->ring;
  int64_t fo = block / 4, po = (block + piv_at) / 4;
  for (int i = 0; i < count; ++i) {
    r.h_offs[at + i] = fo;
    r.h_ns[at + i] = ns[i];
    r.h_poffs[at + i] = po;
    fo += int64_t(ns[i]) * ns[i];
    po += ns[i];
  }

  Request rq;
  rq.ticket = s->next_ticket++;
  rq.rule = rule;
  rq.flags = flags;
  rq.count = count;
  rq.block = block;
  rq.block_end = new_head;
  rq.floats = floats;
  rq.slot = slot;
  rq.t_submit = Clock::now();
  s->head = new_head;
  r.head = slot + count;
  r.queued.push_back(rq.ticket);
  s->reqs.push_back(rq);
  *ticket = rq.ticket;
  return LUB_OK;
}

int lu_server_progress(lub_server* s, int* launched, int* completed) {
  if (s == nullptr) return LUB_ENULL;
  if (launched) *launched = 0;
  if (completed) *completed = 0;
  int rc = retire(s, completed);
  if (rc != LUB_OK) return rc;
  for (int rule = 0; rule < kRules; ++rule) {
    rc = launch(s, rule, launched);
    if (rc != LUB_OK) return rc;
  }
  return LUB_OK;
}

int lu_server_query(lub_server* s, uint64_t ticket, int* state, float** result, int** piv, double* latency_ms) {
  if (s == nullptr || state == nullptr) return LUB_ENULL;
  const Request* rq = find(s, ticket);
  if (rq == nullptr || rq->released) return LUB_EINVAL_ARG;
  *state = rq->state;
  if (result) *result = reinterpret_cast<float*>(s->arena + rq->block);
  if (piv)
    *piv = rq->rule == LUB_PIVOT_NONE ? nullptr
                                      : reinterpret_cast<int*>(s->arena + rq->block + round_up(rq->floats * 4, 16));
  if (latency_ms)
    *latency_ms = rq->state == kDone
                      ? std::chrono::duration<double, std::milli>(rq->t_done - rq->t_submit).count()
                      : -1.0;
  return LUB_OK;
}

int lu_server_release(lub_server* s, uint64_t ticket) {
  if (s == nullptr) return LUB_ENULL;
  Request* rq = find(s, ticket);
  if (rq == nullptr || rq->released || rq->state != kDone) return LUB_EINVAL_ARG;
  rq->released = true;
  reclaim(s);
  return LUB_OK;
}

int lu_server_stats(lub_server* s, int64_t* queued, int64_t* running, int64_t* resident_bytes) {
  if (s == nullptr) return LUB_ENULL;
  int64_t q = 0, run = 0;
  for (const Request& rq : s->reqs) {
    if (rq.released) continue;
    if (rq.state == kQueued) q += rq.count;
    if (rq.state == kRunning) run += rq.count;
  }
  if (queued) *queued = q;
  if (running) *running = run;
  if (resident_bytes) *resident_bytes = s->head - s->tail;
  return LUB_OK;
}

}  // extern "C"
