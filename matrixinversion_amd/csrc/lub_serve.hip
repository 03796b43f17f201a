// lub_serve.hip -- continuous-batching server over the ragged path
// (include/lubatch.h, lu_server_*).
//
// The reference has no serving layer: its run.py compiles, launches and times
// one batch per process (parallel_pivot/run.py:198-234).  Here requests of
// any size mix arrive one at a time and are inverted in HBM-resident batches:
//
//   create    everything the server will use is made here, so that no
//             request ever waits for an allocation: the arena (the HBM
//             budget), a copy stream, kSlots launch slots (a compute stream,
//             a ragged plan for max_batch matrices and two events each), and
//             per pivot rule a device slot ring plus its pinned staging ring.
//   submit    a request (count matrices, packed row-major, sizes ns[]) is
//             admitted if its block fits the HBM budget: the block (its
//             matrices, then its pivot vectors) is carved from a ring arena in
//             submission order and filled by an async copy on the copy stream;
//             the request's slot entries (matrix offset, n, pivot offset) are
//             written to the pinned staging ring of its pivot rule.
//             Otherwise LUB_EBUSY: the caller keeps the request and retries
//             after progress/release.
//   progress  retires batches whose completion event has fired, then, while a
//             launch slot is free, launches the queued requests of a pivot
//             rule (the front run of its ring) as one ragged batch on that
//             slot: the batch's slot entries go up in one copy per array, the
//             slot's stream waits for the copies, runs the ragged plan and
//             records the batch's completion event.  Batches on different
//             slots -- of one rule or of several -- run concurrently.
//   poll      the tickets finished since the last poll, in completion order.
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
#include "lubatch_util.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kRules = 5;                 // LUB_PIVOT_NONE .. LUB_PIVOT_GEPP
constexpr int kSlots = 3;                 // batches in flight at once (+ the copy stream = 4 streams)
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

// A launch slot: one batch at a time on its own stream and plan.
struct Slot {
  hipStream_t stream = nullptr;
  lub_ragged_plan* plan = nullptr;
  hipEvent_t copied = nullptr;   // the batch's slot entries are on the device
  hipEvent_t done = nullptr;     // the batch is inverted
  bool busy = false;
  int rule = -1;
  int64_t ring_end = 0;          // monotonic ring position after the batch
  std::vector<uint64_t> tickets;
};

// A launched batch's run of its rule's ring, in launch (= ring) order.
struct RingRun {
  int64_t end = 0;
  bool done = false;
};

struct Rule {
  int64_t* d_offs = nullptr;      // device slot ring
  int* d_ns = nullptr;
  int64_t* d_poffs = nullptr;
  int64_t* h_offs = nullptr;      // pinned staging ring
  int* h_ns = nullptr;
  int64_t* h_poffs = nullptr;
  int64_t head = 0, tail = 0;     // monotonic slot positions
  std::deque<uint64_t> queued;    // tickets in submission order
  std::deque<RingRun> runs;       // launched batches not yet reclaimed
  int inflight = 0;               // batches of this rule running now
};

}  // namespace

struct lub_server {
  int device = 0;
  char* arena = nullptr;
  int64_t arena_bytes = 0;
  int64_t head = 0, tail = 0;     // monotonic arena positions
  int64_t max_batch = 0;
  int64_t ring = 0;               // slots per rule ring
  hipStream_t copy = nullptr;
  Slot slots[kSlots];
  Rule rules[kRules];
  std::deque<Request> reqs;       // reqs[i].ticket == first_ticket + i
  uint64_t first_ticket = 1;
  uint64_t next_ticket = 1;
  std::vector<uint64_t> finished; // done since the last poll
  int64_t batches = 0;            // launched in all
  int peak_rule_inflight = 0;     // most batches of one rule running at once
};

namespace {

Request* find(lub_server* s, uint64_t t) {
  if (t < s->first_ticket || t >= s->next_ticket) return nullptr;
  return &s->reqs[t - s->first_ticket];
}

int slot_init(lub_server* s, Slot& sl) {
  if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) return LUB_EHIP;
  int rc = lu_ragged_plan_create(&sl.plan, s->max_batch);
  if (rc != LUB_OK) return rc;
  if (hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
    return LUB_EHIP;
  sl.tickets.reserve(1024);
  return LUB_OK;
}

void slot_free(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  if (sl.plan) lu_ragged_plan_destroy(sl.plan);
  if (sl.copied) (void)hipEventDestroy(sl.copied);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  sl = Slot{};
}

int rule_init(lub_server* s, Rule& r) {
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
  if (r.d_offs) (void)hipFree(r.d_offs);
  if (r.d_ns) (void)hipFree(r.d_ns);
  if (r.d_poffs) (void)hipFree(r.d_poffs);
  if (r.h_offs) (void)hipHostFree(r.h_offs);
  if (r.h_ns) (void)hipHostFree(r.h_ns);
  if (r.h_poffs) (void)hipHostFree(r.h_poffs);
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
  for (Slot& sl : s->slots) {
    if (!sl.busy) continue;
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return LUB_EHIP;
    for (uint64_t t : sl.tickets) {
      Request* rq = find(s, t);
      rq->state = kDone;
      rq->t_done = now;
      s->finished.push_back(t);
    }
    if (completed) *completed += static_cast<int>(sl.tickets.size());
    Rule& r = s->rules[sl.rule];
    for (RingRun& run : r.runs)
      if (run.end == sl.ring_end) run.done = true;
    // ring space comes back in ring order
    while (!r.runs.empty() && r.runs.front().done) {
      r.tail = r.runs.front().end;
      r.runs.pop_front();
    }
    --r.inflight;
    sl.busy = false;
    sl.tickets.clear();
  }
  return LUB_OK;
}

// Launch the front run of `rule`'s queue on `sl`.  Tickets leave the queue
// only once every enqueue has succeeded; on an error they stay queued (and
// the slot stays free), so a later progress call retries them.
int launch(lub_server* s, int rule, Slot& sl, int* launched) {
  Rule& r = s->rules[rule];
  const Request* first = find(s, r.queued.front());
  const int64_t begin = first->slot;
  const int flags = first->flags;
  int64_t end = begin;
  size_t take = 0;
  for (uint64_t t : r.queued) {
    const Request* rq = find(s, t);
    // a batch is one contiguous run of ring slots with one set of flags
    if (rq->slot != end || rq->flags != flags || end - begin + rq->count > s->max_batch) break;
    if ((begin % s->ring) + (end - begin) + rq->count > s->ring) break;
    end += rq->count;
    ++take;
  }
  // the batch's slot entries in one upload per array (staged by submit)
  const int64_t at = begin % s->ring;
  const size_t cnt = static_cast<size_t>(end - begin);
  if (hipMemcpyAsync(r.d_offs + at, r.h_offs + at, cnt * 8, hipMemcpyHostToDevice, s->copy) != hipSuccess ||
      hipMemcpyAsync(r.d_ns + at, r.h_ns + at, cnt * 4, hipMemcpyHostToDevice, s->copy) != hipSuccess ||
      hipMemcpyAsync(r.d_poffs + at, r.h_poffs + at, cnt * 8, hipMemcpyHostToDevice, s->copy) != hipSuccess)
    return LUB_EHIP;
  if (hipEventRecord(sl.copied, s->copy) != hipSuccess) return LUB_EHIP;
  if (hipStreamWaitEvent(sl.stream, sl.copied, 0) != hipSuccess) return LUB_EHIP;
  float* A = reinterpret_cast<float*>(s->arena);
  int* piv = rule == LUB_PIVOT_NONE ? nullptr : reinterpret_cast<int*>(s->arena);
  int rc = lu_batched_inverse_ragged_plan(sl.plan, A, r.d_offs + at, r.d_ns + at, piv,
                                          piv ? r.d_poffs + at : nullptr, end - begin, flags, sl.stream);
  if (rc != LUB_OK) return rc;  // the inversion kernel did not start: the batch can run again later
  // From here the batch is committed (its matrices are being inverted in
  // place, so it must never be launched a second time).
  const bool recorded = hipEventRecord(sl.done, sl.stream) == hipSuccess;
  if (!recorded) (void)hipStreamSynchronize(sl.stream);  // then `done` (never recorded) reads as complete

  sl.tickets.clear();
  for (size_t i = 0; i < take; ++i) {
    const uint64_t t = r.queued.front();
    r.queued.pop_front();
    find(s, t)->state = kRunning;
    sl.tickets.push_back(t);
  }
  sl.busy = true;
  sl.rule = rule;
  sl.ring_end = end;
  r.runs.push_back(RingRun{end, false});
  ++r.inflight;
  ++s->batches;
  s->peak_rule_inflight = std::max(s->peak_rule_inflight, r.inflight);
  if (launched) ++*launched;
  return recorded ? LUB_OK : LUB_EHIP;
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
  int rc = LUB_OK;
  if (hipGetDevice(&s->device) != hipSuccess || hipMalloc(&s->arena, s->arena_bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess)
    rc = LUB_EHIP;
  for (int i = 0; rc == LUB_OK && i < kSlots; ++i) rc = slot_init(s, s->slots[i]);
  for (int i = 0; rc == LUB_OK && i < kRules; ++i) rc = rule_init(s, s->rules[i]);
  if (rc != LUB_OK) {
    lu_server_destroy(s);
    return rc;
  }
  *out = s;
  return LUB_OK;
}

int lu_server_destroy(lub_server* s) {
  if (s == nullptr) return LUB_ENULL;
  if (s->copy) (void)hipStreamSynchronize(s->copy);
  for (Slot& sl : s->slots) slot_free(sl);
  for (Rule& r : s->rules) rule_free(r);
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
  // Free slots go to the rules with queued work, the rule with the fewest
  // batches running first (ties: lower rule number).
  for (Slot& sl : s->slots) {
    if (sl.busy) continue;
    int pick = -1;
    for (int rule = 0; rule < kRules; ++rule) {
      if (s->rules[rule].queued.empty()) continue;
      if (pick < 0 || s->rules[rule].inflight < s->rules[pick].inflight) pick = rule;
    }
    if (pick < 0) break;
    rc = launch(s, pick, sl, launched);
    if (rc != LUB_OK) return rc;
  }
  return LUB_OK;
}

int lu_server_poll(lub_server* s, uint64_t* tickets, int capacity, int* count) {
  if (s == nullptr || count == nullptr || (tickets == nullptr && capacity > 0)) return LUB_ENULL;
  if (capacity < 0) return LUB_EINVAL_ARG;
  const int take = static_cast<int>(std::min<size_t>(s->finished.size(), static_cast<size_t>(capacity)));
  std::copy(s->finished.begin(), s->finished.begin() + take, tickets);
  s->finished.erase(s->finished.begin(), s->finished.begin() + take);
  *count = take;
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

int lu_server_counters(lub_server* s, int64_t* batches, int* peak_rule_inflight) {
  if (s == nullptr) return LUB_ENULL;
  if (batches) *batches = s->batches;
  if (peak_rule_inflight) *peak_rule_inflight = s->peak_rule_inflight;
  return LUB_OK;
}


// Benchmark driver: replays a request schedule against `server` from a host
// loop with no interpreter in it (declared in lubatch_util.h).  Request i
// becomes due arrival_s[i] seconds after the call starts (ascending); it is
// count[i] matrices packed at src[i] (an address) with sizes
// ns[ns_off[i] .. ns_off[i] + count[i]) and flags[i].  The loop submits every
// due request (LUB_EBUSY: the rest wait for the next pass), calls progress,
// and on completions polls and releases; latency_ms[i] is the time from
// arrival_s[i] to the pass that saw request i done.  Tickets of requests
// submitted before the call are taken off the poll queue but not released
// (lu_server_query still reports them).  stats[0] = batches
// launched, stats[1] = LUB_EBUSY refusals.  Gives up with LUB_EHIP if the
// stream has not drained 30 s after its last arrival.
int lubx_serve_replay(lub_server* s, int n_req, const double* arrival_s, const uint64_t* src, const int* ns,
                      const int64_t* ns_off, const int* count, const int* flags, double* latency_ms, int64_t* stats) {
  if (s == nullptr || arrival_s == nullptr || src == nullptr || ns == nullptr || ns_off == nullptr ||
      count == nullptr || flags == nullptr || latency_ms == nullptr || stats == nullptr)
    return LUB_ENULL;
  if (n_req < 0) return LUB_EINVAL_ARG;
  std::vector<std::pair<uint64_t, int>> live;  // (ticket, request), few at a time
  std::vector<uint64_t> polled(256);
  const double give_up = (n_req > 0 ? arrival_s[n_req - 1] : 0.0) + 30.0;
  const Clock::time_point t0 = Clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  int nxt = 0, done = 0;
  int64_t batches = 0, busy = 0;
  while (done < n_req) {
    const double now = elapsed();
    if (now > give_up) return LUB_EHIP;
    while (nxt < n_req && arrival_s[nxt] <= now) {
      uint64_t t = 0;
      const int rc = lu_server_submit(s, reinterpret_cast<const float*>(src[nxt]), ns + ns_off[nxt], count[nxt],
                                      flags[nxt], &t);
      if (rc == LUB_EBUSY) {
        ++busy;
        break;
      }
      if (rc != LUB_OK) return rc;
      live.emplace_back(t, nxt++);
    }
    int launched = 0, completed = 0;
    if (int rc = lu_server_progress(s, &launched, &completed); rc != LUB_OK) return rc;
    batches += launched;
    if (completed > 0) {
      const double seen = elapsed();
      for (;;) {
        int got = 0;
        if (int rc = lu_server_poll(s, polled.data(), static_cast<int>(polled.size()), &got); rc != LUB_OK) return rc;
        for (int k = 0; k < got; ++k) {
          auto it = std::find_if(live.begin(), live.end(), [&](const auto& e) { return e.first == polled[k]; });
          if (it == live.end()) continue;  // a ticket the caller submitted before the replay: left to the caller
          latency_ms[it->second] = (seen - arrival_s[it->second]) * 1e3;
          if (int rc = lu_server_release(s, it->first); rc != LUB_OK) return rc;
          *it = live.back();
          live.pop_back();
          ++done;
        }
        if (got < static_cast<int>(polled.size())) break;
      }
    }
  }
  stats[0] = batches;
  stats[1] = busy;
  return LUB_OK;
}

}  // extern "C"
