// Serving loop: continuous batching of encode requests on one device.
//
// Requests (device-resident squares, as zoda_square) queue in arrival order.
// Each of `slots` launch slots owns a stream, a completion event and a work
// counter.  Whenever a slot is free it takes the next launch:
//   * a large request (a side the device's tuning table runs direct: its own
//     fused kernel fills the device) at the head of the queue goes alone
//     through the single-square
//     kernel; at most slots - 1 of them run at once, in arrival order; while
//     the head is a large request that has to wait (for a slot or memory),
//     nothing passes it;
//   * otherwise (a small request at the head) every queued small request
//     that fits the HBM budget next to the requests on the GPU goes out as
//     one persistent batched launch whose descriptors travel in the kernel
//     arguments (nothing is copied before the launch; the counter is never
//     reset).
// Requests that arrive while others run are admitted as soon as a slot and
// memory free up, and their launches overlap the ones in flight.  Twiddle
// tables stay resident in the per-device cache.  Completion times come from
// the slot's event, measured on the GPU against an anchor event that is
// re-recorded every few seconds (a float of milliseconds since one fixed
// event would lose microseconds as the server stays up).
//
// Output memory: a request whose `out` is NULL gets its output from the
// server's device pool at admission.  The pool holds blocks cached by size;
// everything it holds, cached or in use, stays within the HBM budget, and a
// request waits in the queue while its block does not fit.  The block goes
// back to the pool at completion (trace replay) or when the caller releases
// it after reading it (zoda_server_output / zoda_server_release).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "encode_plans.hpp"

namespace zoda {

namespace {


double now_us() {
  using namespace std::chrono;
  return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

uint64_t output_bytes(const zoda_square& s) {
  const uint64_t n_kn = 1ull << log2_ceil((uint64_t)s.k + s.n);
  return n_kn * s.cols * 4;
}

uint64_t request_bytes(const zoda_square& s) { return (uint64_t)s.k * s.cols * 4 + output_bytes(s); }

// A request's shape checks; a NULL output asks for a pool block.
int check_request(const zoda_square& s, uint32_t i) {
  zoda_square shape = s;
  if (!shape.out) shape.out = const_cast<uint32_t*>(shape.in);  // validated as present below
  return check_batch_square(shape, i);
}

struct Request {
  zoda_square sq;
  double arrival_us = 0;   // host clock
  double done_us = -1;     // host clock, from the GPU event
  uint64_t bytes = 0;
  bool big = false;
  bool pooled = false;     // output from the pool (sq.out set at admission)
  bool release_on_done = false;
  bool polled = false;     // the caller has seen its final state
  int state = 0;           // 0 queued, 1 on the GPU, 2 done, 3 rejected
};

// Device blocks for server-owned outputs, cached by size.  held() counts every
// block allocated (in use or cached) and never exceeds the budget given to
// get(): when a new block does not fit, cached blocks of other sizes are
// freed first, and if it still does not fit the caller waits.
class OutPool {
 public:
  ~OutPool() { clear(); }

  // *out = a block of `bytes`, or nullptr when it does not fit the budget now.
  int get(uint64_t bytes, uint64_t budget, uint32_t** out) {
    *out = nullptr;
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
      *out = it->second.back();
      it->second.pop_back();
      return 0;
    }
    for (auto f = free_.begin(); held_ + bytes > budget && f != free_.end(); ++f) {
      while (!f->second.empty() && held_ + bytes > budget) {
        ZODA_HIP(hipFree(f->second.back()));
        f->second.pop_back();
        held_ -= f->first;
      }
    }
    if (held_ + bytes > budget) return 0;
    if (const hipError_t e = hipMalloc((void**)out, bytes)) {
      *out = nullptr;
      if (e == hipErrorOutOfMemory) {  // the device is fuller than the budget assumed: wait
        (void)hipGetLastError();
        return 0;
      }
      return alloc_fail(e, "serving output block");
    }
    held_ += bytes;
    peak_ = std::max(peak_, held_);
    return 0;
  }

  void put(uint32_t* p, uint64_t bytes) { free_[bytes].push_back(p); }

  void clear() {
    for (auto& f : free_)
      for (uint32_t* p : f.second) (void)hipFree(p);
    free_.clear();
    held_ = 0;
  }

  uint64_t held() const { return held_; }
  uint64_t peak() const { return peak_; }

 private:
  std::map<uint64_t, std::vector<uint32_t*>> free_;
  uint64_t held_ = 0, peak_ = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint32_t* d_counter = nullptr;    // work counter of this slot's batched launches
  uint32_t base = 0;                // its value when the next launch starts
  std::vector<uint64_t> members;    // request ids in this launch
  uint64_t bytes = 0;
  bool busy = false;
  bool big = false;
};

}  // namespace

class Server {
 public:
  ~Server() { release(); }

  int init(uint64_t budget, uint32_t nslots) {
    ZODA_TRY(current_state(&st_));
    budget_ = budget;
    tabs_.assign((BATCH_MAX_LOG + 1) * BATCH_VARIANTS, EncodeTabs{});
    have_.reset(new bool[tabs_.size()]());
    ZODA_HIP(hipMalloc((void**)&d_tabs_, tabs_.size() * sizeof(EncodeTabs)));
    ZODA_HIP(hipMemset(d_tabs_, 0, tabs_.size() * sizeof(EncodeTabs)));
    ZODA_HIP(hipEventCreate(&start_));
    slots_.resize(nslots);
    for (Slot& s : slots_) {
      ZODA_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
      ZODA_HIP(hipEventCreate(&s.done));
      if (const hipError_t ea = hipMalloc((void**)&s.d_counter, sizeof(uint32_t)))
        return alloc_fail(ea, "serving slot allocation");
      ZODA_HIP(hipMemset(s.d_counter, 0, sizeof(uint32_t)));
    }
    // The slot streams do not wait on the null stream: the zero fills must
    // be complete before the first launch on any of them.
    ZODA_HIP(hipDeviceSynchronize());
    ZODA_HIP(hipStreamCreateWithFlags(&clock_, hipStreamNonBlocking));
    return anchor();
  }

  // Reference point of the GPU clock: an event on the otherwise idle clock
  // stream and the host time right after it completes.  Completion events
  // are timed against the latest anchor, so the float milliseconds between
  // them stay small (microsecond resolution) however long the server runs.
  int anchor() {
    ZODA_HIP(hipEventRecord(start_, clock_));
    ZODA_HIP(hipEventSynchronize(start_));
    start_host_us_ = now_us();
    return 0;
  }

  // Thread-safe enqueue; returns the request id.
  // release_on_done: a pool output goes back to the pool at completion (no
  // caller reads it).
  int submit(const zoda_square& sq, double arrival_us, uint64_t* id, bool release_on_done = false) {
    ZODA_TRY(check_request(sq, 0));
    std::lock_guard<std::mutex> lock(mu_);
    Request r;
    r.sq = sq;
    r.arrival_us = arrival_us;
    r.bytes = request_bytes(sq);
    {
      std::lock_guard<std::recursive_mutex> dl(st_->mu);
      r.big = st_->batch_direct[log2_ceil(sq.k)];
    }
    r.pooled = sq.out == nullptr;
    r.release_on_done = r.pooled && release_on_done;
    if (!r.big) {  // resolve the batched plan and its tables now, off the launch path
      SquareDesc d;
      ZODA_TRY(describe_square(st_, sq, &d, tabs_.data(), have_.get()));
      ZODA_TRY(sync_tables());
    }
    if (r.bytes > budget_) r.state = 3;  // could never be admitted
    first_arrival_us_ = std::min(first_arrival_us_, arrival_us);
    reqs_.push_back(r);
    *id = first_id_ + reqs_.size() - 1;
    if (r.state == 0) queue_.push_back(*id);
    else rejected_++;
    return 0;
  }

  // One scheduler step: retire finished launches, then fill free slots.
  int pump() {
    std::lock_guard<std::mutex> lock(mu_);
    if (now_us() - start_host_us_ > ANCHOR_US) ZODA_TRY(anchor());
    for (Slot& s : slots_) {
      if (!s.busy) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) return fail((int)q, "serving launch failed: %s", hipGetErrorString(q));
      float ms = 0;
      ZODA_HIP(hipEventElapsedTime(&ms, start_, s.done));
      const double t = start_host_us_ + ms * 1000.0;
      for (uint64_t id : s.members) {
        Request& r = req(id);
        r.done_us = t;
        r.state = 2;
        last_done_us_ = std::max(last_done_us_, t);
        completed_++;
        if (r.release_on_done) give_back(r);
      }
      in_flight_bytes_ -= s.bytes;
      if (s.big) big_in_flight_--;
      s.members.clear();
      s.bytes = 0;
      s.busy = false;
      s.big = false;
    }
    for (Slot& s : slots_) {
      if (s.busy) continue;
      if (queue_.empty()) break;
      bool launched = false;
      ZODA_TRY(fill_slot(s, &launched));
      if (!launched) break;
    }
    retire();
    return 0;
  }

  bool idle() {
    std::lock_guard<std::mutex> lock(mu_);
    if (!queue_.empty()) return false;
    for (Slot& s : slots_)
      if (s.busy) return false;
    return true;
  }

  // Requests wait while nothing runs: only pool memory held by finished,
  // unreleased requests can keep the head out then.
  bool stalled() {
    std::lock_guard<std::mutex> lock(mu_);
    if (queue_.empty()) return false;
    for (Slot& s : slots_)
      if (s.busy) return false;
    return true;
  }

  int poll(uint64_t id, int* state, double* latency_us) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!known(id)) return fail(ZODA_EINVAL, "unknown or retired request %llu", (unsigned long long)id);
    Request& r = req(id);
    *state = r.state;
    *latency_us = r.state == 2 ? r.done_us - r.arrival_us : -1.0;
    if (r.state >= 2) r.polled = true;
    return 0;
  }

  // The pool output of a finished request (hand-off: valid until released).
  int output(uint64_t id, uint32_t** d_out) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!known(id)) return fail(ZODA_EINVAL, "unknown or retired request %llu", (unsigned long long)id);
    Request& r = req(id);
    if (r.state != 2) return fail(ZODA_ESTATE, "request %llu is not done", (unsigned long long)id);
    if (!r.pooled || !r.sq.out) return fail(ZODA_ESTATE, "request %llu holds no pool output", (unsigned long long)id);
    *d_out = r.sq.out;
    return 0;
  }

  // Returns a finished request's pool output to the pool (no-op for outputs
  // the caller owns or already released).
  int release_output(uint64_t id) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!known(id)) return fail(ZODA_EINVAL, "unknown or retired request %llu", (unsigned long long)id);
    Request& r = req(id);
    if (r.state == 0 || r.state == 1) return fail(ZODA_ESTATE, "request %llu is not done", (unsigned long long)id);
    give_back(r);
    return 0;
  }

  void fill_stats(zoda_serve_stats* s) {
    std::lock_guard<std::mutex> lock(mu_);
    s->batches = batches_;
    s->max_batch = max_batch_;
    s->peak_bytes = peak_bytes_;
    s->rejected = rejected_;
    s->makespan_us = last_done_us_ > first_arrival_us_ ? last_done_us_ - first_arrival_us_ : 0;
    s->pool_peak_bytes = pool_.peak();
    s->pool_held_bytes = pool_.held();
  }

  double latency(uint64_t id) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!known(id)) return -1.0;
    const Request& r = req(id);
    return r.state == 2 ? r.done_us - r.arrival_us : -1.0;
  }

  // Requests are retired (their records dropped) once finished, polled and
  // holding no pool output -- only with the online API (retire_ set), whose
  // callers poll each ticket; the trace replay reads every latency at the end.
  void set_retire(bool on) { retire_ = on; }

  // Background scheduler for the online API.
  void start_thread() {
    stop_ = false;
    thread_ = std::thread([this] {
      while (!stop_.load()) {
        if (pump() != 0) {
          std::lock_guard<std::mutex> lock(mu_);
          thread_error_ = true;
          break;
        }
        std::this_thread::yield();
      }
    });
  }

  void stop_thread() {
    if (thread_.joinable()) {
      stop_ = true;
      thread_.join();
    }
  }

  bool thread_error() {
    std::lock_guard<std::mutex> lock(mu_);
    return thread_error_;
  }

 private:
  static constexpr double ANCHOR_US = 5e6;

  bool known(uint64_t id) const { return id >= first_id_ && id - first_id_ < reqs_.size(); }
  Request& req(uint64_t id) { return reqs_[id - first_id_]; }

  void give_back(Request& r) {
    if (r.pooled && r.sq.out) {
      pool_.put(r.sq.out, output_bytes(r.sq));
      r.sq.out = nullptr;
    }
  }

  void retire() {
    if (!retire_) return;
    while (!reqs_.empty()) {
      const Request& r = reqs_.front();
      if (r.state < 2 || !r.polled || (r.pooled && r.sq.out)) break;
      reqs_.pop_front();
      first_id_++;
    }
  }

  // A pool block for a request that asks for one; false: it does not fit now.
  int admit_output(Request& r, bool* ok) {
    *ok = true;
    if (!r.pooled || r.sq.out) return 0;
    uint32_t* p = nullptr;
    ZODA_TRY(pool_.get(output_bytes(r.sq), budget_, &p));
    r.sq.out = p;
    *ok = p != nullptr;
    return 0;
  }

  // The next launch for free slot s (see the top of this file).  A large head
  // that waits blocks the queue: letting small requests pass it put their
  // batches on the GPU ahead of it and raised the p99 latency (the large
  // requests are the tail).
  int fill_slot(Slot& s, bool* launched) {
    *launched = false;
    const uint64_t head = queue_.front();
    if (req(head).big) {
      Request& r = req(head);
      const bool slot_free = slots_.size() == 1 || big_in_flight_ + 1 < slots_.size();
      bool ok = slot_free && in_flight_bytes_ + r.bytes <= budget_;
      if (ok) ZODA_TRY(admit_output(r, &ok));
      if (ok) {
        const zoda_square& q = r.sq;
        ZODA_TRY(launch_encode(st_, q.in, q.k, q.cols, q.n, q.out, ZODA_LAYOUT_ROW, s.stream));
        queue_.pop_front();
        s.big = true;
        big_in_flight_++;
        return commit(s, std::vector<uint64_t>{head}, r.bytes, launched);
      }
      return 0;
    }
    ServeBatch b{};
    std::vector<uint64_t> taken;
    uint64_t bytes = 0;
    for (auto it = queue_.begin(); it != queue_.end() && taken.size() < SERVE_MAX_BATCH;) {
      Request& r = req(*it);
      if (r.big) {
        ++it;
        continue;
      }
      if (in_flight_bytes_ + bytes + r.bytes > budget_) break;
      bool ok = true;
      ZODA_TRY(admit_output(r, &ok));
      if (!ok) break;
      bytes += r.bytes;
      taken.push_back(*it);
      it = queue_.erase(it);
    }
    if (taken.empty()) return 0;
    // Largest sides first inside the launch: long tiles start early, short
    // ones fill the tail.
    std::stable_sort(taken.begin(), taken.end(), [&](uint64_t x, uint64_t y) {
      return req(x).sq.k > req(y).sq.k;
    });
    uint64_t items = 0;
    uint32_t lds_bytes = 4, max_log = 0;
    for (uint32_t q = 0; q < taken.size(); q++) {
      ZODA_TRY(describe_square(st_, req(taken[q]).sq, &b.sq[q], tabs_.data(), have_.get()));
      b.first[q] = (uint32_t)items;
      items += b.sq[q].tiles;
      lds_bytes = std::max(lds_bytes, batch_lds_bytes(b.sq[q].L, b.sq[q].variant));
      max_log = std::max(max_log, b.sq[q].L);
    }
    if (items >= (1ull << 31)) return fail(ZODA_EINVAL, "batch of %llu tiles is too large", (unsigned long long)items);
    b.n = (uint32_t)taken.size();
    b.first[b.n] = (uint32_t)items;
    b.items = (uint32_t)items;
    b.base = s.base;
    b.counter = s.d_counter;
    b.tabs = d_tabs_;
    ZODA_TRY(sync_tables());
    uint32_t advance = 0, grid = 0;
    ZODA_TRY(batch_grid(st_, lds_bytes, max_log, &grid));
    ZODA_TRY(launch_serve_batch(b, grid, lds_bytes, max_log, s.stream, &advance));
    s.base += advance;
    return commit(s, std::move(taken), bytes, launched);
  }

  int commit(Slot& s, std::vector<uint64_t> members, uint64_t bytes, bool* launched) {
    ZODA_HIP(hipEventRecord(s.done, s.stream));
    for (uint64_t id : members) req(id).state = 1;
    max_batch_ = std::max(max_batch_, (uint32_t)members.size());
    s.members = std::move(members);
    s.bytes = bytes;
    s.busy = true;
    in_flight_bytes_ += bytes;
    peak_bytes_ = std::max(peak_bytes_, in_flight_bytes_);
    batches_++;
    *launched = true;
    return 0;
  }

  // The device copy of the table pointers, indexed by (side, variant), only
  // grows.  A new entry is uploaded synchronously, before any launch that
  // names it: launches on every slot stream read the array, and a copy
  // ordered on one stream would not be ordered before a kernel on another.
  // Entries already on the device are rewritten with the same bytes, so a
  // kernel reading them meanwhile is unaffected.
  int sync_tables() {
    size_t n = 0;
    for (size_t i = 0; i < tabs_.size(); i++) n += have_[i];
    if (n == tabs_uploaded_) return 0;
    ZODA_HIP(hipMemcpy(d_tabs_, tabs_.data(), tabs_.size() * sizeof(EncodeTabs), hipMemcpyHostToDevice));
    tabs_uploaded_ = n;
    return 0;
  }

  void release() {
    stop_thread();
    for (Slot& s : slots_) {
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      (void)hipFree(s.d_counter);
    }
    slots_.clear();
    for (Request& r : reqs_) give_back(r);
    pool_.clear();
    if (start_) (void)hipEventDestroy(start_);
    start_ = nullptr;
    if (clock_) (void)hipStreamDestroy(clock_);
    clock_ = nullptr;
    (void)hipFree(d_tabs_);
    d_tabs_ = nullptr;
  }

  DeviceState* st_ = nullptr;
  uint64_t budget_ = 0, in_flight_bytes_ = 0, peak_bytes_ = 0;
  uint32_t batches_ = 0, max_batch_ = 0, rejected_ = 0;
  size_t big_in_flight_ = 0;
  uint64_t completed_ = 0;
  std::vector<EncodeTabs> tabs_;
  std::unique_ptr<bool[]> have_;
  size_t tabs_uploaded_ = 0;
  EncodeTabs* d_tabs_ = nullptr;
  hipEvent_t start_ = nullptr;
  hipStream_t clock_ = nullptr;
  double start_host_us_ = 0;
  double first_arrival_us_ = 1e300, last_done_us_ = 0;
  OutPool pool_;
  bool retire_ = false;
  uint64_t first_id_ = 0;
  std::vector<Slot> slots_;
  std::deque<Request> reqs_;
  std::deque<uint64_t> queue_;
  std::mutex mu_;
  std::thread thread_;
  std::atomic<bool> stop_{false};
  bool thread_error_ = false;
};

}  // namespace zoda

using namespace zoda;

struct zoda_server {
  Server impl;
};

extern "C" int zoda_server_create(uint64_t hbm_budget, uint32_t slots, zoda_server** out) {
  clear_error();
  if (!out || slots == 0 || slots > 16 || hbm_budget == 0) return fail(ZODA_EINVAL, "bad arguments");
  *out = nullptr;
  zoda_server* s = new zoda_server();
  int rc = s->impl.init(hbm_budget, slots);
  if (rc) {
    delete s;
    return rc;
  }
  s->impl.set_retire(true);
  s->impl.start_thread();
  *out = s;
  return 0;
}

extern "C" int zoda_server_submit(zoda_server* srv, const zoda_square* req, uint64_t* ticket) {
  clear_error();
  if (!srv || !req || !ticket) return fail(ZODA_EINVAL, "bad arguments");
  if (srv->impl.thread_error()) return fail(ZODA_ESTATE, "the scheduler stopped after an error");
  const int rc = srv->impl.submit(*req, now_us(), ticket);
  if (rc) return rc;
  if (request_bytes(*req) > 0 && srv->impl.latency(*ticket) < 0) {
    int state = 0;
    double lat = 0;
    srv->impl.poll(*ticket, &state, &lat);
    if (state == 3) return fail(ZODA_EBUDGET, "request of %llu bytes exceeds the HBM budget",
                                (unsigned long long)request_bytes(*req));
  }
  return 0;
}

// state: 0 queued, 1 on the GPU, 2 done, 3 rejected; latency_us valid when done.
extern "C" int zoda_server_poll(zoda_server* srv, uint64_t ticket, int* state, double* latency_us) {
  clear_error();
  if (!srv || !state || !latency_us) return fail(ZODA_EINVAL, "bad arguments");
  return srv->impl.poll(ticket, state, latency_us);
}

extern "C" int zoda_server_output(zoda_server* srv, uint64_t ticket, uint32_t** d_out) {
  clear_error();
  if (!srv || !d_out) return fail(ZODA_EINVAL, "bad arguments");
  return srv->impl.output(ticket, d_out);
}

extern "C" int zoda_server_release(zoda_server* srv, uint64_t ticket) {
  clear_error();
  if (!srv) return fail(ZODA_EINVAL, "bad arguments");
  return srv->impl.release_output(ticket);
}

extern "C" int zoda_server_wait_all(zoda_server* srv) {
  clear_error();
  if (!srv) return fail(ZODA_EINVAL, "bad arguments");
  int stalled_for = 0;
  while (!srv->impl.idle()) {
    if (srv->impl.thread_error()) return fail(ZODA_ESTATE, "the scheduler stopped after an error");
    // a queue that stays put with nothing on the GPU waits for pool memory
    // that only zoda_server_release can free: report it instead of waiting
    stalled_for = srv->impl.stalled() ? stalled_for + 1 : 0;
    if (stalled_for > 2000)
      return fail(ZODA_EBUDGET, "queued requests wait for pool memory held by finished requests; "
                                "release their outputs (zoda_server_release)");
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  return 0;
}

extern "C" int zoda_server_stats(zoda_server* srv, zoda_serve_stats* stats) {
  clear_error();
  if (!srv || !stats) return fail(ZODA_EINVAL, "bad arguments");
  srv->impl.fill_stats(stats);
  return 0;
}

extern "C" int zoda_server_destroy(zoda_server* srv) {
  clear_error();
  if (srv) {
    srv->impl.stop_thread();
    delete srv;
  }
  return 0;
}

// Trace replay: requests arrive at arrival_us[i] (microseconds after the
// start, non-decreasing); this thread plays both the arrival process and the
// scheduler.  mode 0 is the continuous-batching server above; mode 1 is the
// plain loop, one single-square encode per request in arrival order on one
// stream.
extern "C" int zoda_hip_serve_trace(const zoda_square* reqs, const double* arrival_us, uint32_t count,
                                    uint64_t hbm_budget, uint32_t slots, int mode,
                                    double* latency_us, zoda_serve_stats* stats) {
  clear_error();
  if (!reqs || !arrival_us || !latency_us || count == 0 || (mode != 0 && mode != 1))
    return fail(ZODA_EINVAL, "bad arguments");
  for (uint32_t i = 0; i < count; i++) {
    ZODA_TRY(check_request(reqs[i], i));
    if (i && arrival_us[i] < arrival_us[i - 1]) return fail(ZODA_EINVAL, "arrivals must be sorted");
  }
  if (mode == 0) {
    if (slots == 0 || slots > 16) return fail(ZODA_EINVAL, "slots must be 1..16");
    Server srv;
    ZODA_TRY(srv.init(hbm_budget, slots));
    const double t0 = now_us();
    uint32_t next = 0;
    std::vector<uint64_t> ids(count);
    while (next < count || !srv.idle()) {
      const double t = now_us() - t0;
      while (next < count && arrival_us[next] <= t) {
        ZODA_TRY(srv.submit(reqs[next], t0 + arrival_us[next], &ids[next], true));
        next++;
      }
      ZODA_TRY(srv.pump());
    }
    for (uint32_t i = 0; i < count; i++) latency_us[i] = srv.latency(ids[i]);
    if (stats) srv.fill_stats(stats);
    return 0;
  }
  // plain loop
  DeviceState* st = nullptr;
  ZODA_TRY(current_state(&st));
  hipStream_t s;
  ZODA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::vector<hipEvent_t> ev(count + 1, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (!rc) hip_note(rc, hipEventCreate(&e), "event creation");
  // Requests without an output write to one block per output size, allocated
  // before the start: the loop runs one request at a time on one stream, so
  // consecutive requests of a size never overlap on the block.
  std::map<uint64_t, uint32_t*> scratch;
  uint64_t scratch_bytes = 0;
  for (uint32_t i = 0; i < count && !rc; i++) {
    if (reqs[i].out) continue;
    const uint64_t b = output_bytes(reqs[i]);
    if (scratch.count(b)) continue;
    uint32_t* p = nullptr;
    if (const hipError_t ea = hipMalloc((void**)&p, b)) {
      rc = alloc_fail(ea, "plain-loop output block");
      break;
    }
    scratch[b] = p;
    scratch_bytes += b;
  }
  double start_host = 0, t0 = 0;
  if (!rc) hip_note(rc, hipEventRecord(ev[count], s), "start event");
  if (!rc) hip_note(rc, hipEventSynchronize(ev[count]), "start event");
  if (!rc) {
    start_host = now_us();
    t0 = start_host;
  }
  uint64_t peak = 0;
  for (uint32_t i = 0; i < count && !rc; i++) {
    while (now_us() - t0 < arrival_us[i]) {
    }
    const zoda_square& q = reqs[i];
    uint32_t* out = q.out ? q.out : scratch[output_bytes(q)];
    rc = launch_encode(st, q.in, q.k, q.cols, q.n, out, ZODA_LAYOUT_ROW, s);
    if (!rc) hip_note(rc, hipEventRecord(ev[i], s), "event record");
    peak = std::max(peak, request_bytes(q));
  }
  if (!rc) hip_note(rc, hipStreamSynchronize(s), "plain loop");
  double last = 0;
  for (uint32_t i = 0; i < count && !rc; i++) {
    float ms = 0;
    hip_note(rc, hipEventElapsedTime(&ms, ev[count], ev[i]), "event timing");
    if (rc) break;
    const double done = start_host + ms * 1000.0;
    latency_us[i] = done - (t0 + arrival_us[i]);
    last = std::max(last, done);
  }
  if (!rc && stats) {
    stats->makespan_us = last - (t0 + arrival_us[0]);
    stats->batches = count;
    stats->max_batch = 1;
    stats->peak_bytes = peak;
    stats->rejected = 0;
    stats->pool_peak_bytes = scratch_bytes;
    stats->pool_held_bytes = scratch_bytes;
  }
  (void)hipStreamSynchronize(s);
  for (auto& b : scratch) (void)hipFree(b.second);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s);
  return rc;
}
