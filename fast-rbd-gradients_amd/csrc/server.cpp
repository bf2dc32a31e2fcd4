// Serving loop (rbd_server_*): continuous batching of forward-dynamics-gradient
// requests onto the device.
//
// A request is one trajectory: host input n x 21, host output n x 105, both
// caller-owned and valid until the request completes.  The dispatcher thread
// admits queued requests in arrival order while the fixed memory budget
// allows: a batch holds one contiguous range [inputs of all its requests |
// outputs of all its requests] (n x 504 bytes per request, the two halves
// rounded up to 256 bytes) of device memory, and the same range of pinned
// staging, until its results are delivered.  It runs the batch on a free
// in-flight slot (own stream, own rbd_batch work queue): inputs -> pinned
// staging, one H2D copy, one persistent-kernel launch, one D2H copy, an event.  While batches
// are on the device the dispatcher keeps admitting newly arrived requests into
// the next free slot, so work from several batches overlaps on the GPU.  The
// completion thread waits for slots in launch order, copies results to the
// callers' buffers, returns the memory to the budget and records each
// request's latency (submit -> result in the caller's buffer).
//
// Every batch of a server runs with the same knot-to-lane mapping (fixed at
// creation), so a request's result does not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "context.hpp"

using namespace rbdi;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

namespace rbdi {

// First-fit allocator over [0, size) with coalescing of freed ranges.
class RangeAllocator {
 public:
  explicit RangeAllocator(size_t size) : size_(size) { free_[0] = size; }
  size_t alloc(size_t n) {                 // offset, or SIZE_MAX if no range fits
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        used_ += n;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t n) {
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest() const {                 // longest free range
    size_t m = 0;
    for (const auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }
  size_t free_ranges() const { return free_.size(); }

 private:
  size_t size_;
  size_t used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Request {
  uint64_t id = 0;
  int64_t n = 0;
  const float* h_in = nullptr;   // caller's buffers, valid until completion
  float* h_out = nullptr;
  size_t in_off = 0, out_off = 0;  // offsets of its input and output in the pools
  size_t in_bytes = 0, out_bytes = 0, bytes = 0;
  Clock::time_point t_submit;
  int status = 0;                // 0 pending, 1 done, -1 failed
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  rbd_batch* batch = nullptr;
  std::vector<std::shared_ptr<Request>> reqs;
  size_t off = 0, bytes = 0;            // the batch's range in the pools
  size_t in_bytes = 0, out_bytes = 0;   // packed inputs, packed outputs
  bool busy = false;
};

}  // namespace rbdi

struct rbd_server {
  rbd_context* ctx = nullptr;          // the caller's context (device, flags, precision)
  rbd_context* run_ctx = nullptr;      // same, with the knot-to-lane mapping fixed
  rbd_server_config cfg{};
  char* d_pool = nullptr;              // device memory of the budget
  char* h_pool = nullptr;              // pinned staging of the same size
  std::unique_ptr<RangeAllocator> alloc;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_dispatch;  // new request, returned memory, freed slot, stop
  std::condition_variable cv_complete;  // a slot was launched, stop
  std::condition_variable cv_done;      // a request finished
  std::deque<std::shared_ptr<Request>> queue;                    // submitted, not admitted
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;    // not yet collected
  std::deque<int> launched;                                       // slot indices in launch order
  uint64_t next_id = 1;
  uint64_t mem_gen = 0;                 // bumped whenever memory or a slot is returned
  bool stop = false;
  rbd_status async_error = RBD_OK;

  // statistics
  std::vector<double> latencies_us;
  int64_t knots_done = 0, requests_done = 0, batches = 0, inflight_max = 0, admitted_knots_max = 0;
  size_t budget_used_max = 0;
  Clock::time_point t_first{}, t_last{};
  bool any_done = false;

  std::thread dispatcher, completer;
};

namespace {

void dispatch_loop(rbd_server* s) {
  if (use_device(s->ctx) != RBD_OK) return;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_dispatch.wait(lk, [s] {
      if (s->stop) return true;
      if (s->queue.empty()) return false;
      for (const Slot& sl : s->slots)
        if (!sl.busy) return true;
      return false;
    });
    if (s->stop) return;
    // optional batching window: let requests right behind the head join
    if (s->cfg.batch_window_us > 0) {
      const auto deadline = s->queue.front()->t_submit +
                            std::chrono::duration_cast<Clock::duration>(
                                std::chrono::duration<double, std::micro>(s->cfg.batch_window_us));
      s->cv_dispatch.wait_until(lk, deadline, [s] {
        return s->stop || int64_t(s->queue.size()) >= s->cfg.max_batch_trajectories;
      });
      if (s->stop) return;
    }
    int slot = -1;
    for (int i = 0; i < int(s->slots.size()); ++i)
      if (!s->slots[i].busy) { slot = i; break; }
    if (slot < 0) continue;
    // admit in arrival order while the budget and the batch limits allow; the
    // batch takes one range, so it must fit the longest free range
    std::vector<std::shared_ptr<Request>> admitted;
    int64_t knots = 0;
    size_t in_bytes = 0, out_bytes = 0;
    const size_t room = s->alloc->largest();
    for (const auto& r : s->queue) {
      if (int64_t(admitted.size()) >= s->cfg.max_batch_trajectories) break;
      if (knots + r->n > s->cfg.max_batch_knots) break;
      if (round_up(in_bytes + r->in_bytes) + round_up(out_bytes + r->out_bytes) > room) break;
      r->in_off = in_bytes;                                 // relative to the batch, fixed up below
      r->out_off = out_bytes;
      in_bytes += r->in_bytes;
      out_bytes += r->out_bytes;
      knots += r->n;
      admitted.push_back(r);
    }
    size_t off = SIZE_MAX;
    if (!admitted.empty()) off = s->alloc->alloc(round_up(in_bytes) + round_up(out_bytes));
    if (off == SIZE_MAX) {
      // the head request does not fit yet: sleep until memory comes back
      const uint64_t gen = s->mem_gen;
      s->cv_dispatch.wait(lk, [s, gen] { return s->stop || s->mem_gen != gen; });
      continue;
    }
    s->queue.erase(s->queue.begin(), s->queue.begin() + std::ptrdiff_t(admitted.size()));
    for (auto& r : admitted) {
      r->in_off += off;
      r->out_off += off + round_up(in_bytes);
    }
    s->budget_used_max = std::max(s->budget_used_max, s->alloc->used());
    s->admitted_knots_max = std::max(s->admitted_knots_max, knots);
    Slot& sl = s->slots[size_t(slot)];
    sl.busy = true;
    sl.reqs = std::move(admitted);
    sl.off = off;
    sl.bytes = round_up(in_bytes) + round_up(out_bytes);
    sl.in_bytes = in_bytes;
    sl.out_bytes = out_bytes;
    lk.unlock();
    // stage inputs, one H2D copy, one batched launch, one D2H copy, event
    std::vector<rbd_traj> trajs;
    trajs.reserve(sl.reqs.size());
    for (auto& r : sl.reqs) {
      std::memcpy(s->h_pool + r->in_off, r->h_in, r->in_bytes);
      trajs.push_back(rbd_traj{reinterpret_cast<const float*>(s->d_pool + r->in_off),
                               reinterpret_cast<float*>(s->d_pool + r->out_off), r->n});
    }
    const size_t out0 = off + round_up(in_bytes);
    hipError_t e = hipMemcpyAsync(s->d_pool + off, s->h_pool + off, in_bytes, hipMemcpyHostToDevice, sl.stream);
    rbd_status st = RBD_OK;
    if (e == hipSuccess)
      st = rbd_fdgrad_batched_f32(sl.batch, trajs.data(), int(trajs.size()), s->cfg.use_graph, sl.stream);
    if (e == hipSuccess && st == RBD_OK)
      e = hipMemcpyAsync(s->h_pool + out0, s->d_pool + out0, out_bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess && st == RBD_OK) e = hipEventRecord(sl.done, sl.stream);
    lk.lock();
    if (e != hipSuccess || st != RBD_OK) s->async_error = st != RBD_OK ? st : RBD_ERR_HIP;
    s->launched.push_back(slot);
    s->batches++;
    int64_t inflight = 0;
    for (const Slot& x : s->slots) inflight += x.busy ? 1 : 0;
    s->inflight_max = std::max(s->inflight_max, inflight);
    s->cv_complete.notify_all();
  }
}

void complete_loop(rbd_server* s) {
  if (use_device(s->ctx) != RBD_OK) return;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_complete.wait(lk, [s] { return s->stop || !s->launched.empty(); });
    if (s->launched.empty()) return;       // stop with nothing in flight
    const int slot = s->launched.front();
    Slot& sl = s->slots[size_t(slot)];
    const bool failed_launch = s->async_error != RBD_OK;
    lk.unlock();
    const hipError_t e = failed_launch ? hipErrorUnknown : hipEventSynchronize(sl.done);
    if (e == hipSuccess)
      for (auto& r : sl.reqs) std::memcpy(r->h_out, s->h_pool + r->out_off, r->out_bytes);
    const auto now = Clock::now();
    lk.lock();
    s->launched.pop_front();
    s->alloc->release(sl.off, sl.bytes);
    for (auto& r : sl.reqs) {
      r->status = (e == hipSuccess) ? 1 : -1;
      if (e == hipSuccess) {
        s->latencies_us.push_back(std::chrono::duration<double, std::micro>(now - r->t_submit).count());
        s->knots_done += r->n;
        s->requests_done++;
        if (!s->any_done) { s->t_first = r->t_submit; s->any_done = true; }
        s->t_last = now;
      }
    }
    if (e != hipSuccess && s->async_error == RBD_OK) s->async_error = RBD_ERR_HIP;
    sl.reqs.clear();
    sl.busy = false;
    s->mem_gen++;
    s->cv_done.notify_all();
    s->cv_dispatch.notify_all();
  }
}

double percentile(std::vector<double>& v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p / 100.0 * double(v.size() - 1);
  const size_t i = size_t(pos);
  const double f = pos - double(i);
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

void rbd_server_config_default(rbd_server_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->slots = 4;
  cfg->memory_budget_bytes = int64_t(64) << 20;
  cfg->max_batch_trajectories = 1024;
  cfg->max_batch_knots = 1 << 17;
  cfg->batch_window_us = 0.0;
  cfg->use_graph = 0;
  cfg->lanes_per_knot = rbd::kLanesPerKnotGroup;
}

rbd_status rbd_server_create(rbd_context* ctx, const rbd_server_config* cfg, rbd_server** out) {
  if (!out) return fail(RBD_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  if (rbd_status st = check_ctx(ctx)) return st;
  rbd_server_config c;
  if (cfg) c = *cfg; else rbd_server_config_default(&c);
  if (c.slots < 1 || c.slots > 64) return fail(RBD_ERR_INVALID_ARGUMENT, "slots must be in [1, 64]");
  if (c.memory_budget_bytes < int64_t(kAlign)) return fail(RBD_ERR_INVALID_ARGUMENT, "memory budget too small");
  if (c.max_batch_trajectories < 1 || c.max_batch_knots < 1)
    return fail(RBD_ERR_INVALID_ARGUMENT, "batch limits must be positive");
  if (c.lanes_per_knot != 1 && c.lanes_per_knot != rbd::kLanesPerKnotGroup)
    return fail(RBD_ERR_INVALID_ARGUMENT, "server lanes_per_knot must be 1 or 16");
  if (!(c.batch_window_us >= 0)) return fail(RBD_ERR_INVALID_ARGUMENT, "negative batch window");
  if (rbd_status st = use_device(ctx)) return st;
  auto* s = new (std::nothrow) rbd_server();
  if (!s) return fail(RBD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->cfg = c;
  rbd_config rc;
  rbd_config_default(&rc);
  rc.device = ctx->device;
  rc.flags = ctx->flags;
  rc.precision = ctx->precision;
  rc.lanes_per_knot = c.lanes_per_knot;
  rbd_status st = rbd_create(&rc, &s->run_ctx);
  const size_t budget = round_up(size_t(c.memory_budget_bytes));
  hipError_t e = hipSuccess;
  if (st == RBD_OK) e = hipMalloc(&s->d_pool, budget);
  if (st == RBD_OK && e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->h_pool), budget, hipHostMallocDefault);
  if (st == RBD_OK && e == hipSuccess) {
    s->alloc.reset(new RangeAllocator(budget));
    s->slots.resize(size_t(c.slots));
    for (Slot& sl : s->slots) {
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
      if (e == hipSuccess && st == RBD_OK)
        st = rbd_batch_create(s->run_ctx, int(c.max_batch_trajectories), &sl.batch);
    }
  }
  if (st != RBD_OK || e != hipSuccess) {
    const std::string msg = st != RBD_OK ? std::string(rbd_last_error()) : std::string(hipGetErrorString(e));
    rbd_server_destroy(s);
    return fail(st != RBD_OK ? st : RBD_ERR_OUT_OF_MEMORY, "server setup: " + msg);
  }
  s->dispatcher = std::thread(dispatch_loop, s);
  s->completer = std::thread(complete_loop, s);
  *out = s;
  return RBD_OK;
}

rbd_status rbd_server_destroy(rbd_server* s) {
  if (!s) return RBD_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // requests already on the device finish; queued ones are failed
    for (auto& r : s->queue) r->status = -1;
    s->queue.clear();
    s->stop = true;
  }
  s->cv_dispatch.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  s->cv_complete.notify_all();
  if (s->completer.joinable()) s->completer.join();
  s->cv_done.notify_all();
  (void)use_device(s->ctx);
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.batch) rbd_batch_destroy(sl.batch);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  if (s->d_pool) (void)hipFree(s->d_pool);
  if (s->h_pool) (void)hipHostFree(s->h_pool);
  if (s->run_ctx) rbd_destroy(s->run_ctx);
  delete s;
  return RBD_OK;
}

rbd_status rbd_server_submit(rbd_server* s, const float* h_in, float* h_out, int64_t n_knots, uint64_t* id) {
  if (!s) return fail(RBD_ERR_INVALID_ARGUMENT, "null server");
  if (n_knots <= 0) return fail(RBD_ERR_INVALID_ARGUMENT, "a request needs at least one knot");
  if (!h_in || !h_out || !id) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  auto r = std::make_shared<Request>();
  r->n = n_knots;
  r->h_in = h_in;
  r->h_out = h_out;
  r->in_bytes = size_t(n_knots) * rbd::kInPerKnot * sizeof(float);
  r->out_bytes = size_t(n_knots) * rbd::kOutPerKnot * sizeof(float);
  r->bytes = round_up(r->in_bytes) + round_up(r->out_bytes);
  if (r->bytes > s->alloc->size()) return fail(RBD_ERR_CAPACITY, "request larger than the server's memory budget");
  if (n_knots > s->cfg.max_batch_knots) return fail(RBD_ERR_CAPACITY, "request longer than max_batch_knots");
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(RBD_ERR_INVALID_ARGUMENT, "server is stopping");
    if (s->async_error != RBD_OK) return fail(s->async_error, "server failed earlier");
    r->id = s->next_id++;
    *id = r->id;
    s->live[r->id] = r;
    s->queue.push_back(r);
  }
  s->cv_dispatch.notify_all();
  return RBD_OK;
}

int rbd_server_poll(rbd_server* s, uint64_t id) {
  if (!s) return -2;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(id);
  if (it == s->live.end()) return -2;
  const int st = it->second->status;
  if (st != 0) s->live.erase(it);
  return st;
}

rbd_status rbd_server_wait(rbd_server* s, uint64_t id, double timeout_s) {
  if (!s) return fail(RBD_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(id);
  if (it == s->live.end()) return fail(RBD_ERR_INVALID_ARGUMENT, "unknown or already collected request id");
  std::shared_ptr<Request> r = it->second;
  const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s));
  if (!s->cv_done.wait_until(lk, deadline, [&r] { return r->status != 0; }))
    return fail(RBD_ERR_CAPACITY, "timed out waiting for the request");
  s->live.erase(id);
  return r->status == 1 ? RBD_OK : fail(s->async_error != RBD_OK ? s->async_error : RBD_ERR_HIP, "request failed");
}

rbd_status rbd_server_get_stats(rbd_server* s, rbd_server_stats* out) {
  if (!s || !out) return fail(RBD_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  std::vector<double> lat = s->latencies_us;
  out->requests_completed = s->requests_done;
  out->knots_completed = s->knots_done;
  out->batches = s->batches;
  out->max_inflight_batches = s->inflight_max;
  out->max_batch_knots = s->admitted_knots_max;
  out->max_budget_bytes_used = int64_t(s->budget_used_max);
  out->queued = int64_t(s->queue.size());
  out->p50_latency_us = percentile(lat, 50.0);
  out->p99_latency_us = percentile(lat, 99.0);
  out->max_latency_us = lat.empty() ? 0.0 : lat.back();
  out->mean_latency_us = lat.empty() ? 0.0 : [&lat] { double a = 0; for (double v : lat) a += v; return a / double(lat.size()); }();
  const double span = s->any_done ? std::chrono::duration<double>(s->t_last - s->t_first).count() : 0.0;
  out->knots_per_s = span > 0 ? double(s->knots_done) / span : 0.0;
  return RBD_OK;
}

rbd_status rbd_server_reset_stats(rbd_server* s) {
  if (!s) return fail(RBD_ERR_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->latencies_us.clear();
  s->knots_done = s->requests_done = s->batches = s->inflight_max = s->admitted_knots_max = 0;
  s->budget_used_max = 0;
  s->any_done = false;
  return RBD_OK;
}

// Open-loop load: requests arrive as a Poisson process of `rate_per_s` for
// `duration_s`; each is one trajectory whose length is drawn uniformly from
// lengths[0..n_lengths); its inputs come from `pool_in` (pool_knots x 21, used
// cyclically) and results go to a private buffer.  Waits for every request,
// then fills `out` with the server's statistics over this run.  The arrival
// times are scheduled in advance (the generator sleeps until each), so a busy
// server makes requests queue rather than slowing arrivals down.
rbd_status rbd_server_run_poisson(rbd_server* s, double rate_per_s, double duration_s, const int64_t* lengths,
                                  int n_lengths, const float* pool_in, int64_t pool_knots, uint64_t seed,
                                  rbd_server_stats* out) {
  if (!s || !lengths || n_lengths <= 0 || !pool_in || !out) return fail(RBD_ERR_INVALID_ARGUMENT, "null argument");
  if (!(rate_per_s > 0) || !(duration_s > 0)) return fail(RBD_ERR_INVALID_ARGUMENT, "rate and duration must be > 0");
  int64_t maxlen = 0;
  for (int i = 0; i < n_lengths; ++i) maxlen = std::max(maxlen, lengths[i]);
  if (maxlen <= 0 || maxlen > pool_knots) return fail(RBD_ERR_INVALID_ARGUMENT, "lengths must be in [1, pool_knots]");
  if (rbd_status st = rbd_server_reset_stats(s)) return st;
  std::mt19937_64 rng(seed);
  std::exponential_distribution<double> gap(rate_per_s);
  std::uniform_int_distribution<int> pick(0, n_lengths - 1);
  const int64_t expected = int64_t(rate_per_s * duration_s * 1.2) + 16;
  // result buffers are recycled once their request completed
  std::vector<std::unique_ptr<float[]>> outbufs;
  std::deque<std::pair<uint64_t, size_t>> pending;   // (id, outbuf index)
  std::vector<size_t> free_bufs;
  (void)expected;
  const auto t0 = Clock::now();
  double t = 0.0;
  int64_t cursor = 0;
  rbd_status st = RBD_OK;
  for (;;) {
    t += gap(rng);
    if (t > duration_s) break;
    // spin until the arrival time (sleep granularity is too coarse at 10 us gaps)
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t));
    while (Clock::now() < due) {
    }
    // collect finished requests to recycle buffers
    while (!pending.empty()) {
      const int p = rbd_server_poll(s, pending.front().first);
      if (p == 0) break;
      if (p < 0) { st = RBD_ERR_HIP; break; }
      free_bufs.push_back(pending.front().second);
      pending.pop_front();
    }
    if (st != RBD_OK) break;
    const int64_t n = lengths[pick(rng)];
    if (cursor + n > pool_knots) cursor = 0;
    size_t bi;
    if (!free_bufs.empty()) { bi = free_bufs.back(); free_bufs.pop_back(); }
    else { outbufs.emplace_back(new float[size_t(maxlen) * rbd::kOutPerKnot]); bi = outbufs.size() - 1; }
    uint64_t id = 0;
    st = rbd_server_submit(s, pool_in + cursor * rbd::kInPerKnot, outbufs[bi].get(), n, &id);
    if (st != RBD_OK) break;
    pending.emplace_back(id, bi);
    cursor += n;
  }
  for (auto& p : pending) {
    const rbd_status w = rbd_server_wait(s, p.first, 60.0);
    if (w != RBD_OK && st == RBD_OK) st = w;
  }
  if (st != RBD_OK) return st;
  return rbd_server_get_stats(s, out);
}

}  // extern "C"
