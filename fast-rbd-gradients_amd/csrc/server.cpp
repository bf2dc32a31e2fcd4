// Serving loop (rbd_server_*): continuous batching of forward-dynamics-gradient
// requests onto the device.
//
// A request is one trajectory: host input n x 21, host output n x 105, both
// caller-owned and valid until the request completes.  The dispatcher thread
// admits queued requests in arrival order while the fixed memory budget
// allows: a batch holds one contiguous range [queue header and trajectory
// descriptors | inputs of all its requests | outputs of all its requests]
// (n x 504 bytes per request plus 32 per descriptor, each part rounded up to
// 256 bytes) of device memory, and the same range of pinned staging, until its
// results are delivered.  It runs the batch on a free in-flight slot (own
// stream, own rbd_batch work queue): descriptors and inputs -> pinned staging,
// one H2D copy, one persistent-kernel launch, one D2H copy, an event.  While
// batches are on the device the dispatcher keeps admitting newly arrived
// requests into the next free slot, so work from several batches overlaps on
// the GPU.  Each slot has a completion thread: it polls the batch's
// completion word (zero-copy batches: the kernel's last wave writes it to
// host-mapped memory) or the slot's event, copies results to the callers' buffers, returns the memory to the budget,
// frees the slot and records each request's latency (submit -> result in the
// caller's buffer).  Slots complete independently, so copying out one large
// batch does not hold up the others.
//
// Every batch of a server runs with the same knot-to-lane mapping (fixed at
// creation), so a request's result does not depend on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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
  size_t stage_bytes = 0;               // header + descriptors (rounded)
  size_t in_bytes = 0, out_bytes = 0;   // packed inputs, packed outputs
  bool busy = false;                    // admitted batch not yet delivered
  bool launched = false;                // its device work is enqueued
  bool poll_word = false;               // the batch publishes its completion word
  uint64_t seq = 0;                     // that word's value for the batch in flight
  rbd_batch_record rec{};               // trace of the batch in flight
  std::thread completer;                // delivers this slot's batches
};

constexpr size_t kTraceRecords = 65536;
constexpr int kWarmRounds = 512;       // warm-up batches per slot at creation
// zero_copy = 2: batches of more knots than this are copied.  Measured at 1M
// req/s on MI355X: 9.3e7-9.4e7 knots/s with p50 1.2 ms at this threshold;
// 5e7-9.6e7 (the runtime's copy stalls) at 4096, 7e7-8.6e7 with no copies.
constexpr int64_t kZeroCopyMaxKnots = 16384;
constexpr size_t kDoneStride = 16;     // 64-bit words between slots' completion words (one 128-byte line each)
// Zero-copy batches of up to this many chunks (waves) publish a completion
// word.  Every wave of such a launch writes back its L2 and waits before it is
// counted: about 2 us less from launch to delivery for the small batches of
// light load, but a collapse for large batches beside other launches (zero-copy
// at 300k req/s: p50 58 us -> 4 ms), which keep the event.
constexpr int64_t kWordMaxChunks = 64;

}  // namespace rbdi

struct rbd_server {
  rbd_context* ctx = nullptr;          // the caller's context (device, flags, precision)
  rbd_context* run_ctx = nullptr;      // same, with the knot-to-lane mapping fixed
  rbd_server_config cfg{};
  char* d_pool = nullptr;              // device memory of the budget (copied batches)
  char* h_pool = nullptr;              // pinned staging of the same size
  char* dh_pool = nullptr;             // device address of h_pool (zero-copy batches)
  // completion words, one line per slot (host-mapped, coherent; zero-copy
  // batches): the batch kernel's last wave writes the slot's sequence number
  uint64_t* h_done = nullptr;
  uint64_t* dh_done = nullptr;         // their device address
  // where the kernels find a batch's queue, inputs and outputs: the device
  // pool (copy mode) or the pinned pool itself (zero-copy mode)
  char* work_pool(bool copy) const { return copy ? d_pool : dh_pool; }
  // whether a batch of this many knots goes through device memory (copies) or
  // is read and written in the pinned pool (zero-copy)
  bool copied(int64_t knots) const {
    return cfg.zero_copy == 0 || (cfg.zero_copy == 2 && knots > zero_copy_max_knots);
  }
  int64_t zero_copy_max_knots = 0;
  std::unique_ptr<RangeAllocator> alloc;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_dispatch;  // new request, returned memory, freed slot, stop
  std::atomic<uint64_t> wake{0};        // bumped with every cv_dispatch notification (polling dispatcher)
  std::condition_variable cv_complete;  // a slot was launched, stop (one completion thread per slot)
  std::condition_variable cv_done;      // a request finished
  std::deque<std::shared_ptr<Request>> queue;                    // submitted, not admitted
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;    // not yet collected
  uint64_t next_id = 1;
  uint64_t mem_gen = 0;                 // bumped whenever memory or a slot is returned
  bool stop = false;
  rbd_status async_error = RBD_OK;

  // statistics
  Clock::time_point t_created = Clock::now();
  std::vector<rbd_batch_record> trace;  // ring of the last kTraceRecords batches
  uint64_t trace_count = 0;
  std::vector<double> latencies_us;
  int64_t knots_done = 0, requests_done = 0, batches = 0, inflight_max = 0, admitted_knots_max = 0;
  size_t budget_used_max = 0;
  Clock::time_point t_first{}, t_last{};
  bool any_done = false;

  std::thread dispatcher;
};

namespace {

double since(const rbd_server* s, Clock::time_point t) {
  return std::chrono::duration<double, std::micro>(t - s->t_created).count();
}

void dispatch_loop(rbd_server* s) {
  if (use_device(s->ctx) != RBD_OK) return;
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [s] {
    if (s->stop) return true;
    if (s->queue.empty()) return false;
    for (const Slot& sl : s->slots)
      if (!sl.busy) return true;
    return false;
  };
  for (;;) {
    if (s->cfg.poll_dispatch) {
      // poll instead of sleeping: a condition-variable wake-up costs about
      // 10 us, more than the batch's kernel at low load
      while (!ready()) {
        const uint64_t w = s->wake.load(std::memory_order_acquire);
        lk.unlock();
        while (s->wake.load(std::memory_order_acquire) == w) std::this_thread::yield();
        lk.lock();
      }
    } else {
      s->cv_dispatch.wait(lk, ready);
    }
    if (s->stop) return;
    if (s->async_error != RBD_OK) {
      // a batch failed: admit nothing more, fail what is queued (submit refuses new requests)
      for (auto& r : s->queue) r->status = -1;
      s->queue.clear();
      s->cv_done.notify_all();
      continue;
    }
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
      const int next = int(admitted.size()) + 1;
      if (next > s->cfg.max_batch_trajectories) break;
      if (knots + r->n > s->cfg.max_batch_knots) break;
      if (round_up(batch_stage_bytes(next)) + round_up(in_bytes + r->in_bytes) + round_up(out_bytes + r->out_bytes) >
          room)
        break;
      r->in_off = in_bytes;                                 // relative to the batch, fixed up below
      r->out_off = out_bytes;
      in_bytes += r->in_bytes;
      out_bytes += r->out_bytes;
      knots += r->n;
      admitted.push_back(r);
    }
    const size_t stage_bytes = round_up(batch_stage_bytes(int(admitted.size())));
    size_t off = SIZE_MAX;
    if (!admitted.empty()) off = s->alloc->alloc(stage_bytes + round_up(in_bytes) + round_up(out_bytes));
    if (off == SIZE_MAX) {
      // the head request does not fit yet: sleep until memory comes back
      const uint64_t gen = s->mem_gen;
      s->cv_dispatch.wait(lk, [s, gen] { return s->stop || s->mem_gen != gen; });
      continue;
    }
    s->queue.erase(s->queue.begin(), s->queue.begin() + std::ptrdiff_t(admitted.size()));
    for (auto& r : admitted) {
      r->in_off += off + stage_bytes;
      r->out_off += off + stage_bytes + round_up(in_bytes);
    }
    s->budget_used_max = std::max(s->budget_used_max, s->alloc->used());
    s->admitted_knots_max = std::max(s->admitted_knots_max, knots);
    Slot& sl = s->slots[size_t(slot)];
    sl.busy = true;
    sl.reqs = std::move(admitted);
    sl.off = off;
    sl.stage_bytes = stage_bytes;
    sl.bytes = stage_bytes + round_up(in_bytes) + round_up(out_bytes);
    sl.in_bytes = in_bytes;
    sl.out_bytes = out_bytes;
    sl.rec = rbd_batch_record{};
    sl.rec.oldest_submit_us = since(s, sl.reqs.front()->t_submit);
    sl.rec.admit_us = since(s, Clock::now());
    sl.rec.requests = int64_t(sl.reqs.size());
    sl.rec.knots = knots;
    sl.rec.slot = slot;
    lk.unlock();
    // stage descriptors and inputs; copy mode: one H2D copy, the launch, one
    // D2H copy; zero-copy mode: the launch alone; then the event
    const bool copy = s->copied(knots);
    char* const work = s->work_pool(copy);
    std::vector<rbd_traj> trajs;
    trajs.reserve(sl.reqs.size());
    for (auto& r : sl.reqs) {
      std::memcpy(s->h_pool + r->in_off, r->h_in, r->in_bytes);
      trajs.push_back(rbd_traj{reinterpret_cast<const float*>(work + r->in_off),
                               reinterpret_cast<float*>(work + r->out_off), r->n});
    }
    const Shape shape = shape_for_batch(s->run_ctx, knots);
    const int lanes = shape.lanes;
    const int n_traj = int(trajs.size());
    const int64_t chunks = batch_stage(s->h_pool + off, trajs.data(), n_traj, lanes);
    const size_t out0 = off + stage_bytes + round_up(in_bytes);
    // a zero-copy batch's results are final when its kernel ends: it publishes
    // a completion word, which the completer polls instead of the event
    sl.poll_word = !copy && s->h_done && knots > 0 && chunks <= kWordMaxChunks;
    unsigned long long* word = nullptr;
    if (sl.poll_word) {
      ++sl.seq;
      word = reinterpret_cast<unsigned long long*>(s->dh_done + size_t(slot) * kDoneStride);
    }
    hipError_t e = hipSuccess;
    if (copy)
      e = hipMemcpyAsync(s->d_pool + off, s->h_pool + off, stage_bytes + in_bytes, hipMemcpyHostToDevice, sl.stream);
    rbd_status st = RBD_OK;
    if (e == hipSuccess)
      st = batch_launch_staged(sl.batch, work + off, trajs.data(), n_traj, knots, chunks, shape, sl.stream, word,
                               sl.seq);
    if (copy && e == hipSuccess && st == RBD_OK)
      e = hipMemcpyAsync(s->h_pool + out0, s->d_pool + out0, out_bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess && st == RBD_OK) e = hipEventRecord(sl.done, sl.stream);
    sl.rec.enqueued_us = since(s, Clock::now());
    lk.lock();
    if (e != hipSuccess || st != RBD_OK) s->async_error = st != RBD_OK ? st : RBD_ERR_HIP;
    sl.launched = true;
    s->batches++;
    int64_t inflight = 0;
    for (const Slot& x : s->slots) inflight += x.busy ? 1 : 0;
    s->inflight_max = std::max(s->inflight_max, inflight);
    s->cv_complete.notify_all();
  }
}

inline void cpu_relax() {
#if defined(__x86_64__)
  __builtin_ia32_pause();
#else
  std::this_thread::yield();
#endif
}

// Waits for the slot's batch.  A zero-copy batch is polled on its completion
// word (the kernel writes it after every output is visible to the host), with
// the event as a backstop: a launch that failed never writes the word, and the
// event then reports the error.  Other batches are polled on the event, rather
// than blocked on: a blocking wait can sleep past the event by far more than a
// batch takes.
hipError_t wait_batch(rbd_server* s, int slot) {
  Slot& sl = s->slots[size_t(slot)];
  if (sl.poll_word) {
    const uint64_t* w = s->h_done + size_t(slot) * kDoneStride;
    for (unsigned spin = 1;; ++spin) {
      if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == sl.seq) return hipSuccess;
      if (spin % 256 == 0) {
        const hipError_t e = hipEventQuery(sl.done);
        if (e != hipErrorNotReady) return e;
      }
      cpu_relax();
    }
  }
  hipError_t e = hipEventQuery(sl.done);
  while (e == hipErrorNotReady) {
    std::this_thread::yield();
    e = hipEventQuery(sl.done);
  }
  return e;
}

// One per slot: waits for the slot's batch, delivers its results, returns its
// memory and the slot.  Slots complete independently, so a large batch's
// delivery does not hold up the others.
void complete_loop(rbd_server* s, int slot) {
  if (use_device(s->ctx) != RBD_OK) return;
  Slot& sl = s->slots[size_t(slot)];
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_complete.wait(lk, [s, &sl] { return s->stop || sl.launched; });
    if (!sl.launched) return;              // stop with nothing in flight
    const bool failed_launch = s->async_error != RBD_OK;
    lk.unlock();
    hipError_t e = failed_launch ? hipErrorUnknown : wait_batch(s, slot);
    // a failed batch's copies and kernel may still be queued or running on the
    // slot's stream: wait for them before its range of the pools is reused
    if (e != hipSuccess) (void)hipStreamSynchronize(sl.stream);
    // a two-wave kernel that gave up a handoff wait leaves invalid results
    const rbd_status dev = e == hipSuccess ? device_status(s->run_ctx) : RBD_OK;
    if (dev != RBD_OK) e = hipErrorUnknown;
    sl.rec.done_us = since(s, Clock::now());
    if (e == hipSuccess)
      for (auto& r : sl.reqs) std::memcpy(r->h_out, s->h_pool + r->out_off, r->out_bytes);
    const auto now = Clock::now();
    sl.rec.delivered_us = since(s, now);
    lk.lock();
    s->trace[size_t(s->trace_count % kTraceRecords)] = sl.rec;
    s->trace_count++;
    sl.launched = false;
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
    if (e != hipSuccess && s->async_error == RBD_OK) s->async_error = dev != RBD_OK ? dev : RBD_ERR_HIP;
    sl.reqs.clear();
    sl.busy = false;
    s->mem_gen++;
    s->cv_done.notify_all();
    s->wake.fetch_add(1, std::memory_order_release);
    s->cv_dispatch.notify_all();
  }
}

// Run every kernel a batch can use once on every slot (queues of 1 and of
// kSmallBatch + 1 one-knot trajectories, zero inputs), so first-use costs of
// the runtime fall at creation and not on the first requests.
rbd_status warm_up(rbd_server* s) {
  // touch the whole staging and device pools once, copies both ways included,
  // so no page of either is first used by a request
  const size_t pool = s->alloc->size();
  std::memset(s->h_pool, 0, pool);
  if (s->d_pool) {
    hipError_t e0 = hipMemcpy(s->d_pool, s->h_pool, pool, hipMemcpyHostToDevice);
    if (e0 == hipSuccess) e0 = hipMemcpy(s->h_pool, s->d_pool, pool, hipMemcpyDeviceToHost);
    if (e0 != hipSuccess) return hip_fail(e0, "server warm-up copies");
  }
  const int n_big = kSmallBatchTrajectories + 1;
  const size_t stage = round_up(batch_stage_bytes(n_big));
  const size_t in_sz = round_up(size_t(n_big) * rbd::kInPerKnot * sizeof(float));
  const size_t block = stage + in_sz + round_up(size_t(n_big) * rbd::kOutPerKnot * sizeof(float));
  if (block * s->slots.size() > pool) return RBD_OK;     // budget too small to bother
  const Shape shape = shape_for_batch(s->run_ctx, 1);
  const int lanes = shape.lanes;
  // the full per-batch sequence, many times per slot, each slot in its own
  // block of the pools: the runtime sizes its internal pools (signals, command
  // buffers) on first demand
  for (int round = 0; round < kWarmRounds; ++round)
    for (size_t si = 0; si < s->slots.size(); ++si) {
      Slot& sl = s->slots[si];
      const size_t off = si * block;
      const int n = (round & 1) ? n_big : 1;
      // both paths of a server that copies large batches only
      const bool copy = s->cfg.zero_copy == 2 ? (round & 2) != 0 : s->copied(n);
      char* const work = s->work_pool(copy);
      std::vector<rbd_traj> trajs(static_cast<size_t>(n_big));
      for (int t = 0; t < n_big; ++t)
        trajs[size_t(t)] =
            rbd_traj{reinterpret_cast<const float*>(work + off + stage) + size_t(t) * rbd::kInPerKnot,
                     reinterpret_cast<float*>(work + off + stage + in_sz) + size_t(t) * rbd::kOutPerKnot, 1};
      hipError_t e = hipEventSynchronize(sl.done);        // this slot's previous round is done with its block
      const int64_t chunks = batch_stage(s->h_pool + off, trajs.data(), n, lanes);
      if (copy && e == hipSuccess)
        e = hipMemcpyAsync(s->d_pool + off, s->h_pool + off, stage + in_sz, hipMemcpyHostToDevice, sl.stream);
      if (e != hipSuccess) return hip_fail(e, "server warm-up upload");
      if (rbd_status st = batch_launch_staged(sl.batch, work + off, trajs.data(), n, n, chunks, shape, sl.stream))
        return st;
      if (copy)
        e = hipMemcpyAsync(s->h_pool + off + stage + in_sz, s->d_pool + off + stage + in_sz,
                           size_t(n) * rbd::kOutPerKnot * sizeof(float), hipMemcpyDeviceToHost, sl.stream);
      if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
      if (e != hipSuccess) return hip_fail(e, "server warm-up");
    }
  for (Slot& sl : s->slots) {
    const hipError_t e = hipStreamSynchronize(sl.stream);
    if (e != hipSuccess) return hip_fail(e, "server warm-up");
  }
  return RBD_OK;
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
  cfg->max_batch_knots = 1 << 15;
  cfg->batch_window_us = 0.0;
  cfg->use_graph = 0;
  cfg->zero_copy = 2;                         // zero-copy for batches up to 16384 knots, copies above
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
  if (c.lanes_per_knot != 1 && c.lanes_per_knot != rbd::kLanesPerKnotGroup && c.lanes_per_knot != kTwoWaveLanes &&
      !is_group_lanes(c.lanes_per_knot))
    return fail(RBD_ERR_INVALID_ARGUMENT, "server lanes_per_knot must be 1, 4, 7, 8, 16 or 32");
  if (!(c.batch_window_us >= 0)) return fail(RBD_ERR_INVALID_ARGUMENT, "negative batch window");
  if (c.zero_copy < 0 || c.zero_copy > 2) return fail(RBD_ERR_INVALID_ARGUMENT, "zero_copy must be 0, 1 or 2");
  if (c.zero_copy_max_knots < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative zero_copy_max_knots");
  if (rbd_status st = use_device(ctx)) return st;
  auto* s = new (std::nothrow) rbd_server();
  if (!s) return fail(RBD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->cfg = c;
  s->trace.resize(kTraceRecords);
  rbd_config rc;
  rbd_config_default(&rc);
  rc.device = ctx->device;
  rc.flags = ctx->flags;
  rc.precision = ctx->precision;
  rc.lanes_per_knot = c.lanes_per_knot;
  rc.waves_per_group = ctx->waves_per_group;
  rbd_status st = rbd_create(&rc, &s->run_ctx);
  if (st == RBD_OK) s->run_ctx->opt.handoff_spins = ctx->opt.handoff_spins;   // the caller's wait bound
  const size_t budget = round_up(size_t(c.memory_budget_bytes));
  hipError_t e = hipSuccess;
  s->zero_copy_max_knots = c.zero_copy_max_knots > 0 ? c.zero_copy_max_knots : kZeroCopyMaxKnots;
  if (st == RBD_OK && c.zero_copy != 1) e = hipMalloc(&s->d_pool, budget);
  // zero-copy batches: the kernels read and write the pinned pool directly, so
  // it is mapped and fine-grained (coherent with the host without cache flushes)
  if (st == RBD_OK && e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->h_pool), budget,
                      c.zero_copy ? (hipHostMallocMapped | hipHostMallocCoherent) : hipHostMallocDefault);
  if (st == RBD_OK && e == hipSuccess && c.zero_copy)
    e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dh_pool), s->h_pool, 0);
  const size_t done_bytes = sizeof(uint64_t) * kDoneStride * size_t(c.slots);
  if (st == RBD_OK && e == hipSuccess && c.zero_copy)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->h_done), done_bytes, hipHostMallocMapped | hipHostMallocCoherent);
  if (st == RBD_OK && e == hipSuccess && c.zero_copy) {
    std::memset(s->h_done, 0, done_bytes);
    e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dh_done), s->h_done, 0);
  }
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
  if (st == RBD_OK && e == hipSuccess) st = warm_up(s);
  if (st != RBD_OK || e != hipSuccess) {
    const std::string msg = st != RBD_OK ? std::string(rbd_last_error()) : std::string(hipGetErrorString(e));
    rbd_server_destroy(s);
    return fail(st != RBD_OK ? st : RBD_ERR_OUT_OF_MEMORY, "server setup: " + msg);
  }
  s->dispatcher = std::thread(dispatch_loop, s);
  for (int i = 0; i < int(s->slots.size()); ++i) s->slots[size_t(i)].completer = std::thread(complete_loop, s, i);
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
  s->wake.fetch_add(1, std::memory_order_release);
  s->cv_dispatch.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  s->cv_complete.notify_all();
  for (Slot& sl : s->slots)
    if (sl.completer.joinable()) sl.completer.join();
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
  if (s->h_done) (void)hipHostFree(s->h_done);
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
  // alone in a batch it also needs the queue header and one descriptor
  if (r->bytes + round_up(batch_stage_bytes(1)) > s->alloc->size())
    return fail(RBD_ERR_CAPACITY, "request larger than the server's memory budget");
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
  s->wake.fetch_add(1, std::memory_order_release);
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

rbd_status rbd_server_get_trace(rbd_server* s, rbd_batch_record* out, int64_t max_records, int64_t* n_out) {
  if (!s || !out || !n_out || max_records < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t have = std::min<uint64_t>(s->trace_count, kTraceRecords);
  const uint64_t n = std::min<uint64_t>(have, uint64_t(max_records));
  const uint64_t first = s->trace_count - n;
  for (uint64_t i = 0; i < n; ++i) out[i] = s->trace[size_t((first + i) % kTraceRecords)];
  *n_out = int64_t(n);
  return RBD_OK;
}

rbd_status rbd_server_reset_stats(rbd_server* s) {
  if (!s) return fail(RBD_ERR_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->latencies_us.clear();
  s->knots_done = s->requests_done = s->batches = s->inflight_max = s->admitted_knots_max = 0;
  s->budget_used_max = 0;
  s->any_done = false;
  s->trace_count = 0;
  return RBD_OK;
}

// Open-loop load: requests arrive as a Poisson process of `rate_per_s` for
// `duration_s`; each is one trajectory whose length is drawn uniformly from
// lengths[0..n_lengths); its inputs come from `pool_in` (pool_knots x 21, used
// cyclically) and results go to result buffers the generator recycles, as a
// client would: 2048 are allocated and touched before the first arrival, more
// only if that many are outstanding at once.  Waits for every request, then
// fills `out` with the server's statistics over this run.  The arrival times
// are scheduled in advance (the generator spins until each), so a busy server
// makes requests queue rather than slowing arrivals down.
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
  // result buffers are recycled once their request completed.  They are
  // touched here, before the clock starts (and when more are needed, here in
  // the load generator): an untouched buffer would take its page faults in
  // the server's delivery copy, whose slowdown would then deepen the backlog
  // and with it the number of fresh buffers.  About 20 ms of requests at the
  // offered rate are prepared (at least 2,048).
  const size_t buf_floats = size_t(maxlen) * rbd::kOutPerKnot;
  std::vector<std::unique_ptr<float[]>> outbufs;
  std::deque<std::pair<uint64_t, size_t>> pending;   // (id, outbuf index)
  std::vector<size_t> free_bufs;
  (void)expected;
  const size_t prepared = std::max<size_t>(2048, size_t(rate_per_s * 0.02));
  for (size_t i = 0; i < prepared; ++i) {
    outbufs.emplace_back(new float[buf_floats]);
    std::memset(outbufs.back().get(), 0, buf_floats * sizeof(float));
    free_bufs.push_back(prepared - 1 - i);
  }
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
    else {
      outbufs.emplace_back(new float[buf_floats]);
      std::memset(outbufs.back().get(), 0, buf_floats * sizeof(float));
      bi = outbufs.size() - 1;
    }
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
