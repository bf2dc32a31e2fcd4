// TEST INFRASTRUCTURE ONLY -- the CPU baseline's thread pool.
//
// Persistent worker threads, each with its own task queue; add_task(tid, f)
// hands f to worker tid and sync() blocks until every task handed out so far
// has finished.  Same contract as the reference's ReusableThreads
// (utils/ReusableThreads.h:26-115, addTask :43-52, sync :55-64), written so
// that sync() never reads a queue without holding that queue's lock: the
// number of outstanding tasks is one counter guarded by one mutex, raised
// before a task becomes visible to its worker and lowered after it ran.
#pragma once
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace rbd_oracle {

class ReusablePool {
 public:
  explicit ReusablePool(int n_threads) : n_(n_threads), slots_(new Slot[n_threads]) {
    for (int t = 0; t < n_; ++t) workers_.emplace_back([this, t] { run(t); });
  }
  ~ReusablePool() {
    for (int t = 0; t < n_; ++t) {
      std::lock_guard<std::mutex> g(slots_[t].m);
      slots_[t].stop = true;
      slots_[t].cv.notify_all();
    }
    for (auto& w : workers_) w.join();
  }
  ReusablePool(const ReusablePool&) = delete;
  ReusablePool& operator=(const ReusablePool&) = delete;

  int size() const { return n_; }

  void add_task(int tid, std::function<void()> f) {
    {
      std::lock_guard<std::mutex> g(done_m_);
      ++pending_;
    }
    {
      std::lock_guard<std::mutex> g(slots_[tid].m);
      slots_[tid].q.push_back(std::move(f));
    }
    slots_[tid].cv.notify_one();
  }

  void sync() {
    std::unique_lock<std::mutex> g(done_m_);
    done_cv_.wait(g, [this] { return pending_ == 0; });
  }

 private:
  struct Slot {
    std::mutex m;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    bool stop = false;
  };
  void run(int t) {
    Slot& s = slots_[t];
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> g(s.m);
        s.cv.wait(g, [&s] { return s.stop || !s.q.empty(); });
        if (s.q.empty()) return;  // stop requested and nothing left
        job = std::move(s.q.front());
        s.q.pop_front();
      }
      job();
      {
        std::lock_guard<std::mutex> g(done_m_);
        if (--pending_ == 0) done_cv_.notify_all();
      }
    }
  }

  int n_;
  std::unique_ptr<Slot[]> slots_;
  std::vector<std::thread> workers_;
  std::mutex done_m_;
  std::condition_variable done_cv_;
  long pending_ = 0;
};

}  // namespace rbd_oracle
