// Persistent host worker pool for per-job host loops (transcripts, hint
// assembly, argument images).  Plain C++, no device code.
//
// Any number of callers may run a loop at once: each run() posts a job, every
// idle worker joins the posted job with the fewest workers, and the caller
// works on its own job too, so a caller never waits for another caller's loop
// and concurrent callers share the cores.  A loop body may itself call run().
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace p3g {

class HostPool {
 public:
  explicit HostPool(unsigned nthreads) {
    for (unsigned i = 1; i < nthreads; i++) threads_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

  unsigned threads() const { return (unsigned)threads_.size() + 1; }

  // Calls f(i) for every i in [0, count) and returns when all calls are done.
  // The first exception thrown by f is rethrown here after the others end.
  void run(size_t count, const std::function<void(size_t)>& f) {
    if (threads_.empty() || count < 2) {
      for (size_t i = 0; i < count; i++) f(i);
      return;
    }
    Job job;
    job.fn = &f;
    job.n = count;
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.push_back(&job);
    }
    cv_.notify_all();
    work(job);
    std::exception_ptr e;
    {
      std::unique_lock<std::mutex> lk(mu_);
      unlink_locked(&job);  // no worker joins after this
      cv_done_.wait(lk, [&] { return job.active == 0; });
      e = job.err;
    }
    if (e) std::rethrow_exception(e);
  }

 private:
  struct Job {
    const std::function<void(size_t)>* fn = nullptr;
    size_t n = 0;
    std::atomic<size_t> next{0};
    unsigned active = 0;  // workers inside work(); guarded by mu_
    std::exception_ptr err;
  };

  void work(Job& j) {
    for (;;) {
      const size_t i = j.next.fetch_add(1);
      if (i >= j.n) return;
      try {
        (*j.fn)(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu_);
        if (!j.err) j.err = std::current_exception();
      }
    }
  }

  void unlink_locked(Job* j) {
    auto it = std::find(jobs_.begin(), jobs_.end(), j);
    if (it != jobs_.end()) jobs_.erase(it);
  }

  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (stop_) return;
      Job* j = jobs_.front();
      for (Job* c : jobs_)
        if (c->active < j->active) j = c;
      j->active++;
      lk.unlock();
      work(*j);
      lk.lock();
      unlink_locked(j);  // exhausted: every index has been claimed
      if (--j->active == 0) cv_done_.notify_all();
    }
  }

  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::vector<Job*> jobs_;
  std::vector<std::thread> threads_;
  bool stop_ = false;
};

}  // namespace p3g
